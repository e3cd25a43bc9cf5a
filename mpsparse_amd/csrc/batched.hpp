// Batched Jacobi-PCG: many independent SPD systems, up to 4 right-hand sides
// each, solved together.
//
// The systems are the diagonal blocks of one CSR matrix (system s owns rows
// [sys_off[s], sys_off[s+1]) and all its columns lie in that range). Rows are
// cut into tiles of at most T rows that never cross a system boundary; every
// kernel walks a list of the tiles of still-active systems with a persistent
// grid, one workgroup per tile at a time. Vectors are interleaved by
// right-hand side ([row][4], 16 bytes per row), so the matrix is read once per
// iteration for all four recurrences and every gather is one 16-byte load.
//
// Per iteration (k = iteration parity; P[k] holds the previous direction):
//   k_bt_spmm:    p = z + beta p_old (z = D^-1 r, formed for gathered columns
//                 too), q = A p, block sums of <p, q>          -> P[k^1], Q
//   k_bt_fin_alpha (one wave per system): alpha = <r,z> / <p,q>
//   k_bt_update:  r -= alpha q, sums of <r,r>, <r,z>; on odd iterations also
//                 x += alpha' p' + alpha p with the even iteration's alpha'
//                 and direction p' (still in P[k^1]'s partner buffer)
//   k_bt_fin_beta: convergence per (system, rhs), beta, system activity
// The iterate is updated every other iteration: the recurrence itself never
// reads x, and the two updates are applied in their original order, so the
// result is the same bit for bit, while x and xl move through memory once per
// two iterations (the update's traffic falls from 132 to 100 B per row and
// iteration). Graphs hold an even number of iterations, so x is current
// whenever the host looks; a system whose last pair stops on an even
// iteration stays active for the odd one that applies its last update.
// Each (system, rhs) pair is its own CG recurrence with its own scalars; a
// pair that stops gets alpha = 0, a system with no running pair is skipped.
// The iterate is compensated (x + xl, the rounding errors of every update kept
// in xl), so it can get below the FP32 rounding floor of the solution; when
// every pair has stopped, a true-residual pass checks the iterate that will be
// returned (exactly x + xl for an FP64 result, fl(x + xl) for an FP32 one) and
// restarts the recurrence of every pair that stopped short.
// Only k_bt_fin_beta and the verification finisher change a system's
// activity, so a producer and the finisher reading its sums always agree on
// which tiles were processed.
#pragma once

#include "common.hpp"

namespace msp {
namespace batch {

constexpr int kRhs = 4;
constexpr int kEB = 6;  // row entries per load batch of the batched SpMM
constexpr int kUB = 4;  // rows per thread per load batch of the batched update

enum : int32_t { kPairRunning = 0, kPairConverged = 1, kPairBreakdown = 2, kPairMaxIter = 3, kPairStagnated = 4 };

struct PairState {
  float rz, pq, alpha, beta, rr, bb;
  float thresh;   // recurrence target ||r||^2 (lowered after a failed check)
  float target;   // tol^2 ||b||^2, the true-residual goal
  float rr_true;  // ||b - A x||^2 of the last verification
  int32_t status, iter, nverify;
  int32_t collapse;  // fold xl into x at the next update (restart of an FP32 result)
  float alpha_hold;  // alpha of the last even iteration, applied to x on the odd one
};

constexpr int kBD = 8;  // diagonals per system at most in the batched diagonal storage

struct View {
  const uint32_t* __restrict__ row_ptr;
  const uint32_t* __restrict__ col;
  const float* __restrict__ val;
  const float* __restrict__ dinv;
  // batched diagonal storage (nullptr when a system has more than kBD
  // diagonals): plane d, row i = A(i, i + offs[s][d]) for row i of system s
  // (0 where absent), offsets ascending per system, unused slots 0
  const float* __restrict__ planes;  // [kBD][ld]
  uint64_t ld;
  const int32_t* __restrict__ offs;  // [nsys][kBD]
  const uint32_t* __restrict__ sys_off;
  const uint32_t* __restrict__ tile_r0;   // [ntiles + 1]
  const uint32_t* __restrict__ tile_sys;  // [ntiles]
  const uint32_t* __restrict__ active;    // active tile list
  const uint32_t* __restrict__ n_active;  // its length
  const int32_t* __restrict__ sys_active;
};

__device__ __forceinline__ f4 ld4(const float* p, uint32_t i) { return reinterpret_cast<const f4*>(p)[i]; }
__device__ __forceinline__ void st4(float* p, uint32_t i, f4 v) { reinterpret_cast<f4*>(p)[i] = v; }

// Per-tile sums of ND values through LDS; the tile's partials are written by
// thread 0 to partials[tile][ND].
template <int ND>
__device__ __forceinline__ void tile_partials(CSum (&acc)[ND], CSum* partials, uint32_t tile) {
  block_reduce<ND>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) partials[(size_t)tile * ND + d] = acc[d];
  }
}

// Init: x = xl = 0, r = b, P[0] = 0; sums of <b,b>, <r,r>, <r,z> per rhs.
__global__ __launch_bounds__(kBlock) void k_bt_init(View V, const float* __restrict__ b, float* __restrict__ x,
                                                    float* __restrict__ xl, float* __restrict__ r,
                                                    float* __restrict__ p0, uint32_t ntiles,
                                                    CSum* __restrict__ partials) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t r0 = V.tile_r0[t], r1 = V.tile_r0[t + 1];
    CSum acc[3 * kRhs];
#pragma unroll
    for (int d = 0; d < 3 * kRhs; ++d) acc[d] = csum_zero();
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
      const f4 bv = ld4(b, i);
      const float di = V.dinv[i];
      st4(x, i, f4{0.f, 0.f, 0.f, 0.f});
      st4(xl, i, f4{0.f, 0.f, 0.f, 0.f});
      st4(r, i, bv);
      st4(p0, i, f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int k = 0; k < kRhs; ++k) {
        acc[k] = csum_add(acc[k], bv[k] * bv[k]);
        acc[kRhs + k] = csum_add(acc[kRhs + k], bv[k] * bv[k]);
        acc[2 * kRhs + k] = csum_add(acc[2 * kRhs + k], bv[k] * (di * bv[k]));
      }
    }
    tile_partials<3 * kRhs>(acc, partials, t);
    __syncthreads();
  }
}

// Sum of the ND partials of system s's tiles [t0, t1) by one wave; valid in lane 0.
template <int ND>
__device__ __forceinline__ void sys_sum(const CSum* __restrict__ partials, uint32_t t0, uint32_t t1, float (&tot)[ND]) {
  const int lane = threadIdx.x & (kWave - 1);
  CSum a[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) a[d] = csum_zero();
  for (uint32_t t = t0 + lane; t < t1; t += kWave) {
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] = csum_merge(a[d], partials[(size_t)t * ND + d]);
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) tot[d] = csum_val(wave_reduce(a[d]));
}

__device__ __forceinline__ void refresh_system(PairState* ps, int32_t* sys_active, uint32_t s) {
  int32_t any = 0;
#pragma unroll
  for (int k = 0; k < kRhs; ++k) any |= (ps[s * kRhs + k].status == kPairRunning) ? 1 : 0;
  sys_active[s] = any;
}

// Init finisher (one wave per system): thresholds and first scalars. Pairs
// beyond nrhs, and right-hand sides that are already below tol, stop at once.
__global__ __launch_bounds__(kBlock) void k_bt_fin_init(const CSum* __restrict__ partials,
                                                        const uint32_t* __restrict__ sys_tile, uint32_t nsys,
                                                        int32_t nrhs, float tol, int32_t maxiter,
                                                        PairState* __restrict__ ps, int32_t* __restrict__ sys_active) {
  const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (s >= nsys) return;
  float tot[3 * kRhs];
  sys_sum<3 * kRhs>(partials, sys_tile[s], sys_tile[s + 1], tot);
  if ((threadIdx.x & (kWave - 1)) != 0) return;
#pragma unroll
  for (int k = 0; k < kRhs; ++k) {
    PairState& p = ps[s * kRhs + k];
    p.bb = tot[k];
    p.rr = tot[kRhs + k];
    p.rz = tot[2 * kRhs + k];
    p.thresh = tol * tol * tot[k];
    p.target = p.thresh;
    p.alpha = 0.0f;
    p.beta = 0.0f;
    p.pq = 0.0f;
    p.rr_true = tot[kRhs + k];
    p.iter = 0;
    p.nverify = 0;
    p.collapse = 0;
    int32_t st = kPairRunning;
    if (k >= nrhs || !(p.rr > p.thresh)) st = kPairConverged;
    if (!isfinite(tot[k]) || !isfinite(p.rz)) st = kPairBreakdown;
    if (st == kPairRunning && maxiter <= 0) st = kPairMaxIter;
    p.status = st;
  }
  refresh_system(ps, sys_active, s);
}

// Active tile list: tiles of systems with a running pair, in tile order.
// One workgroup (exclusive scan over the systems' tile counts, then fill).
__global__ __launch_bounds__(kBlock) void k_bt_active_list(const uint32_t* __restrict__ sys_tile, uint32_t nsys,
                                                           const int32_t* __restrict__ sys_active,
                                                           uint32_t* __restrict__ list, uint32_t* __restrict__ n_active) {
  __shared__ uint32_t wsum[kWavesPerBlock];
  __shared__ uint32_t base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (uint32_t c = 0; c < nsys; c += kBlock) {
    const uint32_t s = c + threadIdx.x;
    const uint32_t cnt = (s < nsys && sys_active[s]) ? sys_tile[s + 1] - sys_tile[s] : 0u;
    // inclusive scan in the wave
    uint32_t v = cnt;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o = __shfl_up(v, off, kWave);
      if (lane >= off) v += o;
    }
    if (lane == kWave - 1) wsum[wid] = v;
    __syncthreads();
    uint32_t wbase = base;
    for (int w = 0; w < wid; ++w) wbase += wsum[w];
    const uint32_t start = wbase + v - cnt;
    for (uint32_t k = 0; k < cnt; ++k) list[start + k] = sys_tile[s] + k;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) base = start + cnt;
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_active = base;
}

// p = z + beta p_old for this row and every gathered column; q = A p.
// RPT rows per thread in a strided pattern; the 4 right-hand sides as one f4.
template <int RPT>
__global__ __launch_bounds__(kBlock) void k_bt_spmm(View V, const float* __restrict__ r,
                                                    const float* __restrict__ pold, float* __restrict__ pnew,
                                                    float* __restrict__ q, const PairState* __restrict__ ps,
                                                    CSum* __restrict__ partials) {
  const uint32_t na = *V.n_active;
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    const uint32_t t = V.active[a];
    const uint32_t s = V.tile_sys[t];
    if (!V.sys_active[s]) continue;
    const uint32_t r0 = V.tile_r0[t], r1 = V.tile_r0[t + 1];
    f4 beta;
#pragma unroll
    for (int k = 0; k < kRhs; ++k) beta[k] = ps[s * kRhs + k].beta;
    CSum acc[kRhs];
#pragma unroll
    for (int k = 0; k < kRhs; ++k) acc[k] = csum_zero();
    for (uint32_t i0 = r0 + threadIdx.x * RPT; i0 < r1; i0 += kBlock * RPT) {
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const uint32_t i = i0 + u;
        if (i < r1) {
          const f4 pi = V.dinv[i] * ld4(r, i) + beta * ld4(pold, i);
          f4 qi = f4{0.f, 0.f, 0.f, 0.f};
          // kEB entries at a time: every column / value load, then every
          // gather, then the products in entry order (one entry at a time
          // waited two memory round trips per entry)
          const uint32_t e0 = V.row_ptr[i], e1 = V.row_ptr[i + 1];
          for (uint32_t eb = e0; eb < e1; eb += kEB) {
            uint32_t cj[kEB];
            float vj[kEB], dj[kEB];
            f4 rj[kEB], pj[kEB];
#pragma unroll
            for (int u = 0; u < kEB; ++u) {
              const bool in = eb + u < e1;
              cj[u] = in ? V.col[eb + u] : i;
              vj[u] = in ? V.val[eb + u] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kEB; ++u) {
              dj[u] = V.dinv[cj[u]];
              rj[u] = ld4(r, cj[u]);
              pj[u] = ld4(pold, cj[u]);
            }
#pragma unroll
            for (int u = 0; u < kEB; ++u)
              if (eb + u < e1) qi += vj[u] * (dj[u] * rj[u] + beta * pj[u]);
          }
          st4(pnew, i, pi);
          st4(q, i, qi);
#pragma unroll
          for (int k = 0; k < kRhs; ++k) acc[k] = csum_add(acc[k], pi[k] * qi[k]);
        }
      }
    }
    tile_partials<kRhs>(acc, partials, t);
    __syncthreads();
  }
}

// The same product over the batched diagonal storage: a row's plane values and
// its gathers at i + off (clamped into the system; absent entries are stored
// as 0 and contribute nothing) are all issued before the arithmetic, one
// memory round trip per row instead of three (row pointers -> columns ->
// gathers). NDIAG > 0: exactly that many diagonals (5: 2D stencils); 0: kBD.
// MID: every system's middle diagonal is the main one (offset 0), whose
// operands are the row's own (no second load). MIRROR (with MID): every
// system is exactly symmetric, so a lower diagonal -k reads the upper plane
// +k at row i - k (0 above the system's first row): 3 planes stream instead of
// 5, and the two upper planes' second reads come from cache (they stay plain
// loads); planes read once are nontemporal.
template <int RPT, int NDIAG, bool MID = false, bool MIRROR = false>
__global__ __launch_bounds__(kBlock) void k_bt_spmm_dia(View V, const float* __restrict__ r,
                                                        const float* __restrict__ pold, float* __restrict__ pnew,
                                                        float* __restrict__ q, const PairState* __restrict__ ps,
                                                        CSum* __restrict__ partials) {
  constexpr int ND = NDIAG > 0 ? NDIAG : kBD;
  const uint32_t na = *V.n_active;
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    const uint32_t t = V.active[a];
    const uint32_t s = V.tile_sys[t];
    if (!V.sys_active[s]) continue;
    const uint32_t r0 = V.tile_r0[t], r1 = V.tile_r0[t + 1];
    const int64_t lo = V.sys_off[s], hi = V.sys_off[s + 1];
    f4 beta;
#pragma unroll
    for (int k = 0; k < kRhs; ++k) beta[k] = ps[s * kRhs + k].beta;
    int32_t off[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) off[d] = V.offs[(size_t)s * kBD + d];
    CSum acc[kRhs];
#pragma unroll
    for (int k = 0; k < kRhs; ++k) acc[k] = csum_zero();
    for (uint32_t i0 = r0 + threadIdx.x; i0 < r1; i0 += kBlock * RPT) {
      float vv[RPT][ND], dd[RPT][ND], di[RPT];
      f4 rr[RPT][ND], pp[RPT][ND], ri[RPT], pi0[RPT];
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const uint32_t iu = i0 + u * kBlock;
        const uint32_t i = iu < r1 ? iu : i0;
        di[u] = V.dinv[i];
        ri[u] = ld4(r, i);
        pi0[u] = ld4(pold, i);
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          if constexpr (MIRROR) {
            if (d < ND / 2) {
              const int64_t m = (int64_t)i + off[d];
              const float v = V.planes[(uint64_t)(ND - 1 - d) * V.ld + (uint64_t)(m < lo ? lo : m)];
              vv[u][d] = m < lo ? 0.0f : v;
            } else if (d == ND / 2) {
              vv[u][d] = __builtin_nontemporal_load(V.planes + (uint64_t)d * V.ld + i);
            } else {
              vv[u][d] = V.planes[(uint64_t)d * V.ld + i];
            }
          } else {
            vv[u][d] = __builtin_nontemporal_load(V.planes + (uint64_t)d * V.ld + i);
          }
          if (MID && d == ND / 2) continue;  // the row's own operands
          int64_t j = (int64_t)i + off[d];
          j = j < lo ? lo : (j >= hi ? hi - 1 : j);
          dd[u][d] = V.dinv[j];
          rr[u][d] = ld4(r, (uint32_t)j);
          pp[u][d] = ld4(pold, (uint32_t)j);
        }
        if constexpr (MID) {
          dd[u][ND / 2] = di[u];
          rr[u][ND / 2] = ri[u];
          pp[u][ND / 2] = pi0[u];
        }
      }
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const uint32_t i = i0 + u * kBlock;
        if (i >= r1) continue;
        const f4 pi = di[u] * ri[u] + beta * pi0[u];
        f4 qi = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < ND; ++d) qi += vv[u][d] * (dd[u][d] * rr[u][d] + beta * pp[u][d]);
        st4(pnew, i, pi);
        st4(q, i, qi);
#pragma unroll
        for (int k = 0; k < kRhs; ++k) acc[k] = csum_add(acc[k], pi[k] * qi[k]);
      }
    }
    tile_partials<kRhs>(acc, partials, t);
    __syncthreads();
  }
}

// Distinct diagonal offsets (column - row) of every system, one workgroup per
// system: an LDS set of kBD slots filled by compare-and-swap, then sorted.
// A system with more than kBD offsets raises `fail`. nds[s] = its count.
__global__ __launch_bounds__(kBlock) void k_bd_offsets(const uint32_t* __restrict__ row_ptr,
                                                       const uint32_t* __restrict__ col,
                                                       const uint32_t* __restrict__ sys_off, uint32_t nsys,
                                                       int32_t* __restrict__ offs, int32_t* __restrict__ nds,
                                                       uint32_t* __restrict__ fail) {
  constexpr int32_t kEmpty = INT32_MIN;
  __shared__ int32_t set[kBD];
  __shared__ int32_t over;
  for (uint32_t s = blockIdx.x; s < nsys; s += gridDim.x) {
    if (threadIdx.x < kBD) set[threadIdx.x] = kEmpty;
    if (threadIdx.x == 0) over = 0;
    __syncthreads();
    const uint32_t lo = sys_off[s], hi = sys_off[s + 1];
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
      for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const int64_t o64 = (int64_t)col[e] - (int64_t)i;
        const int32_t o = (o64 < INT32_MIN + 1 || o64 > INT32_MAX) ? kEmpty : (int32_t)o64;
        bool placed = o != kEmpty;
        if (placed) {
          placed = false;
          for (int k = 0; k < kBD && !placed; ++k) {
            const int32_t cur = __atomic_load_n(&set[k], __ATOMIC_RELAXED);
            if (cur == o) {
              placed = true;
            } else if (cur == kEmpty) {
              const int32_t prev = atomicCAS(&set[k], kEmpty, o);
              placed = prev == kEmpty || prev == o;
            }
          }
        }
        if (!placed) over = 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t v[kBD];
      int c = 0;
      for (int k = 0; k < kBD; ++k)
        if (set[k] != kEmpty) v[c++] = set[k];
      for (int a = 1; a < c; ++a)  // insertion sort: ascending offsets
        for (int b = a; b > 0 && v[b - 1] > v[b]; --b) {
          const int32_t tmp = v[b];
          v[b] = v[b - 1];
          v[b - 1] = tmp;
        }
      for (int k = 0; k < kBD; ++k) offs[(size_t)s * kBD + k] = k < c ? v[k] : 0;
      nds[s] = c;
      if (over) atomicAdd(fail, 1u);
    }
    __syncthreads();
  }
}

// Symmetry of the 5-plane batched storage (offsets -b, -a, 0, a, b per
// system): rows whose lower value differs bitwise from the partner's upper
// value k rows back (or is not 0 where that row lies before the system).
__global__ __launch_bounds__(kBlock) void k_bd_symcheck(const float* __restrict__ planes, uint64_t ld,
                                                        const uint32_t* __restrict__ sys_off, uint32_t nsys,
                                                        const int32_t* __restrict__ offs,
                                                        uint32_t* __restrict__ misses) {
  uint32_t miss = 0;
  for (uint32_t s = blockIdx.x; s < nsys; s += gridDim.x) {
    const uint32_t lo = sys_off[s], hi = sys_off[s + 1];
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int64_t m = (int64_t)i + offs[(size_t)s * kBD + d];
        const float a = planes[(uint64_t)d * ld + i];
        const float b = m < (int64_t)lo ? 0.0f : planes[(uint64_t)(4 - d) * ld + (uint64_t)m];
        miss += __float_as_uint(a) != __float_as_uint(b) ? 1u : 0u;
      }
    }
  }
#pragma unroll
  for (int sh = kWave / 2; sh > 0; sh >>= 1) miss += __shfl_xor(miss, sh, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && miss) atomicAdd(misses, miss);
}

// Values into the batched diagonal storage (planes zero-filled by the
// caller), one workgroup per system; duplicates added in CSR order.
__global__ __launch_bounds__(kBlock) void k_bd_fill(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                    const uint32_t* __restrict__ sys_off, uint32_t nsys,
                                                    const int32_t* __restrict__ offs, uint64_t ld,
                                                    float* __restrict__ planes) {
  for (uint32_t s = blockIdx.x; s < nsys; s += gridDim.x) {
    int32_t off[kBD];
#pragma unroll
    for (int d = 0; d < kBD; ++d) off[d] = offs[(size_t)s * kBD + d];
    for (uint32_t i = sys_off[s] + threadIdx.x; i < sys_off[s + 1]; i += kBlock) {
      for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const int64_t o = (int64_t)col[e] - (int64_t)i;
#pragma unroll
        for (int d = 0; d < kBD; ++d)
          if (o == off[d]) {
            planes[(uint64_t)d * ld + i] += val[e];
            break;
          }
      }
    }
  }
}

// alpha = <r,z> / <p,q> per running pair (breakdown on a non-positive <p,q>).
__global__ __launch_bounds__(kBlock) void k_bt_fin_alpha(const CSum* __restrict__ partials,
                                                         const uint32_t* __restrict__ sys_tile, uint32_t nsys,
                                                         PairState* __restrict__ ps,
                                                         const int32_t* __restrict__ sys_active, int32_t even) {
  const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (s >= nsys || !sys_active[s]) return;
  float tot[kRhs];
  sys_sum<kRhs>(partials, sys_tile[s], sys_tile[s + 1], tot);
  if ((threadIdx.x & (kWave - 1)) != 0) return;
#pragma unroll
  for (int k = 0; k < kRhs; ++k) {
    PairState& p = ps[s * kRhs + k];
    p.alpha = 0.0f;
    if (p.status == kPairRunning) {
      p.pq = tot[k];
      if (!(tot[k] > 0.0f) || !isfinite(tot[k])) p.status = kPairBreakdown;
      else p.alpha = p.rz / tot[k];
    }
    if (even) p.alpha_hold = p.alpha;
  }
}

// r -= alpha q; sums of <r,r>, <r, D^-1 r>. X (odd iterations): also
// (x, xl) += alpha' p' then += alpha p (compensated, in that order), with
// alpha' = alpha_hold and p' the even iteration's direction.
template <bool X>
__global__ __launch_bounds__(kBlock) void k_bt_update(View V, float* __restrict__ x, float* __restrict__ xl,
                                                      float* __restrict__ r, const float* __restrict__ pprev,
                                                      const float* __restrict__ p, const float* __restrict__ q,
                                                      const PairState* __restrict__ ps,
                                                      CSum* __restrict__ partials) {
  const uint32_t na = *V.n_active;
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    const uint32_t t = V.active[a];
    const uint32_t s = V.tile_sys[t];
    if (!V.sys_active[s]) continue;
    const uint32_t r0 = V.tile_r0[t], r1 = V.tile_r0[t + 1];
    f4 alpha, alpha0;
    bool fold[kRhs];
#pragma unroll
    for (int k = 0; k < kRhs; ++k) {
      alpha[k] = ps[s * kRhs + k].alpha;
      alpha0[k] = ps[s * kRhs + k].alpha_hold;
      fold[k] = ps[s * kRhs + k].collapse != 0;
    }
    CSum acc[2 * kRhs];
#pragma unroll
    for (int d = 0; d < 2 * kRhs; ++d) acc[d] = csum_zero();
    // kUB rows per thread at a time, every load of the group issued before
    // the first update (a row at a time waited a memory round trip per row);
    // rows past the tile load row i0's operands and store nothing
    for (uint32_t i0 = r0 + threadIdx.x; i0 < r1; i0 += kUB * kBlock) {
      [[maybe_unused]] f4 xv[kUB], lv[kUB], p0v[kUB], pv[kUB];
      f4 rv[kUB], qv[kUB];
      float dv[kUB];
#pragma unroll
      for (int u = 0; u < kUB; ++u) {
        const uint32_t iu = i0 + u * kBlock;
        const uint32_t i = iu < r1 ? iu : i0;
        if constexpr (X) {
          xv[u] = ld4(x, i);
          lv[u] = ld4(xl, i);
          p0v[u] = ld4(pprev, i);
          pv[u] = ld4(p, i);
        }
        rv[u] = ld4(r, i);
        qv[u] = ld4(q, i);
        dv[u] = V.dinv[i];
      }
#pragma unroll
      for (int u = 0; u < kUB; ++u) {
        const uint32_t i = i0 + u * kBlock;
        if (i >= r1) continue;
        if constexpr (X) {
          f4 xi = xv[u], li = lv[u];
#pragma unroll
          for (int k = 0; k < kRhs; ++k) {
            float xk = xi[k], lk = li[k];
            if (fold[k]) {
              xk += lk;
              lk = 0.0f;
            }
            comp_axpy(xk, lk, alpha0[k], p0v[u][k]);
            comp_axpy(xk, lk, alpha[k], pv[u][k]);
            xi[k] = xk;
            li[k] = lk;
          }
          st4(x, i, xi);
          st4(xl, i, li);
        }
        const f4 ri = rv[u] - alpha * qv[u];
        st4(r, i, ri);
        const float di = dv[u];
#pragma unroll
        for (int k = 0; k < kRhs; ++k) {
          acc[k] = csum_add(acc[k], ri[k] * ri[k]);
          acc[kRhs + k] = csum_add(acc[kRhs + k], ri[k] * (di * ri[k]));
        }
      }
    }
    tile_partials<2 * kRhs>(acc, partials, t);
    __syncthreads();
  }
}

// Convergence of every running pair, beta, then (odd iterations, after the
// iterate's update) the system's activity.
__global__ __launch_bounds__(kBlock) void k_bt_fin_beta(const CSum* __restrict__ partials,
                                                        const uint32_t* __restrict__ sys_tile, uint32_t nsys,
                                                        int32_t maxiter, PairState* __restrict__ ps,
                                                        int32_t* __restrict__ sys_active, int32_t odd) {
  const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (s >= nsys || !sys_active[s]) return;
  float tot[2 * kRhs];
  sys_sum<2 * kRhs>(partials, sys_tile[s], sys_tile[s + 1], tot);
  if ((threadIdx.x & (kWave - 1)) != 0) return;
#pragma unroll
  for (int k = 0; k < kRhs; ++k) {
    PairState& p = ps[s * kRhs + k];
    if (p.status != kPairRunning) continue;
    p.iter += 1;
    if (odd) p.collapse = 0;  // folded by this iteration's update
    p.rr = tot[k];
    const float rz_new = tot[kRhs + k];
    if (!isfinite(tot[k]) || !isfinite(rz_new)) {
      p.status = kPairBreakdown;
    } else if (tot[k] <= p.thresh) {
      p.status = kPairConverged;
    } else if (p.iter >= maxiter) {
      p.status = kPairMaxIter;
    } else if (!(p.rz > 0.0f) || !(rz_new > 0.0f)) {
      p.status = kPairBreakdown;
    } else {
      p.beta = rz_new / p.rz;
      p.rz = rz_new;
    }
  }
  if (odd) refresh_system(ps, sys_active, s);
}

// True residual r = b - A x of the iterate that would be returned, for every
// row, into r; sums of <r,r>, <r, D^-1 r> per tile (all tiles, active or not).
// EXACT: x = x + xl in FP64 arithmetic (the FP64 result); otherwise
// x = fl(x + xl) with error-free FP32 products and TwoSum rows.
template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_bt_residual(View V, const float* __restrict__ b,
                                                        const float* __restrict__ x, const float* __restrict__ xl,
                                                        float* __restrict__ r, uint32_t ntiles,
                                                        CSum* __restrict__ partials) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t r0 = V.tile_r0[t], r1 = V.tile_r0[t + 1];
    CSum acc[2 * kRhs];
#pragma unroll
    for (int d = 0; d < 2 * kRhs; ++d) acc[d] = csum_zero();
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
      const f4 bi = ld4(b, i);
      f4 ri;
      if constexpr (EXACT) {
        double row[kRhs];
#pragma unroll
        for (int k = 0; k < kRhs; ++k) row[k] = (double)bi[k];
        for (uint32_t e = V.row_ptr[i]; e < V.row_ptr[i + 1]; ++e) {
          const double a = (double)V.val[e];
          const uint32_t j = V.col[e];
          const f4 xj = ld4(x, j), lj = ld4(xl, j);
#pragma unroll
          for (int k = 0; k < kRhs; ++k) row[k] -= a * ((double)xj[k] + (double)lj[k]);
        }
#pragma unroll
        for (int k = 0; k < kRhs; ++k) ri[k] = (float)row[k];
      } else {
        CSum row[kRhs];
#pragma unroll
        for (int k = 0; k < kRhs; ++k) row[k] = CSum{bi[k], 0.0f};
        for (uint32_t e = V.row_ptr[i]; e < V.row_ptr[i + 1]; ++e) {
          const float a = V.val[e];
          const uint32_t j = V.col[e];
          const f4 xj = ld4(x, j) + ld4(xl, j);
#pragma unroll
          for (int k = 0; k < kRhs; ++k) {
            const float pr = a * xj[k];
            const float er = __builtin_fmaf(a, xj[k], -pr);
            row[k] = csum_add(row[k], -pr);
            row[k].c -= er;
          }
        }
#pragma unroll
        for (int k = 0; k < kRhs; ++k) ri[k] = csum_val(row[k]);
      }
      st4(r, i, ri);
      const float di = V.dinv[i];
#pragma unroll
      for (int k = 0; k < kRhs; ++k) {
        acc[k] = csum_add(acc[k], ri[k] * ri[k]);
        acc[kRhs + k] = csum_add(acc[kRhs + k], ri[k] * (di * ri[k]));
      }
    }
    tile_partials<2 * kRhs>(acc, partials, t);
    __syncthreads();
  }
}

// Verification finisher: a pair that stopped as converged but whose true
// residual is above tol restarts from it (beta = 0 makes the next direction
// z); after `max_verify` checks it stops as stagnated. Pairs that stopped for
// other reasons keep their status. rr_true records the true <r,r>.
__global__ __launch_bounds__(kBlock) void k_bt_fin_verify(const CSum* __restrict__ partials,
                                                          const uint32_t* __restrict__ sys_tile, uint32_t nsys,
                                                          int32_t maxiter, int32_t max_verify, int32_t exact,
                                                          PairState* __restrict__ ps,
                                                          int32_t* __restrict__ sys_active) {
  const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (s >= nsys) return;
  float tot[2 * kRhs];
  sys_sum<2 * kRhs>(partials, sys_tile[s], sys_tile[s + 1], tot);
  if ((threadIdx.x & (kWave - 1)) != 0) return;
#pragma unroll
  for (int k = 0; k < kRhs; ++k) {
    PairState& p = ps[s * kRhs + k];
    p.rr_true = tot[k];
    if (p.status != kPairConverged || p.bb == 0.0f) continue;
    if (tot[k] <= p.target) continue;
    p.nverify += 1;
    if (p.nverify > max_verify) {
      p.status = kPairStagnated;
    } else if (p.iter >= maxiter) {
      p.status = kPairMaxIter;
    } else {
      p.status = kPairRunning;
      p.rr = tot[k];
      p.rz = tot[kRhs + k];
      p.beta = 0.0f;
      p.thresh = 0.25f * fminf(p.thresh, p.target);  // aim the restarted recurrence below the target
      p.collapse = exact ? 0 : 1;  // an FP32 result restarts from fl(x + xl)
    }
  }
  refresh_system(ps, sys_active, s);
}

// Inverse diagonal (0 for a missing or zero diagonal: counted in bad).
__global__ __launch_bounds__(kBlock) void k_bt_dinv(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                    uint32_t n, float* __restrict__ dinv, uint32_t* __restrict__ bad) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float d = 0.0f;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
      if (col[e] == i) d += val[e];
    dinv[i] = (d > 0.0f && isfinite(d)) ? 1.0f / d : 0.0f;
    if (!(d > 0.0f && isfinite(d))) atomicAdd(bad, 1u);
  }
}

// Block-diagonal check: every column of row i lies in its system's range
// (sys_of_row via the tile table). Counts violations.
__global__ __launch_bounds__(kBlock) void k_bt_check(const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col,
                                                     const uint32_t* __restrict__ tile_r0,
                                                     const uint32_t* __restrict__ tile_sys,
                                                     const uint32_t* __restrict__ sys_off, uint32_t ntiles,
                                                     uint32_t* __restrict__ bad) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t s = tile_sys[t], lo = sys_off[s], hi = sys_off[s + 1];
    uint32_t nb = 0;
    for (uint32_t i = tile_r0[t] + threadIdx.x; i < tile_r0[t + 1]; i += kBlock)
      for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) nb += (col[e] < lo || col[e] >= hi) ? 1u : 0u;
    if (nb) atomicAdd(bad, nb);
  }
}

// [n][nrhs] <-> [n][4] (zero padding) for callers with fewer right-hand sides.
__global__ __launch_bounds__(kBlock) void k_bt_pack(const float* __restrict__ src, uint32_t n, int32_t nrhs,
                                                    float* __restrict__ dst) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    f4 v = f4{0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < nrhs; ++k) v[k] = src[(size_t)i * nrhs + k];
    st4(dst, i, v);
  }
}

// x out: fl(x + xl) as FP32, or x + xl exactly as FP64, [n][nrhs].
__global__ __launch_bounds__(kBlock) void k_bt_unpack(const float* __restrict__ x, const float* __restrict__ xl,
                                                      uint32_t n, int32_t nrhs, float* __restrict__ dst) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const f4 v = ld4(x, i) + ld4(xl, i);
    for (int k = 0; k < nrhs; ++k) dst[(size_t)i * nrhs + k] = v[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_bt_unpack64(const float* __restrict__ x, const float* __restrict__ xl,
                                                        uint32_t n, int32_t nrhs, double* __restrict__ dst) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const f4 v = ld4(x, i), l = ld4(xl, i);
    for (int k = 0; k < nrhs; ++k) dst[(size_t)i * nrhs + k] = (double)v[k] + (double)l[k];
  }
}

}  // namespace batch
}  // namespace msp
