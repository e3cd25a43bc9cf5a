// Diagonal (offset) storage for stencil-like matrices, used in place of the
// CSR stream when every nonzero lies on one of at most kMaxDiag diagonals
// (column - row = off[d]) and the diagonals are mostly full. Values are stored
// diagonal-major, val[d * ld + i] = A(i, i + off[d]) (zero where the diagonal
// leaves the matrix or has no entry), so a row costs only its values (no
// column indices, no row pointers) and every load is coalesced: the 7-point
// operator moves 28 B of values per row instead of 56 B of values + columns.
// The row results feed the same epilogues as the CSR kernel.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "spmv.hpp"

namespace msp {

constexpr int kMaxDiag = 16;
// rows and columns of a matrix in diagonal storage stay below this (32-bit
// byte offsets into value planes and vectors)
constexpr uint32_t kDiaMaxDim = 1u << 30;
// largest row count walked XCD-contiguously by the diagonal kernel
constexpr uint32_t kDiaXcdRows = 1u << 23;

// Diagonal d (column - row = off[d]) of row i is read from
// val[plane[d] * ld + i + shift[d]]: plane[d] = d, shift[d] = 0 in plain
// storage; for a matrix found exactly symmetric, a lower diagonal (off = -k)
// reads the upper diagonal +k one row block earlier (A(i, i-k) = A(i-k, i):
// plane of +k, shift -k), so the 7-point operator streams 4 value planes
// (16 B per row) instead of 7 (28 B) -- the same values, summed in the same
// order, so the products are bit-identical.
struct DiaOffsets {
  int32_t off[kMaxDiag];
  int32_t plane[kMaxDiag];
  int32_t shift[kMaxDiag];
  int32_t ndiag;
};

// Plain storage: every diagonal reads its own plane.
inline void dia_plain(DiaOffsets& o) {
  for (int d = 0; d < kMaxDiag; ++d) {
    o.plane[d] = d;
    o.shift[d] = 0;
  }
}

template <typename V>
struct DiaView {
  const typename V::storage* __restrict__ val;  // [ndiag][ld]
  const float* __restrict__ dlo;                // 16-bit formats: FP32 remainder of the diagonal
  uint64_t ld;
  uint32_t n, m;
  int32_t mirrored;  // o.plane / o.shift in use (else plane d, shift 0)
  DiaOffsets o;
};

// Groups of U * kBlock consecutive rows (order below); a thread
// takes rows tid, tid + kBlock, ... of its workgroup's group (U rows a full
// grid stride apart were slower: the grid then streams U distant regions at
// once). Plain coalesced loads (a non-temporal hint on these 4-byte lane
// loads more than doubled the time); partial sums of the epilogue go to rb.
// Every load of a thread's U rows is issued before any arithmetic on loaded
// data (value widening, the gather's own arithmetic): arithmetic under the
// per-diagonal branch makes each diagonal wait for its own loads (2.5 TB/s
// on the 7-point operator at 256^3). Columns outside the
// matrix read a clamped in-range x and are then left out of the sum, so the
// result equals the guarded form bit for bit; rows past the end of the
// U-group load row i0's data and store nothing. Requires m > 0.
// A loader with 32-bit gathers (raw32).
template <typename XL, typename = void>
struct HasRaw32 : std::false_type {};
template <typename XL>
struct HasRaw32<XL, std::void_t<decltype(std::declval<const XL&>().raw32(0u))>> : std::true_type {};

template <typename V, int ND, typename XL, typename Epi, int NDIAG, int U>
__device__ __forceinline__ void dia_rows(const DiaView<V>& A, XL xl, Epi epi, GridRed rb,
                                         const int* __restrict__ done) {
  // xl and epi are copies: the prologue sets their step scalars
  // NDIAG > 0: exactly NDIAG diagonals (no per-diagonal branch; stencils and
  // their triangular factors); 0: A.o.ndiag <= kMaxDiag
  constexpr int DMAX = NDIAG > 0 ? NDIAG : kMaxDiag;
  // the gate is tested after the prologue (whose loads overlap its latency)
  const int halt = kernel_prologue(xl, epi, done ? *done : 0);
  if (halt) return;
  constexpr int NA = ND > 0 ? ND : 1;
  CSum acc[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) acc[d] = csum_zero();
  const int nd = NDIAG > 0 ? NDIAG : A.o.ndiag;
  // W: rows, columns and byte offsets in 32 bits (n, m < kDiaMaxDim), each
  // load one offset register against a base in scalar registers, for loaders
  // with 32-bit gathers (raw32) and up to two sums: the 7-point product at
  // 256^3 took 82 -> 74 us with BF16 values, 91 -> 89 us with FP32; at 128^3
  // the FSAI-PCG step's products with sums 11.3 -> 10.1 us (q = A p with
  // <p,q>) and 13.3 -> 13.0 us (the x / r update with w = G r). (Loaders
  // without raw32, such as those forming the direction from two vectors, were
  // 5-15% slower in this form when it was tried.)
  constexpr bool W = HasRaw32<XL>::value && ND <= 2;
  using Row = typename std::conditional<W, uint32_t, uint64_t>::type;
  using Col = typename std::conditional<W, int32_t, int64_t>::type;
  const Col m = (Col)A.m;
  const Row n = A.n;
  // a workgroup's U-group is U contiguous runs of kBlock rows
  const Row stride = kBlock;
  // 16-bit values held one per register (packing two per register would
  // wait for the loads)
  using S = typename V::storage;
  using Held = typename std::conditional<sizeof(S) == 2, uint32_t, float>::type;
  // start of each diagonal's values, relative to row 0: plane * ld + shift.
  // A mirrored lower diagonal (shift -k) reads the plane of +k (index >= 1,
  // after the lower planes), so its rows i < k still address memory inside
  // the array; those values are left out of the sum (column < 0)
  int64_t base[DMAX];
  if (A.mirrored) {
#pragma unroll
    for (int d = 0; d < DMAX; ++d) base[d] = d < nd ? (int64_t)A.o.plane[d] * (int64_t)A.ld + A.o.shift[d] : 0;
  } else {
#pragma unroll
    for (int d = 0; d < DMAX; ++d) base[d] = (int64_t)d * (int64_t)A.ld;
  }
  [[maybe_unused]] const S* vd[DMAX];
  if constexpr (W) {
#pragma unroll
    for (int d = 0; d < DMAX; ++d) vd[d] = scalar_ptr(A.val + base[d]);
  }
  // Row groups of U * kBlock rows. Up to kDiaXcdRows rows they are walked
  // XCD-contiguously: the workgroups on one XCD (dispatch is round-robin,
  // workgroup b on XCD b % 8) take one eighth of the groups in order, so the
  // x gathers at columns i +- k and a diagonal's reads at rows i - k meet
  // lines the same XCD's L2 fetched shortly before (128^3: plain CG 527 ->
  // 506 ms, the learned-FSAI solve 17.4 -> 16.9 ms). Larger matrices are
  // walked grid-stride: at 256^3 the rows between such reuses outgrow an
  // XCD's L2 either way (the 7-point product fetched 494 against 508 MB) and
  // the CG step was 2.5% slower.
  const Row ngroups = (n + (Row)(kBlock * U) - 1) / (Row)(kBlock * U);
  Row gfirst = blockIdx.x, glast = ngroups, gstep = gridDim.x;
  if (gridDim.x >= (uint32_t)kNumXcd && n <= (Row)kDiaXcdRows) {
    const uint32_t xcd = blockIdx.x % kNumXcd, idx = blockIdx.x / kNumXcd;
    gstep = gridDim.x / kNumXcd + (xcd < gridDim.x % kNumXcd ? 1u : 0u);
    gfirst = ngroups * xcd / kNumXcd + idx;
    glast = ngroups * (xcd + 1) / kNumXcd;
  }
  for (Row grp = gfirst; grp < glast; grp += gstep) {
    const Row i0 = grp * (Row)(kBlock * U) + threadIdx.x;
    if (i0 >= n) continue;
    Held a[U][DMAX];
    typename XL::Raw xr[U][DMAX];
    [[maybe_unused]] float dl[U];
    [[maybe_unused]] typename XL::Raw xd[U];
    [[maybe_unused]] typename EpiPre<Epi>::type pre[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const Row iu = i0 + (Row)u * stride;
      const Row i = iu < n ? iu : i0;
#pragma unroll
      for (int d = 0; d < DMAX; ++d) {
        a[u][d] = {};
        xr[u][d] = {};
        if (d < nd) {  // loads only: no arithmetic on loaded data before the last load
          const Col c = (Col)i + A.o.off[d];
          const uint32_t cc = (uint32_t)(c < 0 ? 0 : (c >= m ? m - 1 : c));
          if constexpr (W) {
            a[u][d] = ld_off32(vd[d], i);
            xr[u][d] = xl.raw32(cc);
          } else {
            a[u][d] = A.val[base[d] + (int64_t)i];
            xr[u][d] = xl.raw(cc);
          }
        }
      }
      if constexpr (V::kLowDiag) {
        if constexpr (W) dl[u] = ld_off32(A.dlo, i);
        else dl[u] = A.dlo[i];
        const uint32_t ci = (uint32_t)((Col)i < m ? (Col)i : m - 1);
        if constexpr (W) xd[u] = xl.raw32(ci);
        else xd[u] = xl.raw(ci);
      }
      if constexpr (EpiPre<Epi>::has) pre[u] = epi.pre((uint32_t)i);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const Row i = i0 + (Row)u * stride;
      if (i < n) {
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d < DMAX; ++d) {
          if (d < nd) {
            const Col c = (Col)i + A.o.off[d];
            s = (c >= 0 && c < m) ? s + V::cvt1((S)a[u][d]) * xl.make(xr[u][d]) : s;
          }
        }
        if constexpr (V::kLowDiag) {
          if ((Col)i < m) s += dl[u] * xl.make(xd[u]);
        }
        if constexpr (EpiPre<Epi>::has) epi.row((uint32_t)i, s, acc, pre[u]);
        else epi.row((uint32_t)i, s, acc);
      }
    }
  }
  if constexpr (ND > 0) block_partials<ND>(reinterpret_cast<CSum(&)[ND]>(acc), rb);
}

template <typename V, int ND, typename XL, typename Epi, int NDIAG, int U = 1>
__global__ __launch_bounds__(kBlock) void k_dia(DiaView<V> A, XL xl, Epi epi, GridRed rb,
                                                const int* __restrict__ done) {
  dia_rows<V, ND, XL, Epi, NDIAG, U>(A, xl, epi, rb, done);
}

// The same rows for epilogues with sums (ND > 0), compiled for at most 6 waves
// per SIMD. Aimed at 8 waves (64 VGPRs), the scheduler issued the first row's
// loads of the 7-point product with <p, y> one diagonal at a time, each
// waiting for the last (135 us at 256^3 against 90 us without the sum), in a
// kernel that then took 59 VGPRs anyway; aimed at 6 it keeps every load of
// the U rows in flight.
template <typename V, int ND, typename XL, typename Epi, int NDIAG, int U = 1>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 6))) void k_dia_sum(
    DiaView<V> A, XL xl, Epi epi, GridRed rb, const int* __restrict__ done) {
  dia_rows<V, ND, XL, Epi, NDIAG, U>(A, xl, epi, rb, done);
}

// ---------------------------------------------------------------------------
// Rows marched along the far diagonal pair of symmetric storage. With offsets
// +-c (c the largest), row i's -c entry is the +c plane's value at row i - c
// and its gather x[i - c]; the walk above streams that plane and x a second
// time, c rows after the first read, and once c rows of all the planes outgrow
// an XCD's L2 (the 7-point operator at 256^3: c = 65536 rows, 1.5 MB of planes
// per XCD between the two reads) the second read comes back through the
// fabric (1.45x the minimum traffic). Here a thread owns the rows
// t, t + c, t + 2c, ... of one level range: the +c value and x[i + c] it loads
// for row i are, one level later, the -c value and x[i] of row i + c, carried
// in registers, and x[i - c] is the previous level's x[i]. A work item is a
// tile of kBlock rows within [0, c) times a range of levels; the items go to
// the XCDs in contiguous ranges (one level range, consecutive tiles), so the
// +-1 and +-nx neighbours of a tile are fetched by the same L2. Each row's sum
// takes its diagonals in the same order as dia_rows, so the products are bit
// for bit the same.
struct DiaMarch {
  uint32_t c;       // the far offset
  uint32_t ntile;   // tiles of kBlock rows covering [0, c)
  uint32_t nlev;    // levels: ceil(n / c)
  uint32_t lpc;     // levels per work item
  uint32_t nwork;   // work items (= the grid): ntile * ceil(nlev / lpc)
};

template <typename V, int ND, typename XL, typename Epi, int NDIAG, int UZ>
__device__ __forceinline__ void dia_march_rows(const DiaView<V>& A, const DiaMarch& mk, XL xl, Epi epi, GridRed rb,
                                               const int* __restrict__ done) {
  static_assert(NDIAG >= 3 && (NDIAG & 1), "the march needs -c, 0, +c and mirrored pairs");
  const int halt = kernel_prologue(xl, epi, done ? *done : 0);
  if (halt) return;
  constexpr int NA = ND > 0 ? ND : 1;
  CSum acc[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) acc[d] = csum_zero();
  using S = typename V::storage;
  using Held = typename std::conditional<sizeof(S) == 2, uint32_t, float>::type;
  using XR = typename XL::Raw;
  // offsets ascend and pair up: -c is diagonal 0, +c the last, the main
  // diagonal the middle one (the host checks all three). Every index into the
  // register arrays below is a compile-time constant, so a carried value is
  // never copied while its load is in flight.
  constexpr int dlo = 0, dhi = NDIAG - 1, dmid = NDIAG / 2;
  // rows, columns and byte offsets in 32 bits (n, m < kDiaMaxDim); each
  // diagonal's values from a base in scalar registers
  const uint32_t n = A.n, c = mk.c;
  const int32_t m = (int32_t)A.m;
  const S* vd[NDIAG];
#pragma unroll
  for (int d = 0; d < NDIAG; ++d)
    vd[d] = scalar_ptr(A.val + ((int64_t)A.o.plane[d] * (int64_t)A.ld + A.o.shift[d]));
  const uint32_t w = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t tile = w % mk.ntile, chunk = w / mk.ntile;
  const uint32_t t = tile * (uint32_t)kBlock + threadIdx.x;  // row of level 0
  const uint32_t k0 = chunk * mk.lpc;
  const uint32_t k1 = k0 + mk.lpc < mk.nlev ? k0 + mk.lpc : mk.nlev;
  // rows past n (last level) or t >= c (last tile) load row `last` and store
  // nothing; all later levels of such a thread are past n as well
  const uint32_t last = n - 1;
  if (t < c && k0 < k1) {
    // carried: the +c value of row i - c, x[i - c], x[i]
    const uint32_t i0 = t + k0 * c;  // < nlev * c < n + c < 2^31
    const uint32_t ia = i0 < n ? i0 : last;
    Held vfar = {};
    XR xm = xl.raw(0), xo;
    if (k0 > 0) {
      vfar = ld_off32(vd[dlo], ia);  // the +c plane at row ia - c
      xm = xl.raw(ia - c);
    }
    xo = xl.raw((int32_t)ia < m ? ia : (uint32_t)(m - 1));
    for (uint32_t k = k0; k < k1; k += UZ) {
      Held a[UZ][NDIAG];
      XR xr[UZ][NDIAG];
      [[maybe_unused]] float dl[UZ];
      [[maybe_unused]] typename EpiPre<Epi>::type pre[UZ];
      uint32_t iu[UZ];
      bool ok[UZ];
#pragma unroll
      for (int u = 0; u < UZ; ++u) {
        const uint32_t ii = t + (k + u) * c;
        ok[u] = (k + u < k1) && ii < n;
        const uint32_t i = ok[u] ? ii : last;
        iu[u] = i;
#pragma unroll
        for (int d = 0; d < NDIAG; ++d) {
          if (d == dlo) {  // value and x of -c: carried
            a[u][d] = u == 0 ? vfar : a[u - 1][dhi];
            continue;
          }
          // the diagonal's and the +c plane's values are read once (the
          // +c plane's second use is carried): nontemporal; the other upper
          // planes are read again by the mirrored lower diagonals from L2
          if (d == dmid || d == dhi) a[u][d] = __builtin_nontemporal_load(vd[d] + i);
          else a[u][d] = ld_off32(vd[d], i);
          if (d == dmid) continue;  // x[i]: carried
          const int32_t col = (int32_t)i + A.o.off[d];
          const uint32_t cc = (uint32_t)(col < 0 ? 0 : (col >= m ? m - 1 : col));
          xr[u][d] = xl.raw(cc);
        }
        if constexpr (V::kLowDiag) dl[u] = ld_off32(scalar_ptr(A.dlo), i);
        if constexpr (EpiPre<Epi>::has && !EpiOwnX<Epi>::value) pre[u] = epi.pre(i);
      }
#pragma unroll
      for (int u = 0; u < UZ; ++u) {
        xr[u][dmid] = u == 0 ? xo : xr[u - 1][dhi];
        xr[u][dlo] = u == 0 ? xm : xr[u - 1][dmid];
        if constexpr (EpiOwnX<Epi>::value) pre[u] = xr[u][dmid];
      }
#pragma unroll
      for (int u = 0; u < UZ; ++u) {
        if (ok[u]) {
          const uint32_t i = iu[u];
          float s = 0.0f;
#pragma unroll
          for (int d = 0; d < NDIAG; ++d) {
            const int32_t col = (int32_t)i + A.o.off[d];
            s = (uint32_t)col < (uint32_t)m ? s + V::cvt1((S)a[u][d]) * xl.make(xr[u][d]) : s;
          }
          if constexpr (V::kLowDiag) {
            if ((int32_t)i < m) s += dl[u] * xl.make(xr[u][dmid]);
          }
          if constexpr (EpiPre<Epi>::has) epi.row(i, s, acc, pre[u]);
          else epi.row(i, s, acc);
        }
      }
      vfar = a[UZ - 1][dhi];
      xm = xr[UZ - 1][dmid];
      xo = xr[UZ - 1][dhi];
    }
  }
  if constexpr (ND > 0) block_partials<ND>(reinterpret_cast<CSum(&)[ND]>(acc), rb);
}

template <typename V, int ND, typename XL, typename Epi, int NDIAG, int UZ>
__global__ __launch_bounds__(kBlock) void k_dia_march(DiaView<V> A, DiaMarch mk, XL xl, Epi epi, GridRed rb,
                                                      const int* __restrict__ done) {
  dia_march_rows<V, ND, XL, Epi, NDIAG, UZ>(A, mk, xl, epi, rb, done);
}

// with sums: compiled without k_dia_sum's 6-wave bound (every load of a
// round trip is issued first here anyway; at 256^3 the CG step's product with
// <p, q> took 107.8 us with the bound, 92.6 us without)
template <typename V, int ND, typename XL, typename Epi, int NDIAG, int UZ>
__global__ __launch_bounds__(kBlock) void k_dia_march_sum(DiaView<V> A, DiaMarch mk, XL xl, Epi epi, GridRed rb,
                                                          const int* __restrict__ done) {
  dia_march_rows<V, ND, XL, Epi, NDIAG, UZ>(A, mk, xl, epi, rb, done);
}

// Rows sampled by the diagonal probe: for sample k (one thread each),
// out[k * (kMaxDiag + 1)] = row length, then up to kMaxDiag column indices.
constexpr int kDiaSamples = 64;
struct DiaSample {
  uint32_t row[kDiaSamples];
};

__global__ __launch_bounds__(kDiaSamples) void k_dia_sample(const uint32_t* __restrict__ row_ptr,
                                                            const uint32_t* __restrict__ col, DiaSample smp,
                                                            uint32_t* __restrict__ out) {
  const uint32_t k = threadIdx.x, r = smp.row[k];
  const uint32_t b = row_ptr[r], len = row_ptr[r + 1] - b;
  uint32_t* o = out + k * (kMaxDiag + 1);
  o[0] = len;
  for (uint32_t j = 0; j < len && j < (uint32_t)kMaxDiag; ++j) o[1 + j] = col[b + j];
}

// Entries whose offset is not in the candidate list (counted), per row.
__global__ __launch_bounds__(kBlock) void k_dia_check(const uint32_t* __restrict__ row_ptr,
                                                      const uint32_t* __restrict__ col, uint32_t n, DiaOffsets o,
                                                      uint32_t* __restrict__ misses) {
  uint32_t miss = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const int64_t off = (int64_t)col[e] - (int64_t)i;
      bool hit = false;
      for (int d = 0; d < o.ndiag; ++d) hit |= (off == o.off[d]);
      miss += hit ? 0u : 1u;
    }
  }
#pragma unroll
  for (int sh = kWave / 2; sh > 0; sh >>= 1) miss += __shfl_xor(miss, sh, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && miss) atomicAdd(misses, miss);
}

// Symmetry of filled FP32 diagonal storage: for every lower diagonal d
// (off[d] = -k, partner p with off[p] = +k) count rows i >= k whose value
// differs from the partner's at row i - k (bitwise).
__global__ __launch_bounds__(kBlock) void k_dia_symcheck(const float* __restrict__ val, uint32_t n, DiaOffsets o,
                                                         const int32_t* __restrict__ partner, uint64_t ld,
                                                         uint32_t* __restrict__ misses) {
  uint32_t miss = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    for (int d = 0; d < o.ndiag; ++d) {
      const int64_t k = -(int64_t)o.off[d];
      if (k <= 0 || (int64_t)i < k) continue;
      const uint32_t a = __float_as_uint(val[(uint64_t)d * ld + i]);
      const uint32_t b = __float_as_uint(val[(uint64_t)partner[d] * ld + (uint64_t)((int64_t)i - k)]);
      miss += a != b ? 1u : 0u;
    }
  }
#pragma unroll
  for (int sh = kWave / 2; sh > 0; sh >>= 1) miss += __shfl_xor(miss, sh, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && miss) atomicAdd(misses, miss);
}

// Scatter the FP32 CSR values into diagonal storage (zero-filled by the
// caller); duplicates of an entry are added in CSR order.
__global__ __launch_bounds__(kBlock) void k_dia_fill(const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                     uint32_t n, DiaOffsets o, uint64_t ld, float* __restrict__ out) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const int64_t off = (int64_t)col[e] - (int64_t)i;
      for (int d = 0; d < o.ndiag; ++d) {
        if (off == o.off[d]) {
          out[d * ld + i] += val[e];
          break;
        }
      }
    }
  }
}

// 16-bit diagonal storage with exact row sums, as k_cvt16_rowsum does for CSR:
// off-diagonals rounded, the diagonal (offset 0, index dd; -1 if absent)
// absorbs their rounding defects and is stored as a 16-bit value plus the FP32
// remainder dlo. SUMS = false: every value rounded on its own, dlo = 0.
template <bool BF16, bool SUMS = true>
__global__ __launch_bounds__(kBlock) void k_dia_cvt16(const float* __restrict__ in, uint32_t n, uint32_t m,
                                                      DiaOffsets o, int dd, uint64_t ld, uint16_t* __restrict__ out,
                                                      float* __restrict__ dlo) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if constexpr (!SUMS) {
      for (int d = 0; d < o.ndiag; ++d) {
        const float a = in[d * ld + i];
        out[d * ld + i] = BF16 ? f32_to_bf16(a) : f32_to_f16(a);
      }
      dlo[i] = 0.0f;
      continue;
    }
    double dsum = 0.0;
    for (int d = 0; d < o.ndiag; ++d) {
      const float a = in[d * ld + i];
      if (d == dd) {
        dsum += (double)a;
        continue;
      }
      const uint16_t h = BF16 ? f32_to_bf16(a) : f32_to_f16(a);
      out[d * ld + i] = h;
      dsum += (double)a - (double)(BF16 ? ValBF16::cvt(h) : ValF16::cvt(h));
    }
    float hi = 0.0f;
    if (dd >= 0 && i < m) {
      const uint16_t h = BF16 ? f32_to_bf16((float)dsum) : f32_to_f16((float)dsum);
      out[dd * ld + i] = h;
      hi = BF16 ? ValBF16::cvt(h) : ValF16::cvt(h);
    }
    dlo[i] = i < m ? (float)(dsum - (double)hi) : 0.0f;
  }
}

}  // namespace msp
