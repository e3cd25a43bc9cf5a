// CSR SpMV on row blocks of roughly equal nonzero count.
//
// Replaces the reference's one-simdgroup-per-row `spmv_op` / `fused_spvmv`
// (iterative/BiCGSTAB/spmv.metal:11-132). The row range is cut into blocks
// whose first nonzeros fall in consecutive windows of T nonzeros (T picked
// from the mean row length). One 256-thread workgroup handles a block:
//   1. stream the block's values/column indices with 16-byte nontemporal loads,
//      multiply by the gathered x and park the products in LDS;
//   2. reduce each row from LDS with L lanes per row, L the largest power of
//      two with L * rows <= 256 (1 lane for stencil rows of 5-7 entries when
//      rows > 256, 8-16 lanes for 144-entry rows);
//   3. hand every row result to an epilogue (store, store + dot, Richardson
//      update, ...), whose dot contributions go through one deterministic grid
//      reduction.
// Blocks with more than kWin nonzeros (a row longer than the window) take a
// slower whole-workgroup-per-row path inside the same kernel.
// Workgroups are persistent (grid capped near residency) and walk blocks in
// XCD-contiguous order.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace msp {

constexpr uint32_t kCap = 4096;  // LDS product slots per workgroup

template <typename V>
struct CsrView {
  const uint32_t* __restrict__ row_ptr;
  const uint32_t* __restrict__ col;
  const typename V::storage* __restrict__ val;
  const uint32_t* __restrict__ blk;  // [nblk + 1] first row of each block
  const float* __restrict__ dlo;     // 16-bit formats: FP32 remainder of the diagonal
  uint32_t nblk;
  uint32_t n, m;
};

// x loaders --------------------------------------------------------------
// Four columns at once: one 16-byte load when they are consecutive and
// 4-aligned (block-structured rows), else four gathers. Needs a 16-byte
// aligned vector.
__device__ __forceinline__ bool quad_run(u4 c) {
  return c.y == c.x + 1 && c.z == c.x + 2 && c.w == c.x + 3 && (c.x & 3u) == 0;
}

// Every loader also splits a gather into raw(c) (loads only) and make(raw)
// (the arithmetic), so a kernel can issue many gathers before the first use.
struct XPlain {
  const float* __restrict__ x;
  typedef float Raw;
  __device__ __forceinline__ Raw raw(uint32_t c) const { return x[c]; }
  // raw with a 32-bit byte offset (c < 2^30)
  __device__ __forceinline__ Raw raw32(uint32_t c) const { return ld_off32(x, c); }
  __device__ __forceinline__ float make(Raw w) const { return w; }
  __device__ __forceinline__ float operator()(uint32_t c) const { return x[c]; }
  __device__ __forceinline__ f4 load4(u4 c) const {
    if (quad_run(c)) return *reinterpret_cast<const f4*>(x + c.x);
    return f4{x[c.x], x[c.y], x[c.z], x[c.w]};
  }
};

// Per-workgroup set-up of a gather loader and its epilogue, run by every
// thread before the row loop (a loader whose scalar comes from a grid
// reduction computes it here); `halt` is the kernel's gate as loaded, and the
// returned value is the gate the kernel then tests (nonzero: it returns at
// once). A prologue writes nothing when the gate was set on entry, unless it
// says otherwise. The default does nothing; overloads for specific loader
// types are found by argument-dependent lookup.
template <typename XL, typename Epi>
__device__ __forceinline__ int kernel_prologue(XL&, Epi&, int halt) {
  return halt;
}

// s = r - alpha * v formed on the fly (used to fuse a vector update into the
// SpMV that consumes it); the owning rows also get s written by the epilogue.
struct XAxpy {
  const float* __restrict__ r;
  const float* __restrict__ v;
  const float* __restrict__ alpha;
  struct Raw {
    float r, v;
  };
  __device__ __forceinline__ Raw raw(uint32_t c) const { return Raw{r[c], v[c]}; }
  __device__ __forceinline__ float make(Raw w) const { return w.r - (*alpha) * w.v; }
  __device__ __forceinline__ float operator()(uint32_t c) const { return r[c] - (*alpha) * v[c]; }
  __device__ __forceinline__ f4 load4(u4 c) const {
    const float a = *alpha;
    if (quad_run(c))
      return *reinterpret_cast<const f4*>(r + c.x) - a * *reinterpret_cast<const f4*>(v + c.x);
    return f4{r[c.x] - a * v[c.x], r[c.y] - a * v[c.y], r[c.z] - a * v[c.z], r[c.w] - a * v[c.w]};
  }
};

// Quads of one short block held in registers: kSlots quads per thread from a
// 4-aligned start cover any window of at most kWin nonzeros. The host's block
// size keeps every block below kCap - 4 nonzeros unless a single row is longer.
constexpr int kSlots = (int)(kCap / (4 * kBlock));
constexpr uint32_t kWin = 4u * kBlock * kSlots - 3u;
static_assert(kWin <= kCap && kWin + 4 > kCap - 4, "window and LDS size disagree");

template <typename V>
struct BlockLoad {
  uint32_t r0, r1, nz0, nz1;
  u4 c[kSlots];
  typename V::raw4 v[kSlots];
};

__device__ __forceinline__ bool short_block(uint32_t r0, uint32_t r1, uint32_t nz0, uint32_t nz1) {
  return r0 < r1 && nz1 - nz0 <= kWin;
}

// Issue the loads of block b (bounds, then column indices and values of a
// short block) without waiting for them. Slots past the block re-read its
// first quad instead of branching round the load (a lane-divergent branch per
// slot made the compiler wait for each slot's loads before the next). Column
// index and value padding past nnz must be valid (the library pads with zeros).
template <typename V>
__device__ __forceinline__ void block_fetch(const CsrView<V>& A, uint32_t b, BlockLoad<V>& L) {
  L.r0 = A.blk[b];
  L.r1 = A.blk[b + 1];
  L.nz0 = A.row_ptr[L.r0];
  L.nz1 = A.row_ptr[L.r1];
  if (!short_block(L.r0, L.r1, L.nz0, L.nz1)) return;
  const uint32_t q0 = L.nz0 & ~3u, a0 = q0 + 4u * threadIdx.x;
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const uint32_t i = a0 + (uint32_t)k * 4u * kBlock;
    const uint32_t ic = i < L.nz1 ? i : q0;
    L.c[k] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(A.col + ic));
    L.v[k] = V::ld4(A.val + ic);
  }
}

// Products of a fetched short block into sprod (quads at the block edges and
// the re-read slots gather valid columns of neighbours / padding; only the
// block's own entries are kept).
template <typename V, typename XL>
__device__ __forceinline__ void block_products(const BlockLoad<V>& L, const XL& xl, float* sprod) {
  const uint32_t a0 = (L.nz0 & ~3u) + 4u * threadIdx.x;
  f4 xv[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) xv[k] = xl.load4(L.c[k]);  // every gather in flight at once
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const uint32_t i = a0 + (uint32_t)k * 4u * kBlock;
    if (i < L.nz1) {
      const f4 p = V::cvt4(L.v[k]) * xv[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t j = i + q;
        if (j >= L.nz0 && j < L.nz1) sprod[j - L.nz0] = p[q];
      }
    }
  }
}

// Row sums of a short block from sprod, L lanes per row, L the largest power
// of two with L * rows <= kBlock (at most 64).
template <typename V, typename XL, typename Epi>
__device__ __forceinline__ void block_rows(const CsrView<V>& A, uint32_t r0, uint32_t r1, uint32_t nz0, const XL& xl,
                                           const Epi& epi, const float* sprod, CSum* acc) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nr = r1 - r0;
  uint32_t L = 1;
  if (nr < kBlock) {
    uint32_t q = kBlock / nr;  // >= 1
    L = 1u << (31 - __builtin_clz(q));
    if (L > 64) L = 64;
  }
  const uint32_t G = kBlock / L, g = tid / L, lane = tid % L;
  for (uint32_t rb = 0; rb < nr; rb += G) {
    const uint32_t ri = rb + g;
    float s = 0.0f;
    if (ri < nr) {
      const uint32_t s0 = A.row_ptr[r0 + ri] - nz0, s1 = A.row_ptr[r0 + ri + 1] - nz0;
      for (uint32_t j = s0 + lane; j < s1; j += L) s += sprod[j];
    }
    s = group_reduce_sum(s, L);
    if (lane == 0 && ri < nr) {
      if constexpr (V::kLowDiag) {
        if (r0 + ri < A.m) s += A.dlo[r0 + ri] * xl(r0 + ri);
      }
      epi.row(r0 + ri, s, acc);
    }
  }
}

// A block longer than the LDS window: every row reduced by the whole
// workgroup, straight from HBM.
template <typename V, typename XL, typename Epi>
__device__ __forceinline__ void long_rows(const CsrView<V>& A, uint32_t r0, uint32_t r1, const XL& xl,
                                          const Epi& epi, CSum* acc) {
  __shared__ float wpart[kWavesPerBlock];
  const uint32_t tid = threadIdx.x;
  const int wl = tid & (kWave - 1), wid = tid / kWave;
  for (uint32_t r = r0; r < r1; ++r) {
    const uint32_t s0 = A.row_ptr[r], s1 = A.row_ptr[r + 1];
    float s = 0.0f;
    for (uint32_t j = s0 + tid; j < s1; j += kBlock) s += V::load1(A.val + j) * xl(A.col[j]);
    s = wave_reduce_sum(s);
    if (wl == 0) wpart[wid] = s;
    __syncthreads();
    if (tid == 0) {
      float t = 0.0f;
      for (int w = 0; w < kWavesPerBlock; ++w) t += wpart[w];
      if constexpr (V::kLowDiag) {
        if (r < A.m) t += A.dlo[r] * xl(r);
      }
      epi.row(r, t, acc);
    }
    __syncthreads();
  }
}

// Blocks [lo, hi) of one workgroup. Every quad of a block is requested
// before the first product. (Fetching block b + 1 during block b's row sums
// was measured 15% slower: the row sums' own vector loads (row offsets,
// epilogue operands) wait on the in-order load counter, so they queued behind
// the prefetch.)
template <typename V, typename XL, typename Epi>
__device__ __forceinline__ void csr_blocks(const CsrView<V>& A, uint32_t lo, uint32_t hi, uint32_t step,
                                           const XL& xl, const Epi& epi, float* sprod, CSum* acc) {
  BlockLoad<V> L;
  for (uint32_t b = lo; b < hi; b += step) {
    block_fetch<V>(A, b, L);
    if (short_block(L.r0, L.r1, L.nz0, L.nz1)) {
      block_products<V>(L, xl, sprod);
      __syncthreads();
      block_rows<V>(A, L.r0, L.r1, L.nz0, xl, epi, sprod, acc);
      __syncthreads();
    } else if (L.r0 < L.r1) {
      long_rows<V>(A, L.r0, L.r1, xl, epi, acc);
    }
  }
}

// epilogues ------------------------------------------------------------------
struct EpiStore {
  float* __restrict__ y;
  __device__ __forceinline__ void row(uint32_t r, float s, CSum*) const { y[r] = s; }
};

// Epilogues may declare `Pre`, the per-row operands they read, with
// `pre(i)` loading them and `row(i, s, acc, pre)` using them: the diagonal
// kernel then issues those loads with the row's matrix and vector loads
// instead of after the sum (a second round trip per row).
template <typename Epi, typename = void>
struct EpiPre {
  struct type {};
  static constexpr bool has = false;
};
template <typename Epi>
struct EpiPre<Epi, std::void_t<typename Epi::Pre>> {
  using type = typename Epi::Pre;
  static constexpr bool has = true;
};

// An epilogue whose `pre(i)` is the kernel's own gather of row i's column i
// (`xl.raw(i)`) declares `kOwnX = true`; a kernel that already holds that
// value (the marched diagonal walk carries it from the previous level) passes
// it as `pre` instead of loading it again.
template <typename Epi, typename = void>
struct EpiOwnX {
  static constexpr bool value = false;
};
template <typename Epi>
struct EpiOwnX<Epi, std::void_t<decltype(Epi::kOwnX)>> {
  static constexpr bool value = Epi::kOwnX;
};

// A gather loader's Raw type (float for loaders without one).
template <typename XL, typename = void>
struct RawOf {
  typedef float type;
};
template <typename XL>
struct RawOf<XL, std::void_t<typename XL::Raw>> {
  typedef typename XL::Raw type;
};

// y = A x and <y, w>
struct EpiStoreDot {
  float* __restrict__ y;
  const float* __restrict__ w;
  typedef float Pre;
  __device__ __forceinline__ Pre pre(uint32_t r) const { return w[r]; }
  __device__ __forceinline__ void row(uint32_t r, float s, CSum* acc, Pre wr) const {
    y[r] = s;
    acc[0] = csum_add(acc[0], s * wr);
  }
  __device__ __forceinline__ void row(uint32_t r, float s, CSum* acc) const { row(r, s, acc, pre(r)); }
};

// y = A x and <y, x> (the operand of the dot is the product's own input)
struct EpiStoreDotX {
  float* __restrict__ y;
  const float* __restrict__ x;
  typedef float Pre;
  static constexpr bool kOwnX = true;
  __device__ __forceinline__ Pre pre(uint32_t r) const { return x[r]; }
  __device__ __forceinline__ void row(uint32_t r, float s, CSum* acc, Pre xr) const {
    y[r] = s;
    acc[0] = csum_add(acc[0], s * xr);
  }
  __device__ __forceinline__ void row(uint32_t r, float s, CSum* acc) const { row(r, s, acc, pre(r)); }
};

// t = A s with s formed on the fly (XAxpy); also writes s for owned rows and
// accumulates <t, s>, <t, t>.
struct EpiBicgT {
  float* __restrict__ t;
  float* __restrict__ s_out;
  const float* __restrict__ r;
  const float* __restrict__ v;
  const float* __restrict__ alpha;
  __device__ __forceinline__ void row(uint32_t i, float ts, CSum* acc) const {
    float si = r[i] - (*alpha) * v[i];
    s_out[i] = si;
    t[i] = ts;
    acc[0] = csum_add(acc[0], ts * si);
    acc[1] = csum_add(acc[1], ts * ts);
  }
};

// x_new = x_old + omega * (b - A x_old)
struct EpiRichardson {
  float* __restrict__ x_new;
  const float* __restrict__ x_old;
  const float* __restrict__ b;
  float omega;
  __device__ __forceinline__ void row(uint32_t i, float ax, CSum*) const {
    x_new[i] = x_old[i] + omega * (b[i] - ax);
  }
};

// r = b - A x, accumulating <r, r>
struct EpiResidual {
  float* __restrict__ r;
  const float* __restrict__ b;
  __device__ __forceinline__ void row(uint32_t i, float ax, CSum* acc) const {
    float ri = b[i] - ax;
    r[i] = ri;
    acc[0] = csum_add(acc[0], ri * ri);
  }
};

// Finishers run once, in thread 0 of k_finish, with the grid totals.
struct FinNone {
  __device__ __forceinline__ bool active() const { return false; }
  __device__ __forceinline__ void after() const {}
  template <int ND>
  __device__ __forceinline__ void operator()(const float (&)[ND]) const {}
};

struct FinStore {
  float* out;  // [ND]
  __device__ __forceinline__ bool active() const { return true; }
  __device__ __forceinline__ void after() const {}
  template <int ND>
  __device__ __forceinline__ void operator()(const float (&tot)[ND]) const {
    for (int d = 0; d < ND; ++d) out[d] = tot[d];
  }
};

// Block partials of the epilogue sums go to rb (ND > 0); the caller launches
// k_finish<ND> after this kernel.
template <typename V, int ND, typename XL, typename Epi>
__global__ __launch_bounds__(kBlock) void k_spmv(CsrView<V> A, XL xl, Epi epi, GridRed rb,
                                                 const int* __restrict__ done) {
  // the gate is tested after the prologue (whose loads overlap its latency)
  const int halt = kernel_prologue(xl, epi, done ? *done : 0);
  if (halt) return;
  __shared__ float sprod[kCap];
  constexpr int NA = ND > 0 ? ND : 1;
  CSum acc[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) acc[d] = csum_zero();
  const uint32_t G = gridDim.x;
  // grid-stride over blocks in XCD-swizzled order: at every step each XCD
  // takes a contiguous run of blocks, and the grid as a whole sweeps memory
  // front to back
  const uint32_t wg = xcd_swizzle(blockIdx.x, G);
  csr_blocks<V>(A, wg, A.nblk, G, xl, epi, sprod, acc);
  if constexpr (ND > 0) block_partials<ND>(reinterpret_cast<CSum(&)[ND]>(acc), rb);
}

// Row blocks: blk[b] = first row whose first nonzero index is >= b*T
// (lower bound in row_ptr), blk[nblk] = n.
__global__ __launch_bounds__(kBlock) void k_row_blocks(const uint32_t* __restrict__ row_ptr, uint32_t n,
                                                       uint32_t T, uint32_t nblk, uint32_t* __restrict__ blk) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nblk) return;
  if (b == nblk) {
    blk[b] = n;
    return;
  }
  uint64_t target = (uint64_t)b * T;
  uint32_t lo = 0, hi = n;  // search in row_ptr[0..n]
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo) >> 1);
    if ((uint64_t)row_ptr[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  blk[b] = lo;
}

// Maximum row length (for block sizing): one atomicMax per workgroup (one per
// wave on a single address serialised into ~0.2 ms at 2M rows).
__global__ __launch_bounds__(kBlock) void k_max_row(const uint32_t* __restrict__ row_ptr, uint32_t n,
                                                    uint32_t* __restrict__ out) {
  __shared__ uint32_t wmax[kWavesPerBlock];
  uint32_t m = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    uint32_t l = row_ptr[r + 1] - row_ptr[r];
    m = l > m ? l : m;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(m, off, kWave);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWavesPerBlock; ++w) m = wmax[w] > m ? wmax[w] : m;
    atomicMax(out, m);
  }
}

// Diagonal (duplicates summed) -> inverse for Jacobi; zero diagonal -> flag.
__global__ __launch_bounds__(kBlock) void k_inv_diag(const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                     uint32_t n, float* __restrict__ dinv, int* __restrict__ bad) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    float d = 0.0f;
    for (uint32_t j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
      if (col[j] == r) d += val[j];
    if (d == 0.0f || !isfinite(d)) {
      atomicOr(bad, 1);
      dinv[r] = 0.0f;
    } else {
      dinv[r] = 1.0f / d;
    }
  }
}

// FP32 -> 16-bit value copies for the reduced-precision path.
// 16-bit copy of the values that keeps every row sum: off-diagonal a_ij are
// rounded, the diagonal absorbs their rounding defects,
//   d_i = a_ii + sum_{j != i} (a_ij - round(a_ij)),
// and is stored as round(d_i) in the diagonal slot plus the FP32 remainder
// dlo_i = d_i - round(d_i). For an M-matrix (diffusion) the 16-bit operator is
// then the same operator with every coupling perturbed by at most half an ulp
// of the format, so its energy norm is within that factor of A's. Duplicate
// diagonal entries: the first slot carries round(d_i), the others 0. Rows
// i >= m (no diagonal column) keep plain rounding. One thread per row;
// padding past nnz is written as 0.
template <bool BF16>
__global__ __launch_bounds__(kBlock) void k_cvt16_rowsum(const uint32_t* __restrict__ row_ptr,
                                                          const uint32_t* __restrict__ col,
                                                          const float* __restrict__ in, uint32_t n, uint32_t m,
                                                          uint64_t nnz, uint16_t* __restrict__ out,
                                                          float* __restrict__ dlo) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = nnz + tid; e < nnz + 64; e += stride) out[e] = 0;
  for (uint64_t i = tid; i < n; i += stride) {
    double d = 0.0;
    int64_t dslot = -1;
    for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      const float a = in[k];
      if (col[k] == (uint32_t)i) {
        d += (double)a;
        if (dslot < 0) dslot = k;
        else out[k] = 0;
      } else {
        const uint16_t h = BF16 ? f32_to_bf16(a) : f32_to_f16(a);
        out[k] = h;
        d += (double)a - (double)(BF16 ? ValBF16::cvt(h) : ValF16::cvt(h));
      }
    }
    float hi = 0.0f;
    if (dslot >= 0) {
      const uint16_t h = BF16 ? f32_to_bf16((float)d) : f32_to_f16((float)d);
      out[dslot] = h;
      hi = BF16 ? ValBF16::cvt(h) : ValF16::cvt(h);
    }
    dlo[i] = i < m ? (float)(d - (double)hi) : 0.0f;  // rows past the last column have no diagonal
  }
}

}  // namespace msp
