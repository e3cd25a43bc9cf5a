// Block-CSR (BCSR) kernels: register-blocked SpMV, true residual, diagonal,
// and the GPU CSR -> BCSR conversion.
//
// Replaces the reference's 4x4-only BCSR path (iterative/BiCGSTAB/bcsr_spmv.metal:30-138,
// one thread per block row) and its CPU scipy `tobsr` preprocessing
// (iterative/BiCGSTAB/bcsr_benchmark.py:36-66). Blocks are row-major B x B
// (the scipy BSR `data` layout the reference uploads). For B in {4, 8, 16} a
// 64-lane wave covers 64 / (B*B/4) block rows: lane (i, q) of a block row
// loads the 16-byte quad q of row i of every block of that block row (a wave
// instruction moves 1 KiB of contiguous block values), multiplies it with the
// matching x quad, and the B/4 lanes of a row reduce with two shuffles at most.
#pragma once

#include "common.hpp"

namespace msp {

struct BcsrView {
  const uint32_t* __restrict__ brow_ptr;  // [nbr + 1]
  const uint32_t* __restrict__ bcol;      // [nblk]
  const float* __restrict__ blocks;       // [nblk][B][B]
  uint32_t nbr;
  uint32_t n;  // logical rows
};

struct X4Plain {
  const float* __restrict__ x;
  __device__ __forceinline__ f4 load4(uint32_t c) const { return *reinterpret_cast<const f4*>(x + c); }
};

struct X4Axpy {  // r - alpha v
  const float* __restrict__ r;
  const float* __restrict__ v;
  const float* __restrict__ alpha;
  __device__ __forceinline__ f4 load4(uint32_t c) const {
    const float a = *alpha;
    return *reinterpret_cast<const f4*>(r + c) - a * *reinterpret_cast<const f4*>(v + c);
  }
};

template <int B, int ND, typename XL, typename Epi>
__global__ __launch_bounds__(kBlock) void k_bspmv(BcsrView A, XL xl, Epi epi, GridRed rb,
                                                  const int* __restrict__ done) {
  static_assert(B == 4 || B == 8 || B == 16, "block size");
  if (kernel_prologue(xl, epi, done ? *done : 0)) return;
  constexpr int Q = B / 4;                    // quads per block line
  constexpr int G = B * Q;                    // lanes per block row
  constexpr int W = kWave / G;                // block rows per wave
  constexpr int RPB = W * kWavesPerBlock;     // block rows per workgroup step
  constexpr int NA = ND > 0 ? ND : 1;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int g = lane / G, li = lane % G, i = li / Q, cq = li % Q;
  CSum acc[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) acc[d] = csum_zero();
  const uint32_t units = (A.nbr + RPB - 1) / RPB;
  const uint32_t Gd = gridDim.x;
  // grid-stride in XCD-swizzled order (as the CSR kernel)
  const uint32_t wg = xcd_swizzle(blockIdx.x, Gd);
  for (uint32_t u = wg; u < units; u += Gd) {
    const uint32_t br = u * RPB + wid * W + g;
    float sum = 0.0f;
    if (br < A.nbr) {
      const uint32_t k0 = A.brow_ptr[br], k1 = A.brow_ptr[br + 1];
      const float* base = A.blocks + (size_t)i * B + cq * 4;
      uint32_t k = k0;
      for (; k + 1 < k1; k += 2) {  // two blocks in flight per lane
        f4 a0 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(base + (size_t)k * B * B));
        f4 a1 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(base + (size_t)(k + 1) * B * B));
        const uint32_t c0 = A.bcol[k], c1 = A.bcol[k + 1];
        f4 x0 = xl.load4(c0 * B + cq * 4);
        f4 x1 = xl.load4(c1 * B + cq * 4);
        sum += a0.x * x0.x + a0.y * x0.y + a0.z * x0.z + a0.w * x0.w;
        sum += a1.x * x1.x + a1.y * x1.y + a1.z * x1.z + a1.w * x1.w;
      }
      if (k < k1) {
        f4 a0 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(base + (size_t)k * B * B));
        f4 x0 = xl.load4(A.bcol[k] * B + cq * 4);
        sum += a0.x * x0.x + a0.y * x0.y + a0.z * x0.z + a0.w * x0.w;
      }
    }
#pragma unroll
    for (int off = 1; off < Q; off <<= 1) sum += __shfl_xor(sum, off, kWave);
    if (cq == 0 && br < A.nbr) {
      const uint32_t row = br * B + i;
      if (row < A.n) epi.row(row, sum, acc);
    }
  }
  if constexpr (ND > 0) block_partials<ND>(reinterpret_cast<CSum(&)[ND]>(acc), rb);
}

// Row r of a BCSR matrix as (value, column) pairs, for row-per-thread kernels.
template <typename F>
__device__ __forceinline__ void bcsr_row_foreach(const BcsrView& A, uint32_t B, uint32_t m, uint32_t r, F&& f) {
  const uint32_t br = r / B, ii = r % B;
  for (uint32_t k = A.brow_ptr[br]; k < A.brow_ptr[br + 1]; ++k) {
    const float* blk = A.blocks + (size_t)k * B * B + (size_t)ii * B;
    const uint32_t c0 = A.bcol[k] * B;
    for (uint32_t j = 0; j < B; ++j)
      if (c0 + j < m) f(blk[j], c0 + j);
  }
}

struct RowsBcsr {
  BcsrView A;
  uint32_t B, m;
  template <typename F>
  __device__ __forceinline__ void foreach(uint32_t i, F&& f) const {
    bcsr_row_foreach(A, B, m, i, f);
  }
};

__global__ __launch_bounds__(kBlock) void k_inv_diag_bcsr(BcsrView A, uint32_t B, uint32_t m,
                                                          float* __restrict__ dinv, int* __restrict__ bad) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < A.n; r += gridDim.x * blockDim.x) {
    float d = 0.0f;
    bcsr_row_foreach(A, B, m, r, [&](float a, uint32_t c) {
      if (c == r) d += a;
    });
    if (d == 0.0f || !isfinite(d)) {
      atomicOr(bad, 1);
      dinv[r] = 0.0f;
    } else {
      dinv[r] = 1.0f / d;
    }
  }
}

// ---------------------------------------------------------------------------
// CSR -> BCSR. The CSR's nonzeros are sorted by block key (block row, block
// col) with their CSR position as payload (the radix sort of sort.hpp, stable),
// which leaves the nonzeros of each block adjacent and in CSR order (row-major
// within the block, duplicates of one (row, col) next to each other).

// keys[i] = (row/B << 32) | col/B, payload[i] = i (bit pattern carried as
// float), rows[i] = row. L lanes walk one row (coalesced for long rows).
__global__ __launch_bounds__(kBlock) void k_block_keys(const uint32_t* __restrict__ row_ptr,
                                                       const uint32_t* __restrict__ col, uint32_t n, uint32_t B,
                                                       uint32_t L, uint64_t* __restrict__ keys,
                                                       float* __restrict__ payload, uint32_t* __restrict__ rows) {
  const uint32_t lane = threadIdx.x % L;
  const uint32_t groups = gridDim.x * (kBlock / L);
  for (uint32_t r = blockIdx.x * (kBlock / L) + threadIdx.x / L; r < n; r += groups) {
    for (uint32_t j = row_ptr[r] + lane; j < row_ptr[r + 1]; j += L) {
      keys[j] = ((uint64_t)(r / B) << 32) | (col[j] / B);
      payload[j] = __uint_as_float(j);
      rows[j] = r;
    }
  }
}

// flag[i] = 1 where a new block starts in the block-key-sorted order.
__global__ __launch_bounds__(kBlock) void k_block_heads(const uint32_t* __restrict__ sbrow,
                                                        const uint32_t* __restrict__ sbcol, uint64_t nnz,
                                                        uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (i == 0 || sbrow[i] != sbrow[i - 1] || sbcol[i] != sbcol[i - 1]) ? 1u : 0u;
}

// After an exclusive scan of the head flags (bid = block index of each head):
// write bcol, and sum each run of equal (row, col) entries into its block slot.
__global__ __launch_bounds__(kBlock) void k_block_fill(const uint32_t* __restrict__ sbcol,
                                                       const uint32_t* __restrict__ flag,
                                                       const uint32_t* __restrict__ excl,
                                                       const float* __restrict__ spayload,
                                                       const uint32_t* __restrict__ csr_rows,
                                                       const uint32_t* __restrict__ col,
                                                       const float* __restrict__ val, uint64_t nnz, uint32_t B,
                                                       uint32_t* __restrict__ bcol, float* __restrict__ blocks) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t blk = excl[i] + flag[i] - 1u;  // block index of element i
    if (flag[i]) bcol[blk] = sbcol[i];
    const uint32_t p = __float_as_uint(spayload[i]);
    const uint32_t r = csr_rows[p], c = col[p];
    if (i > 0 && !flag[i]) {
      const uint32_t pp = __float_as_uint(spayload[i - 1]);
      if (csr_rows[pp] == r && col[pp] == c) continue;  // not the head of its (row, col) run
    }
    float s = val[p];
    for (uint64_t j = i + 1; j < nnz && !flag[j]; ++j) {
      const uint32_t q = __float_as_uint(spayload[j]);
      if (csr_rows[q] != r || col[q] != c) break;
      s += val[q];
    }
    blocks[(size_t)blk * B * B + (size_t)(r % B) * B + (c % B)] = s;
  }
}

// brow_ptr[br] = number of blocks before block row br = (heads before the first
// element of block row br) = excl[brp_nnz[br]] (or the total at the end).
__global__ __launch_bounds__(kBlock) void k_brow_ptr(const uint32_t* __restrict__ brp_nnz,
                                                     const uint32_t* __restrict__ excl,
                                                     const uint32_t* __restrict__ flag, uint32_t nbr,
                                                     uint64_t nnz, uint32_t* __restrict__ brow_ptr) {
  for (uint32_t br = blockIdx.x * blockDim.x + threadIdx.x; br <= nbr; br += gridDim.x * blockDim.x) {
    const uint32_t pos = brp_nnz[br];
    brow_ptr[br] = pos < nnz ? excl[pos] : (nnz ? excl[nnz - 1] + flag[nnz - 1] : 0u);
  }
}

}  // namespace msp
