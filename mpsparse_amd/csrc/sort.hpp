// COO -> CSR on the GPU: stable LSD radix sort of packed (row<<32 | col) keys
// carrying one value per key, then row_ptr from the sorted rows.
//
// Replaces the reference's 4-bit LSD pipeline (iterative/cgd/coo_csr.metal:17-193
// driven by iterative/BiCGSTAB/spmv.cpp:62-228). Same result (a stable sort:
// duplicates kept in input order), different machine:
//   * keys are first compacted to (row << cbits) | col with cbits = bitlen(num_cols-1),
//     so only rbits + cbits significant bits are sorted (u32 keys when they fit);
//   * 8-bit digits, 3072-key tiles; per pass: upsweep (tile digit counts in LDS),
//     a per-digit scan of the tile-major count table (three kernels), and a
//     downsweep that ranks keys stably with 64-lane ballots, stages the tile in
//     LDS in digit order and writes runs of equal digits contiguously;
//   * the last pass writes col_ind / values / row directly (no key array), and
//     row_ptr comes from a parallel boundary fill (no serial row fixer).
#pragma once

#include "common.hpp"

namespace msp {
namespace sortk {

constexpr int kItems = 12;
constexpr int kTile = kBlock * kItems;  // 3072 keys per tile (measured best of 2048 / 3072 / 4096 / 6144)
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096 counts per scan block

struct KeyMap {
  uint32_t cbits;
  uint32_t rows, cols;
  uint64_t cmask;
};

__device__ __forceinline__ uint64_t compact_key(uint64_t key, const KeyMap& km, int* err) {
  uint64_t r = key >> 32, c = key & 0xFFFFFFFFull;
  if (r >= km.rows || c >= km.cols) {
    if (err) atomicOr(err, 1);
    r = 0;
    c = 0;
  }
  return (r << km.cbits) | c;
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
  return (uint32_t)(k >> shift) & (kRadix - 1);
}

// Tile digit histogram -> counts[tile * kRadix + d] (tile-major: one
// coalesced 1 KiB row per tile).
template <typename InK, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_upsweep(const InK* __restrict__ keys, uint64_t nnz, KeyMap km,
                                                    int shift, uint32_t num_tiles, uint32_t* __restrict__ counts,
                                                    int* err) {
  __shared__ uint32_t h[kWavesPerBlock][kRadix];
  const int tid = threadIdx.x, wid = tid / kWave;
  for (int i = tid; i < kWavesPerBlock * kRadix; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + tid;
    if (idx < nnz) {
      uint64_t k = (uint64_t)keys[idx];
      if (FIRST) k = compact_key(k, km, err);
      atomicAdd(&h[wid][digit_of(k, shift)], 1u);
    }
  }
  __syncthreads();
  for (int d = tid; d < kRadix; d += kBlock) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) s += h[w][d];
    counts[(uint64_t)blockIdx.x * kRadix + d] = s;
  }
}

// Scatter offsets from the tile-major count table: offset(t, d) = start of
// digit d + sum of counts[t'][d] over tiles t' < t. Three kernels, all with
// one thread per digit and coalesced rows: per-chunk column sums, one block
// that scans them per digit and adds the digit starts, and the per-chunk
// running sums written back in place.
constexpr int kColChunk = 256;  // tiles per column-scan block

__global__ __launch_bounds__(kRadix) void k_colscan_reduce(const uint32_t* __restrict__ counts, uint32_t num_tiles,
                                                           uint32_t* __restrict__ sums) {
  const uint32_t d = threadIdx.x, t0 = blockIdx.x * kColChunk;
  const uint32_t t1 = t0 + kColChunk < num_tiles ? t0 + kColChunk : num_tiles;
  uint32_t s = 0;
#pragma unroll 16
  for (uint32_t t = t0; t < t1; ++t) s += counts[(uint64_t)t * kRadix + d];
  sums[(uint64_t)blockIdx.x * kRadix + d] = s;
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl);

__global__ __launch_bounds__(kRadix) void k_colscan_top(uint32_t* __restrict__ sums, uint32_t nchunks) {
  const uint32_t d = threadIdx.x;
  uint32_t run = 0;
  for (uint32_t c0 = 0; c0 < nchunks; c0 += 16) {  // 16 loads in flight per pass
    uint32_t v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = c0 + q < nchunks ? sums[(uint64_t)(c0 + q) * kRadix + d] : 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (c0 + q < nchunks) sums[(uint64_t)(c0 + q) * kRadix + d] = run;
      run += v[q];
    }
  }
  uint32_t start;
  block_exclusive_scan(run, start);  // digit starts: totals scanned over the digits
  for (uint32_t c = 0; c < nchunks; ++c) sums[(uint64_t)c * kRadix + d] += start;
}

__global__ __launch_bounds__(kRadix) void k_colscan_down(uint32_t* __restrict__ counts, uint32_t num_tiles,
                                                         const uint32_t* __restrict__ sums) {
  const uint32_t d = threadIdx.x, t0 = blockIdx.x * kColChunk;
  const uint32_t t1 = t0 + kColChunk < num_tiles ? t0 + kColChunk : num_tiles;
  uint32_t run = sums[(uint64_t)blockIdx.x * kRadix + d];
  for (uint32_t tb = t0; tb < t1; tb += 16) {
    uint32_t v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = tb + q < t1 ? counts[(uint64_t)(tb + q) * kRadix + d] : 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (tb + q < t1) counts[(uint64_t)(tb + q) * kRadix + d] = run;
      run += v[q];
    }
  }
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl) {
  __shared__ uint32_t wsum[kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  uint32_t wbase = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) {
    uint32_t s = wsum[w];
    if (w < wid) wbase += s;
    total += s;
  }
  excl = wbase + incl - v;
  __syncthreads();
  return total;
}

// Scan phase A: per-block sums of kScanTile consecutive counts.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t N,
                                                        uint32_t* __restrict__ block_sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < N) s += in[idx];
  }
  uint32_t excl;
  uint32_t total = block_exclusive_scan(s, excl);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Scan phase B: one block scans all block sums (exclusive, in place).
__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nb) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kBlock) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t excl;
    uint32_t total = block_exclusive_scan(v, excl);
    if (i < nb) sums[i] = carry + excl;
    carry += total;
  }
}

// Scan phase C: exclusive scan inside each block plus its block offset (in
// place). The block's counts pass through LDS so that both the load and the
// store are coalesced (thread-contiguous runs of kScanItems in HBM were 4x
// slower).
__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ data, uint64_t N,
                                                      const uint32_t* __restrict__ block_off) {
  __shared__ uint32_t buf[kScanTile + kScanTile / 32];  // +1 word per 32 against bank conflicts
  auto at = [](uint32_t k) { return k + (k >> 5); };
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t k = (uint32_t)i * kBlock + threadIdx.x;
    buf[at(k)] = base + k < N ? data[base + k] : 0u;
  }
  __syncthreads();
  const uint32_t k0 = threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = buf[at(k0 + i)];
    s += v[i];
  }
  uint32_t excl;
  block_exclusive_scan(s, excl);
  uint32_t run = block_off[blockIdx.x] + excl;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    buf[at(k0 + i)] = run;
    run += v[i];
  }
  __syncthreads();
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t k = (uint32_t)i * kBlock + threadIdx.x;
    if (base + k < N) data[base + k] = buf[at(k)];
  }
}

// Stable rank + scatter of one tile. Wave w owns tile positions
// [w*64*kItems, (w+1)*64*kItems), item i of lane l at w*64*kItems + i*64 + l,
// so the (wave, item, lane) order is the input order.
template <typename InK, typename OutK, bool FIRST, bool LAST>
__global__ __launch_bounds__(kBlock) void k_downsweep(const InK* __restrict__ kin, const float* __restrict__ vin,
                                                      OutK* __restrict__ kout, float* __restrict__ vout,
                                                      uint32_t* __restrict__ col_out, uint32_t* __restrict__ row_out,
                                                      uint64_t nnz, KeyMap km, int shift, uint32_t num_tiles,
                                                      const uint32_t* __restrict__ offsets) {
  __shared__ uint32_t wcnt[kWavesPerBlock][kRadix];
  __shared__ uint32_t dstart[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint64_t stage[kTile];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  // consecutive tiles on the same XCD: the runs that neighbouring tiles write
  // for one digit are adjacent in the output and then meet in one L2
  const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint64_t tile_base = (uint64_t)tile * kTile;
  const uint32_t tile_n = (uint32_t)((nnz - tile_base) < (uint64_t)kTile ? (nnz - tile_base) : (uint64_t)kTile);

  for (int i = tid; i < kWavesPerBlock * kRadix; i += kBlock) (&wcnt[0][0])[i] = 0;
  for (int d = tid; d < kRadix; d += kBlock) gbase[d] = offsets[(uint64_t)tile * kRadix + d];
  __syncthreads();

  OutK key[kItems];
  float val[kItems];
  uint32_t rank[kItems];
  const uint64_t lane_lt = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
  const uint32_t wave_base = (uint32_t)wid * kWave * kItems;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint32_t pos = wave_base + (uint32_t)i * kWave + lane;
    bool valid = pos < tile_n;
    uint64_t k = 0;
    float v = 0.0f;
    if (valid) {
      k = (uint64_t)kin[tile_base + pos];
      if (FIRST) k = compact_key(k, km, nullptr);
      v = vin[tile_base + pos];
    }
    key[i] = (OutK)k;
    val[i] = v;
    uint32_t d = digit_of(k, shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t below = (uint32_t)__popcll(peers & lane_lt);
    uint32_t before = valid ? wcnt[wid][d] : 0u;
    rank[i] = before + below;
    // highest peer lane publishes the running count for this digit
    if (valid && (peers >> lane) == 1ull) wcnt[wid][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per-digit exclusive scan over waves, then over digits (tile-local starts)
  uint32_t tot = 0;
  {
    const int d = tid;  // kBlock == kRadix
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {
      uint32_t c = wcnt[w][d];
      wcnt[w][d] = run;
      run += c;
    }
    tot = run;
  }
  __syncthreads();
  {
    uint32_t excl;
    block_exclusive_scan(tot, excl);
    dstart[tid] = excl;
  }
  __syncthreads();
  uint32_t lpos[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint32_t pos = wave_base + (uint32_t)i * kWave + lane;
    if (pos < tile_n) {
      uint32_t d = digit_of(key[i], shift);
      lpos[i] = dstart[d] + wcnt[wid][d] + rank[i];
      stage[lpos[i]] = (uint64_t)key[i];
    }
  }
  __syncthreads();
  // keys (or row/col) out in tile-sorted order: consecutive threads hit runs
  // of consecutive global positions
  for (uint32_t j = tid; j < tile_n; j += kBlock) {
    uint64_t k = stage[j];
    uint32_t d = digit_of(k, shift);
    uint32_t out = gbase[d] + (j - dstart[d]);
    if (LAST) {
      col_out[out] = (uint32_t)(k & km.cmask);
      row_out[out] = (uint32_t)(k >> km.cbits);
    } else {
      kout[out] = (OutK)k;
    }
    stage[j] = ((uint64_t)out << 32);  // remember destination for the value pass
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint32_t pos = wave_base + (uint32_t)i * kWave + lane;
    if (pos < tile_n) stage[lpos[i]] |= (uint64_t)__float_as_uint(val[i]);
  }
  __syncthreads();
  for (uint32_t j = tid; j < tile_n; j += kBlock) {
    uint64_t e = stage[j];
    vout[(uint32_t)(e >> 32)] = __uint_as_float((uint32_t)e);
  }
}

// row_ptr from sorted rows: row_ptr[q] = first position whose row >= q.
// Thread for position i (1 <= i < nnz) fills the gap (rows[i-1], rows[i]];
// the head (q <= rows[0]) and tail (q > rows[nnz-1]) are filled grid-stride.
__global__ __launch_bounds__(kBlock) void k_row_ptr(const uint32_t* __restrict__ rows, uint64_t nnz, uint32_t n,
                                                    uint32_t* __restrict__ row_ptr, const int* __restrict__ err) {
  if (*err) return;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nnz == 0) {
    for (uint64_t q = gid; q <= n; q += stride) row_ptr[q] = 0;
    return;
  }
  const uint32_t first = rows[0], last = rows[nnz - 1];
  for (uint64_t q = gid; q <= n; q += stride) {
    if (q <= first) row_ptr[q] = 0;
    else if (q > last) row_ptr[q] = (uint32_t)nnz;
  }
  for (uint64_t i = gid + 1; i < nnz; i += stride) {
    uint32_t r = rows[i], rp = rows[i - 1];
    for (uint32_t q = rp + 1; q <= r; ++q) row_ptr[q] = (uint32_t)i;
  }
}

}  // namespace sortk
}  // namespace msp
