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

constexpr int kItems = 10;
// 2560 keys per tile: the downsweep's LDS (25.6 KB with 16-bit digit counts)
// then fits 6 workgroups per CU (3072 measured best of 2048 / 3072 / 4096 /
// 6144 with 32-bit counts, at 5 per CU)
constexpr int kTile = kBlock * kItems;
static_assert(kTile <= 65535, "16-bit digit counts");
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

// The valid lanes of the wave whose digit equals this lane's: one ballot per
// digit bit, each XORed with this lane's bit spread over a word (0 or ~0);
// any set bit of the OR of those is a lane that differs somewhere. Kept in
// 32-bit halves so that each step is one three-input bitwise operation.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, b, 1);  // bit b as 0 / ~0
    const uint64_t bb = __builtin_amdgcn_uicmp(m, 0u, 33);  // lanes with m != 0 (ICMP_NE)
    lo |= (uint32_t)bb ^ m;
    hi |= (uint32_t)(bb >> 32) ^ m;
  }
  const uint64_t v = __ballot(valid);
  return v & ~(((uint64_t)hi << 32) | lo);
}

// Tile digit histogram -> counts[tile * kRadix + d] (tile-major: one
// coalesced 1 KiB row per tile).
template <typename InK, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_upsweep(const InK* __restrict__ keys, uint64_t nnz, KeyMap km,
                                                    int shift, uint32_t num_tiles, uint32_t* __restrict__ counts,
                                                    int* err) {
  __shared__ uint32_t h[kWavesPerBlock][kRadix];
  const int tid = threadIdx.x, wid = tid / kWave;
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  // every key load of the tile issued before the first count, and before the
  // counts are cleared (the clearing and its barrier overlap the loads)
  InK kv[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + (uint64_t)i * kBlock + tid;
    kv[i] = idx < nnz ? keys[idx] : InK(0);
  }
  for (int i = tid; i < kWavesPerBlock * kRadix; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + (uint64_t)i * kBlock + tid;
    if (idx < nnz) {
      uint64_t k = (uint64_t)kv[i];
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

// sums[c][d] <- sum of the chunk sums of digit d before chunk c; the digit
// starts (totals scanned over the digits) go to dstart[0..255] and nnz to
// dstart[256], for k_colscan_down to add and for passes that need them.
__global__ __launch_bounds__(kRadix) void k_colscan_top(uint32_t* __restrict__ sums, uint32_t nchunks,
                                                        uint32_t* __restrict__ dstart, uint64_t nnz) {
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
  dstart[d] = start;
  if (d == 0) dstart[kRadix] = (uint32_t)nnz;
}

__global__ __launch_bounds__(kRadix) void k_colscan_down(uint32_t* __restrict__ counts, uint32_t num_tiles,
                                                         const uint32_t* __restrict__ sums,
                                                         const uint32_t* __restrict__ dstart) {
  const uint32_t d = threadIdx.x, t0 = blockIdx.x * kColChunk;
  const uint32_t t1 = t0 + kColChunk < num_tiles ? t0 + kColChunk : num_tiles;
  uint32_t run = sums[(uint64_t)blockIdx.x * kRadix + d] + dstart[d];
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

// Key transforms of a downsweep (the two-level sort's phase 1 with two
// passes over keys of 33..40 bits, so that its keys travel as 32 bits):
//   kXfNone: the key is written as read;
//   kXfDrop: the pass's own digit (bits [shift, shift + 8)) is removed, the
//            bits above it move down by 8 -- the digit is recoverable from
//            the output position through the pass's digit starts;
//   kXfSwap: the digit at [shift, shift + 8) (the pass's own) is replaced by
//            the previous pass's digit, recovered from the input position
//            through that pass's digit starts `prev` (257 entries).
enum : int { kXfNone = 0, kXfDrop = 1, kXfSwap = 2 };

__device__ __forceinline__ uint32_t digit_at(const uint32_t* __restrict__ starts, uint32_t lo, uint32_t hi,
                                             uint64_t p) {
  // the largest d in [lo, hi] with starts[d] <= p
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if ((uint64_t)starts[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Stable rank + scatter of one tile. Wave w owns tile positions
// [w*64*kItems, (w+1)*64*kItems), item i of lane l at w*64*kItems + i*64 + l,
// so the (wave, item, lane) order is the input order. Compiled for 6 waves
// per SIMD (the LDS allows 6 workgroups per CU; 151M keys, the pass that
// drops its digit: 1027 -> 854 us), the digit-swapping pass for 4 (its search
// of the previous digit's range spills at 6 and 5; 1000 us at 4 or 5).
template <typename InK, typename OutK, bool FIRST, bool LAST, int XF = kXfNone>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(XF == kXfSwap ? 4 : 6))) void k_downsweep(
                                                      const InK* __restrict__ kin, const float* __restrict__ vin,
                                                      OutK* __restrict__ kout, float* __restrict__ vout,
                                                      uint32_t* __restrict__ col_out, uint32_t* __restrict__ row_out,
                                                      uint64_t nnz, KeyMap km, int shift, uint32_t num_tiles,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ prev = nullptr) {
  static_assert(XF == kXfNone || (sizeof(OutK) == 4 && !LAST), "key transforms write 32-bit keys");
  __shared__ uint16_t wcnt[kWavesPerBlock][kRadix];  // (at most kTile per digit)
  __shared__ uint32_t dstart[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint64_t stage[kTile];
  [[maybe_unused]] __shared__ uint32_t pstart[kRadix + 1];
  [[maybe_unused]] __shared__ uint32_t prange[2];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  // consecutive tiles on the same XCD: the runs that neighbouring tiles write
  // for one digit are adjacent in the output and then meet in one L2
  const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint64_t tile_base = (uint64_t)tile * kTile;
  const uint32_t tile_n = (uint32_t)((nnz - tile_base) < (uint64_t)kTile ? (nnz - tile_base) : (uint64_t)kTile);

  for (int i = tid; i < kWavesPerBlock * kRadix; i += kBlock) (&wcnt[0][0])[i] = 0;
  for (int d = tid; d < kRadix; d += kBlock) gbase[d] = offsets[(uint64_t)tile * kRadix + d];
  if constexpr (XF == kXfSwap) {
    for (int d = tid; d <= kRadix; d += kBlock) pstart[d] = prev[d];
    __syncthreads();
    // the previous digits of the tile's first and last positions bound every
    // search below (a tile usually lies within one or two digits)
    if (tid < 2) prange[tid] = digit_at(pstart, 0, kRadix - 1, tile_base + (tid ? tile_n - 1 : 0));
  }
  __syncthreads();

  // with a transform, key[i] holds (digit << 32) | transformed key
  using HeldK = typename std::conditional<XF == kXfNone, OutK, uint64_t>::type;
  HeldK key[kItems];
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
    val[i] = v;
    uint32_t d = digit_of(k, shift);
    if constexpr (XF == kXfNone) {
      key[i] = (OutK)k;
    } else {
      const uint64_t low = k & ((1ull << shift) - 1ull);
      uint64_t t;
      if constexpr (XF == kXfDrop) {
        t = ((k >> (shift + kRadixBits)) << shift) | low;
      } else {
        const uint64_t pd = valid ? digit_at(pstart, prange[0], prange[1], tile_base + pos) : 0u;
        t = ((k >> (shift + kRadixBits)) << (shift + kRadixBits)) | (pd << shift) | low;
      }
      key[i] = ((uint64_t)d << 32) | (uint32_t)t;
    }
    const uint64_t peers = match_digit(d, valid);
    uint32_t below = (uint32_t)__popcll(peers & lane_lt);
    uint32_t before = valid ? wcnt[wid][d] : 0u;
    rank[i] = before + below;
    // highest peer lane publishes the running count for this digit
    if (valid && (peers >> lane) == 1ull) wcnt[wid][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
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
      wcnt[w][d] = (uint16_t)run;
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
      uint32_t d = XF == kXfNone ? digit_of(key[i], shift) : (uint32_t)((uint64_t)key[i] >> 32);
      lpos[i] = dstart[d] + wcnt[wid][d] + rank[i];
      stage[lpos[i]] = (uint64_t)key[i];
    }
  }
  __syncthreads();
  // keys (or row/col) out in tile-sorted order: consecutive threads hit runs
  // of consecutive global positions
  for (uint32_t j = tid; j < tile_n; j += kBlock) {
    uint64_t k = stage[j];
    uint32_t d = XF == kXfNone ? digit_of(k, shift) : (uint32_t)(k >> 32);
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

// ---------------------------------------------------------------------------
// Two-level COO -> CSR. The rows are grouped into segments of 2^s consecutive
// rows (s chosen so that a segment holds a few thousand keys). Phase 1 sorts
// the keys stably by segment only (LSD passes over the segment bits, the
// kernels above); k_seg_bounds finds each segment's range in that order; and
// k_segment_sort sorts every segment completely in LDS, one workgroup per
// segment, writing col_ind, the values, row_ptr of its rows (and the rows when
// asked) directly. Inside a segment each key carries its position within the
// segment in its low kSegIdxBits bits: stable LSD passes over the local bits
// (row within the segment, column) then leave equal keys in input order, and
// the value is fetched once, at the end, through that position. A segment
// larger than kSegMax keys raises the overflow flag (the host then runs the
// single-level sort instead).
constexpr int kSegThreads = 512;             // workgroup of the segment sort (8 waves)
constexpr int kSegWaves = kSegThreads / kWave;
constexpr int kSegItems = 8;
constexpr int kSegMax = kSegThreads * kSegItems;  // 4096 keys per segment at most
constexpr int kSegIdxBits = 12;                   // 2^12 >= kSegMax

// seg_start[g] = first position whose segment id (key >> seg_shift) >= g, for
// g in [0, G]; with `whole` one segment spans everything.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_seg_bounds(const K* __restrict__ keys, uint64_t nnz, int seg_shift,
                                                       uint32_t G, bool whole, uint32_t* __restrict__ seg_start) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += (uint64_t)gridDim.x * blockDim.x) {
    if (whole) {
      seg_start[g] = g == 0 ? 0u : (uint32_t)nnz;
      continue;
    }
    uint64_t lo = 0, hi = nnz;
    while (lo < hi) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (((uint64_t)keys[mid] >> seg_shift) < g) lo = mid + 1;
      else hi = mid;
    }
    seg_start[g] = (uint32_t)lo;
  }
}

// The same bounds after phase-1 passes with key transforms: the keys are
// grouped by the last pass's digit (the segment id's high digit), whose
// ranges are hi_starts[0..256], and within a range ordered by the low digit
// held at bits [shift, shift + 8) of the key. Segment g = (hi << 8) | lo.
__global__ __launch_bounds__(kBlock) void k_seg_bounds_split(const uint32_t* __restrict__ keys, uint64_t nnz,
                                                             int shift, uint32_t G,
                                                             const uint32_t* __restrict__ hi_starts,
                                                             uint32_t* __restrict__ seg_start) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += (uint64_t)gridDim.x * blockDim.x) {
    if (g == G) {
      seg_start[g] = (uint32_t)nnz;
      continue;
    }
    const uint32_t hi = (uint32_t)(g >> kRadixBits), lo_d = (uint32_t)(g & (kRadix - 1));
    uint64_t lo = hi_starts[hi], top = hi_starts[hi + 1];
    while (lo < top) {
      const uint64_t mid = lo + ((top - lo) >> 1);
      if (digit_of(keys[mid], shift) < lo_d) lo = mid + 1;
      else top = mid;
    }
    seg_start[g] = (uint32_t)lo;
  }
}

// One workgroup per segment (grid-stride over segments). FIRST: the input is
// the caller's raw (row << 32 | col) keys (no phase-1 pass ran: one segment),
// compacted and range-checked here. A segment of cnt keys is spread over
// ieff = ceil(cnt / kSegThreads) items per thread; element order inside the
// workgroup is (wave, item, lane) -- wave w holds positions
// [w * 64 * ieff, (w + 1) * 64 * ieff) -- so the ballot ranks are stable. The
// values stay in registers during the passes and are placed in LDS (over the
// keys' buffer) only once the final order is known.
template <typename K, bool FIRST>
__global__ __launch_bounds__(kSegThreads, 8) void k_segment_sort(const K* __restrict__ keys,
                                                              const float* __restrict__ vals,
                                                              const uint32_t* __restrict__ seg_start, uint32_t G,
                                                              KeyMap km, int s, uint32_t n, uint64_t nnz,
                                                              uint32_t* __restrict__ col_out,
                                                              float* __restrict__ val_out,
                                                              uint32_t* __restrict__ rows_out,
                                                              uint32_t* __restrict__ row_ptr,
                                                              int* __restrict__ overflow, int* __restrict__ err) {
  __shared__ uint64_t skey[kSegMax];
  // per-wave digit counts in 16 bits (at most kSegMax keys): 37 KB of LDS a
  // workgroup and at most 64 registers a lane, so 4 segments per CU instead of 3
  __shared__ uint16_t wcnt[kSegWaves][kRadix];
  static_assert(kSegMax <= 65535, "16-bit counts");
  __shared__ uint32_t dstart[kRadix];
  float* sval = reinterpret_cast<float*>(skey);  // after the last pass
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  const uint64_t lane_lt = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
  const uint32_t lbits = km.cbits + (uint32_t)s;  // local key bits: row within the segment, column
  const uint64_t lmask = lbits >= 64 ? ~0ull : ((1ull << lbits) - 1ull);
  const int npass = (int)((lbits + kRadixBits - 1) / kRadixBits);
  if (blockIdx.x == 0 && tid == 0) row_ptr[n] = (uint32_t)nnz;
  for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
    const uint32_t start = seg_start[g], end = seg_start[g + 1], cnt = end - start;
    const uint32_t row0 = g << s;
    const uint32_t nrow = (n - row0) < (1u << s) ? (n - row0) : (1u << s);
    if (cnt > (uint32_t)kSegMax) {
      if (tid == 0) atomicOr(overflow, 1);
      continue;
    }
    const uint32_t ieff = (cnt + kSegThreads - 1) / kSegThreads;
    const uint32_t wave_base = (uint32_t)wid * kWave * ieff;
    uint64_t key[kSegItems];
    float val[kSegItems];
#pragma unroll
    for (int i = 0; i < kSegItems; ++i) {
      key[i] = 0;
      val[i] = 0.0f;
      if ((uint32_t)i < ieff) {
        const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
        uint64_t k = 0;
        if (j < cnt) {
          k = (uint64_t)keys[start + j];
          val[i] = vals[start + j];
          if (FIRST) k = compact_key(k, km, err);
        }
        key[i] = ((k & lmask) << kSegIdxBits) | j;
      }
    }
    for (int p = 0; p < npass; ++p) {
      const int shift = kSegIdxBits + p * kRadixBits;
      for (int i = tid; i < kSegWaves * kRadix; i += kSegThreads) (&wcnt[0][0])[i] = 0;
      __syncthreads();
      uint32_t rank[kSegItems];
#pragma unroll
      for (int i = 0; i < kSegItems; ++i) {
        rank[i] = 0;
        if ((uint32_t)i < ieff) {
          const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
          const bool valid = j < cnt;
          const uint32_t d = digit_of(key[i], shift);
          const uint64_t peers = match_digit(d, valid);
          const uint32_t below = (uint32_t)__popcll(peers & lane_lt);
          const uint32_t before = valid ? wcnt[wid][d] : 0u;
          rank[i] = before + below;
          if (valid && (peers >> lane) == 1ull) wcnt[wid][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
          __builtin_amdgcn_wave_barrier();
        }
      }
      __syncthreads();
      // per digit: exclusive scan over the waves (in place), digit totals
      // scanned by wave 0 (4 digits per lane) into the tile-local starts
      uint32_t tot = 0;
      if (tid < kRadix) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kSegWaves; ++w) {
          const uint32_t c = wcnt[w][tid];
          wcnt[w][tid] = (uint16_t)run;
          run += c;
        }
        tot = run;
        dstart[tid] = tot;
      }
      __syncthreads();
      if (wid == 0) {
        const uint32_t t0 = dstart[4 * lane], t1 = dstart[4 * lane + 1], t2 = dstart[4 * lane + 2],
                       t3 = dstart[4 * lane + 3];
        const uint32_t sum = t0 + t1 + t2 + t3;
        uint32_t incl = sum;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          const uint32_t t = __shfl_up(incl, off, kWave);
          if (lane >= off) incl += t;
        }
        const uint32_t base = incl - sum;
        dstart[4 * lane] = base;
        dstart[4 * lane + 1] = base + t0;
        dstart[4 * lane + 2] = base + t0 + t1;
        dstart[4 * lane + 3] = base + t0 + t1 + t2;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kSegItems; ++i) {
        if ((uint32_t)i < ieff) {
          const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
          if (j < cnt) {
            const uint32_t d = digit_of(key[i], shift);
            skey[dstart[d] + wcnt[wid][d] + rank[i]] = key[i];
          }
        }
      }
      __syncthreads();
      if (p + 1 < npass) {
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
          if ((uint32_t)i < ieff) {
            const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
            if (j < cnt) key[i] = skey[j];
          }
        }
        __syncthreads();
      }
    }
    if (npass == 0) {  // no local bits: the input order is the order
#pragma unroll
      for (int i = 0; i < kSegItems; ++i) {
        if ((uint32_t)i < ieff) {
          const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
          if (j < cnt) skey[j] = key[i];
        }
      }
      __syncthreads();
    }
    // the sorted keys of this thread's output positions q = tid + k * kSegThreads
    // (and the previous key's local row, + 1), then the values over the key
    // buffer at their input positions
    uint64_t out[kSegItems];
    uint32_t prev[kSegItems];
#pragma unroll
    for (int k = 0; k < kSegItems; ++k) {
      out[k] = 0;
      prev[k] = 0;
      const uint32_t q = (uint32_t)tid + (uint32_t)k * kSegThreads;
      if ((uint32_t)k < ieff && q < cnt) {
        out[k] = skey[q];
        prev[k] = q == 0 ? 0u : (uint32_t)((skey[q - 1] >> kSegIdxBits) >> km.cbits) + 1u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSegItems; ++i) {
      if ((uint32_t)i < ieff) {
        const uint32_t j = wave_base + (uint32_t)i * kWave + lane;
        if (j < cnt) sval[j] = val[i];  // val[i] was loaded from input position j
      }
    }
    __syncthreads();
    // outputs (coalesced); row_ptr: rows [prev, r] start at the first element
    // of row r, rows after the last element start at the segment's end
#pragma unroll
    for (int k = 0; k < kSegItems; ++k) {
      const uint32_t q = (uint32_t)tid + (uint32_t)k * kSegThreads;
      if ((uint32_t)k < ieff && q < cnt) {
        const uint64_t e = out[k];
        const uint64_t lk = e >> kSegIdxBits;
        const uint32_t lr = (uint32_t)(lk >> km.cbits);
        col_out[start + q] = (uint32_t)(lk & km.cmask);
        val_out[start + q] = sval[(uint32_t)(e & ((1u << kSegIdxBits) - 1u))];
        if (rows_out) rows_out[start + q] = row0 + lr;
        for (uint32_t r = prev[k]; r <= lr; ++r) row_ptr[row0 + r] = start + q;
        if (q == cnt - 1)
          for (uint32_t r = lr + 1; r < nrow; ++r) row_ptr[row0 + r] = end;
      }
    }
    if (cnt == 0)
      for (uint32_t r = tid; r < nrow; r += kSegThreads) row_ptr[row0 + r] = start;
    __syncthreads();  // skey is reused by the next segment
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
