// Device-side building blocks shared by every kernel of the library:
// vector types, wave/block reductions with compensated FP32 sums, the
// deterministic one-pass grid reduction (per-block partials + last-block
// finisher), XCD-aware block numbering and value-type loaders.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace msp {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));

constexpr int kWave = 64;
constexpr int kBlock = 256;           // threads per workgroup for streaming kernels
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNumXcd = 8;            // MI355X: 8 XCDs, workgroups dispatched round-robin

// ---------------------------------------------------------------------------
// Compensated (TwoSum) FP32 accumulation: s + c carries ~2x the precision of a
// plain float sum. The error-free transformations below must not be contracted
// into FMAs with a product computed by the caller (hipcc contracts across
// inlined statements by default), hence `fp contract(off)` in each body.
struct CSum {
  float s, c;
};

__device__ __forceinline__ CSum csum_zero() { return CSum{0.0f, 0.0f}; }

__device__ __forceinline__ CSum csum_add(CSum x, float v) {
#pragma clang fp contract(off)
  float s = x.s + v;
  float bp = s - x.s;
  float e = (x.s - (s - bp)) + (v - bp);
  return CSum{s, x.c + e};
}

__device__ __forceinline__ CSum csum_merge(CSum x, CSum y) {
#pragma clang fp contract(off)
  float s = x.s + y.s;
  float bp = s - x.s;
  float e = (x.s - (s - bp)) + (y.s - bp);
  return CSum{s, (x.c + y.c) + e};
}

__device__ __forceinline__ float csum_val(CSum x) { return x.s + x.c; }

// (x, xl) += a * p with the rounding errors of the product (FMA) and of the sum
// (TwoSum) carried in xl: x + xl holds the iterate to about twice FP32 precision.
__device__ __forceinline__ void comp_axpy(float& x, float& xl, float a, float p) {
#pragma clang fp contract(off)
  const float t = a * p;
  const float te = __builtin_fmaf(a, p, -t);
  const float s = x + t;
  const float bp = s - x;
  const float e = (x - (s - bp)) + (t - bp);
  x = s;
  xl += e + te;
}

// DPP lane move (v_mov_dpp): lanes of rows outside ROW_MASK get 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF,
                                                               false));
}
constexpr int kDppXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // lane i <- 7 - i within 8 lanes
constexpr int kDppMirror = 0x140;      // lane i <- 15 - i within 16 lanes
constexpr int kDppBcast15 = 0x142;     // row r <- lane 15 of row r - 1
constexpr int kDppBcast31 = 0x143;     // rows 2, 3 <- lane 31

// Wave-wide sums by DPP lane moves (VALU; the ds_bpermute shuffles they
// replaced each took an LDS round trip): after the quad, half-row and row
// steps every lane holds its 16-lane row's sum; row_bcast:15 adds row 0 into
// row 1 and row 2 into row 3, row_bcast:31 adds rows 0-1 into row 3, whose
// lane 63 then holds the wave's total, read out to every lane. The tree is
// fixed and csum_merge is symmetric, so every wave (and every workgroup)
// gets the same bits for the same inputs.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ CSum csum_dpp(CSum v) {
  return csum_merge(v, CSum{dpp_mov<CTRL, ROW_MASK>(v.s), dpp_mov<CTRL, ROW_MASK>(v.c)});
}
__device__ __forceinline__ CSum wave_reduce(CSum v) {
  v = csum_dpp<kDppXor1>(v);
  v = csum_dpp<kDppXor2>(v);
  v = csum_dpp<kDppHalfMirror>(v);
  v = csum_dpp<kDppMirror>(v);
  v = csum_dpp<kDppBcast15, 0xA>(v);
  v = csum_dpp<kDppBcast31, 0xC>(v);
  return CSum{__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.s), kWave - 1)),
              __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.c), kWave - 1))};
}

__device__ __forceinline__ float wave_reduce_sum(float v) {
  v += dpp_mov<kDppXor1>(v);
  v += dpp_mov<kDppXor2>(v);
  v += dpp_mov<kDppHalfMirror>(v);
  v += dpp_mov<kDppMirror>(v);
  v += dpp_mov<kDppBcast15, 0xA>(v);
  v += dpp_mov<kDppBcast31, 0xC>(v);
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), kWave - 1));
}

// Sum over groups of L consecutive lanes (L a power of two <= 64); every lane
// of the wave must call it. Result valid in every lane of the group.
__device__ __forceinline__ float group_reduce_sum(float v, uint32_t L) {
  for (uint32_t off = L >> 1; off > 0; off >>= 1) v += __shfl_xor(v, (int)off, kWave);
  return v;
}

// Block-wide reduction of ND compensated sums; result valid in thread 0.
template <int ND, int NW = kWavesPerBlock>
__device__ __forceinline__ void block_reduce(CSum (&acc)[ND]) {
  __shared__ CSum red[NW][ND > 0 ? ND : 1];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = wave_reduce(acc[d]);
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) red[wid][d] = acc[d];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      CSum t = red[0][d];
      for (int w = 1; w < NW; ++w) t = csum_merge(t, red[w][d]);
      acc[d] = t;
    }
  }
  __syncthreads();
}

// Scratch of a grid reduction: one CSum per (block, value).
struct GridRed {
  CSum* partials;  // [gridDim.x][ND] of the producing kernel
};

// Deterministic grid-wide reduction in two kernels. The producing kernel ends
// with block_partials (plain stores of each block's sums, no fences or
// atomics); the host then launches k_finish on the same stream, whose single
// workgroup sums the partials in block order and hands the totals to the
// finisher. The kernel boundary makes the partials visible. (The one-kernel
// last-block-done variant needs a device-scope fence and a same-address atomic
// in every workgroup; on this chip that costs ~50 ns per workgroup, 100 us for
// a 2048-workgroup launch.)
template <int ND>
__device__ __forceinline__ void block_partials(CSum (&acc)[ND], GridRed rb) {
  block_reduce<ND>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) rb.partials[blockIdx.x * ND + d] = acc[d];
  }
}

// This thread's share of nb block partials (ND consecutive values at a
// stride of `stride` CSums per block), merged into a[] in block order i = tid,
// tid + kBlock, ...: the loads of a batch are all issued before its first
// merge (a merge waiting on each load in turn would pay the memory latency
// once per partial).
constexpr int kPartialBatch = 8;
template <int ND>
__device__ __forceinline__ void merge_partials(const CSum* __restrict__ p, uint32_t nb, CSum (&a)[ND],
                                               uint32_t stride = ND) {
  for (uint32_t base = 0; base < nb; base += kPartialBatch * kBlock) {
    CSum v[kPartialBatch][ND];
#pragma unroll
    for (int j = 0; j < kPartialBatch; ++j) {
      const uint32_t i = base + j * kBlock + threadIdx.x;
#pragma unroll
      for (int d = 0; d < ND; ++d) v[j][d] = i < nb ? p[i * stride + d] : csum_zero();
    }
#pragma unroll
    for (int j = 0; j < kPartialBatch; ++j) {
      const bool in = base + j * kBlock + threadIdx.x < nb;
#pragma unroll
      for (int d = 0; d < ND; ++d) a[d] = in ? csum_merge(a[d], v[j][d]) : a[d];
    }
  }
}

// Finisher protocol: `active()` repeats the producer's skip test (a producer
// that returned early wrote no partials), `operator()(tot)` then `after()` run
// in thread 0.
// Finishers that own a solver state (a member `S`) run on an LDS copy of it:
// its words are loaded with the partials, so the two latencies overlap, and
// stored back after the bookkeeping (the state's own loads used to wait until
// after the reduction).
template <typename Fin, typename = void>
struct FinHasState : std::false_type {};
template <typename Fin>
struct FinHasState<Fin, std::void_t<decltype(std::declval<Fin&>().S)>> : std::true_type {};

template <int ND, typename Fin>
__global__ __launch_bounds__(kBlock) void k_finish(GridRed rb, uint32_t nb, Fin fin) {
  if constexpr (FinHasState<Fin>::value) {
    using St = typename std::remove_reference<decltype(*fin.S)>::type;
    constexpr int kWords = (int)(sizeof(St) / sizeof(uint32_t));
    static_assert(sizeof(St) % sizeof(uint32_t) == 0 && kWords <= kBlock, "state words");
    __shared__ uint32_t sw[kWords];
    uint32_t word = 0;
    if ((int)threadIdx.x < kWords) word = reinterpret_cast<const uint32_t*>(fin.S)[threadIdx.x];
    CSum a[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] = csum_zero();
    merge_partials<ND>(rb.partials, nb, a);  // (stale when the producer was gated off; then unused)
    if ((int)threadIdx.x < kWords) sw[threadIdx.x] = word;
    __syncthreads();
    Fin lf = fin;
    lf.S = reinterpret_cast<St*>(sw);
    if (!lf.active()) return;  // the same LDS word for every thread
    block_reduce<ND>(a);
    if (threadIdx.x == 0) {
      float tot[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) tot[d] = csum_val(a[d]);
      lf(tot);
      lf.after();
    }
    __syncthreads();
    if ((int)threadIdx.x < kWords) reinterpret_cast<uint32_t*>(fin.S)[threadIdx.x] = sw[threadIdx.x];
  } else {
    if (!fin.active()) return;
    CSum a[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] = csum_zero();
    merge_partials<ND>(rb.partials, nb, a);
    block_reduce<ND>(a);
    if (threadIdx.x == 0) {
      float tot[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) tot[d] = csum_val(a[d]);
      fin(tot);
      fin.after();
    }
  }
}

// Workgroups are dispatched round-robin over the XCDs; renumber them so that
// each XCD walks a contiguous range of work items (better L2 reuse of the
// gathered vector for banded matrices). Bijective on [0, nb).
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t wg, uint32_t nb) {
  const uint32_t xcd = wg % kNumXcd, idx = wg / kNumXcd;
  const uint32_t per = nb / kNumXcd, rem = nb % kNumXcd;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// p[c] from global memory with the byte offset c * sizeof(T) formed in 32 bits
// (the caller guarantees it is below 2^32): for a p held in scalar registers
// the load takes one 32-bit offset register instead of a 64-bit address.
template <typename T>
__device__ __forceinline__ T ld_off32(const T* __restrict__ p, uint32_t c) {
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) T gT;
  return *reinterpret_cast<gT*>(reinterpret_cast<gchar*>((gT*)p) + c * (uint32_t)sizeof(T));
}

// p, which must be the same in every lane, materialised in scalar registers:
// keeps the compiler from folding a per-lane offset into it (one 64-bit
// address register pair per load instead of one shared offset register).
template <typename T>
__device__ __forceinline__ const T* scalar_ptr(const T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<const T*>(((uint64_t)hi << 32) | lo);
}

// ---------------------------------------------------------------------------
// Matrix value types: FP32, or 16-bit storage (BF16 / FP16) widened to FP32.
// Value formats of the CSR SpMV. The 16-bit formats keep the diagonal exact to
// FP32 through a per-row remainder (kLowDiag): y_i gains dlo_i * x_i.
struct ValF32 {
  typedef float storage;
  static constexpr bool kLowDiag = false;
  static constexpr int kType = 0;  // msp_value_type
  static __device__ __forceinline__ f4 load4(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
  }
  static __device__ __forceinline__ float load1(const float* p) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ float cvt1(float v) { return v; }
  // four values as loaded (raw4) and converted (cvt4), so a load can be issued
  // well before its use
  typedef f4 raw4;
  static __device__ __forceinline__ raw4 ld4(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
  }
  static __device__ __forceinline__ f4 cvt4(raw4 w) { return w; }
};

struct ValBF16 {
  typedef uint16_t storage;
  static constexpr bool kLowDiag = true;
  static constexpr int kType = 1;
  static __device__ __forceinline__ float cvt(uint32_t h) { return __uint_as_float(h << 16); }
  static __device__ __forceinline__ f4 load4(const uint16_t* p) {
    u2 w = __builtin_nontemporal_load(reinterpret_cast<const u2*>(p));
    return f4{cvt(w.x & 0xFFFFu), cvt(w.x >> 16), cvt(w.y & 0xFFFFu), cvt(w.y >> 16)};
  }
  static __device__ __forceinline__ float load1(const uint16_t* p) { return cvt(*p); }
  static __device__ __forceinline__ float cvt1(uint16_t v) { return cvt(v); }
  typedef u2 raw4;
  static __device__ __forceinline__ raw4 ld4(const uint16_t* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const u2*>(p));
  }
  static __device__ __forceinline__ f4 cvt4(raw4 w) {
    return f4{cvt(w.x & 0xFFFFu), cvt(w.x >> 16), cvt(w.y & 0xFFFFu), cvt(w.y >> 16)};
  }
};

struct ValF16 {
  typedef uint16_t storage;
  static constexpr bool kLowDiag = true;
  static constexpr int kType = 2;
  static __device__ __forceinline__ float cvt(uint32_t h) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
  }
  static __device__ __forceinline__ f4 load4(const uint16_t* p) {
    u2 w = __builtin_nontemporal_load(reinterpret_cast<const u2*>(p));
    return f4{cvt(w.x & 0xFFFFu), cvt(w.x >> 16), cvt(w.y & 0xFFFFu), cvt(w.y >> 16)};
  }
  static __device__ __forceinline__ float load1(const uint16_t* p) { return cvt(*p); }
  static __device__ __forceinline__ float cvt1(uint16_t v) { return cvt(v); }
  typedef u2 raw4;
  static __device__ __forceinline__ raw4 ld4(const uint16_t* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const u2*>(p));
  }
  static __device__ __forceinline__ f4 cvt4(raw4 w) {
    return f4{cvt(w.x & 0xFFFFu), cvt(w.x >> 16), cvt(w.y & 0xFFFFu), cvt(w.y >> 16)};
  }
};

// BF16 values rounded one by one, without the diagonal remainder: for a
// preconditioner factor stored with its transpose (the two round to exact
// transposes of each other, so G^T G stays symmetric). Internal value type 3
// (not a public msp_value_type).
struct ValBF16Plain : ValBF16 {
  static constexpr bool kLowDiag = false;
  static constexpr int kType = 3;
};
constexpr int32_t kValuesBF16Plain = ValBF16Plain::kType;

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint16_t f32_to_f16(float f) {
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}

}  // namespace msp
