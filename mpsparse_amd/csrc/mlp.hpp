// Edge MLP of the learned FSAI preconditioner: 8 -> 64 -> 64 -> 1, ReLU, BF16
// MFMA operands with FP32 accumulation (v_mfma_f32_32x32x8_bf16 for layer 1
// and the backward products, v_mfma_f32_32x32x16_bf16 for the forward's
// layer 2: A lane l holds A[l % 32][8 (l / 32) .. + 7], B lane l holds
// B[8 (l / 32) .. + 7][l % 32], C as below).
//
// Parameters (FP32 master copy, flat): W1[64][8] | b1[64] | W2[64][64] | b2[64]
// | w3[64] | b3 -> 4801 floats. Each wave handles tiles of 32 edges.
//
// MFMA 32x32x8 operand layout (lane l, h = l / 32, c = l % 32):
//   A (32 x 8, rows = M):   lane holds A[c][4h .. 4h+3]
//   B (8 x 32, cols = N):   lane holds B[4h .. 4h+3][c]
//   C (32 x 32):            lane holds C[8j + 4h + i][c] in acc[4j + i]
// so a C tile read lane-wise is directly the A operand of its transpose (rows
// of A = columns of C) and the B operand of a product that sums over C's rows.
#pragma once

#include "common.hpp"

namespace msp {
namespace mlp {

constexpr int kIn = 8;
constexpr int kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
constexpr int kParams = oB3 + 1;  // 4801

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16 with the hardware converter (v_cvt_pk_bf16_f32: round to
// nearest even, two values per instruction), which replaced ~5 integer
// instructions per value of the portable rounding.
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ short bf16_bits(float f) { return __builtin_bit_cast(short, (__bf16)f); }

__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
  const bf2 lo = __builtin_convertvector(fl2{a, b}, bf2), hi = __builtin_convertvector(fl2{c, d}, bf2);
  const uint32_t l = __builtin_bit_cast(uint32_t, lo), h = __builtin_bit_cast(uint32_t, hi);
  return s4{(short)(l & 0xFFFFu), (short)(l >> 16), (short)(h & 0xFFFFu), (short)(h >> 16)};
}

__device__ __forceinline__ f16 mfma(s4 a, s4 b, f16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f16 mfma16(bf8 a, bf8 b, f16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Row stride of the per-wave BF16 staging tile (32 rows): 72 halves keeps
// each lane's 8 consecutive values one aligned 16-byte LDS read.
constexpr int kStage = 72;

__device__ __forceinline__ f16 zero16() {
  f16 z;
#pragma unroll
  for (int i = 0; i < 16; ++i) z[i] = 0.0f;
  return z;
}

// Weights of the forward pass as B operands (kept in registers).
struct FwdWeights {
  s4 b1[2];     // layer 1: B[k][n] = W1[n][k], n-tile t
  bf8 b2[2][4];  // layer 2: B[k][n] = W2[n][k], n-tile t, 16-deep k-chunk q
  float bias1[2], bias2[2], w3[2];
  float bias3;
};

__device__ __forceinline__ void load_fwd(const float* __restrict__ P, FwdWeights& W) {
  const int l = threadIdx.x & 63, h = l >> 5, c = l & 31;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int n = 32 * t + c;
    const float* w1 = P + oW1 + n * kIn + 4 * h;
    W.b1[t] = pack4(w1[0], w1[1], w1[2], w1[3]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* w2 = P + oW2 + n * kH + 16 * q + 8 * h;
#pragma unroll
      for (int i = 0; i < 8; ++i) W.b2[t][q][i] = (__bf16)w2[i];
    }
    W.bias1[t] = P[oB1 + n];
    W.bias2[t] = P[oB2 + n];
    W.w3[t] = P[oW3 + n];
  }
  W.bias3 = P[oB3];
}

// Forward of one 32-edge tile. Returns H1, H2 (post-ReLU, C layout, n-tiles 0/1)
// and leaves the per-lane layer-3 partial sums in `part` (row 8j+4h+i).
// `sh` is this wave's 32 x 64 BF16 staging area (row stride kStage halves,
// 16-byte aligned).
__device__ __forceinline__ void forward_tile(const float* __restrict__ feat, uint64_t e0, uint64_t ne,
                                             const FwdWeights& W, short* sh, f16 (&H1)[2], f16 (&H2)[2]) {
  const int l = threadIdx.x & 63, h = l >> 5, c = l & 31;
  // layer 1: A = features of edge e0 + c, k = 4h..4h+3
  s4 a1;
  {
    const uint64_t e = e0 + c;
    if (e < ne) {
      const f4 f = *reinterpret_cast<const f4*>(feat + e * kIn + 4 * h);
      a1 = pack4(f.x, f.y, f.z, f.w);
    } else {
      a1 = s4{0, 0, 0, 0};
    }
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    H1[t] = mfma(a1, W.b1[t], zero16());
#pragma unroll
    for (int r = 0; r < 16; ++r) H1[t][r] = fmaxf(H1[t][r] + W.bias1[t], 0.0f);
  }
  // stage H1 row-major (m, k) as BF16 to build layer-2 A operands
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) sh[(8 * j + 4 * h + i) * kStage + 32 * t + c] = bf16_bits(H1[t][4 * j + i]);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
  for (int t = 0; t < 2; ++t) H2[t] = zero16();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bf8 a2 = *reinterpret_cast<const bf8*>(sh + c * kStage + 16 * q + 8 * h);
#pragma unroll
    for (int t = 0; t < 2; ++t) H2[t] = mfma16(a2, W.b2[t][q], H2[t]);
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) H2[t][r] = fmaxf(H2[t][r] + W.bias2[t], 0.0f);
}

// Sum over each 32-lane half of the wave, complete in lanes 16-31 and 48-63:
// quad swaps (xor 1, xor 2), half-row and row mirrors (8- then 16-lane sums),
// then row 0 into row 1 and row 2 into row 3 (VALU lane moves: the
// ds_bpermute shuffles they replace went through LDS).
__device__ __forceinline__ float half_wave_sum(float v) {
  v += dpp_mov<kDppXor1>(v);
  v += dpp_mov<kDppXor2>(v);
  v += dpp_mov<kDppHalfMirror>(v);
  v += dpp_mov<kDppMirror>(v);
  v += dpp_mov<kDppBcast15, 0xA>(v);  // rows 1 and 3 only
  return v;
}

// o[m] for the tile's rows: sum over the 32 lanes of a half-wave of w3[n]
// H2[m][n]; valid in lanes c >= 16 of each half.
__device__ __forceinline__ void layer3(const FwdWeights& W, const f16 (&H2)[2], float (&o)[16]) {
#pragma unroll
  for (int r = 0; r < 16; ++r) o[r] = W.w3[0] * H2[0][r] + W.w3[1] * H2[1][r];
#pragma unroll
  for (int r = 0; r < 16; ++r) o[r] = half_wave_sum(o[r]) + W.bias3;
}

// out[e] = MLP(feat[e]) for all edges. Bounded to 128 registers (4 waves per
// SIMD): without the bound the compiler kept 132 and ran 3 waves per SIMD, and
// the MFMA chain of one tile has too little independent work to hide the
// conversions, the LDS staging and the layer-3 reduction with fewer waves.
__global__ __launch_bounds__(kBlock, 4) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ feat,
                                                        uint64_t ne, float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) short stage[kWavesPerBlock][32 * kStage];
  const int l = threadIdx.x & 63, h = l >> 5, c = l & 31, wid = threadIdx.x / kWave;
  FwdWeights W;
  load_fwd(P, W);
  const uint64_t tiles = (ne + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.x * kWavesPerBlock + wid; tile < tiles;
       tile += (uint64_t)gridDim.x * kWavesPerBlock) {
    const uint64_t e0 = tile * 32;
    f16 H1[2], H2[2];
    forward_tile(feat, e0, ne, W, stage[wid], H1, H2);
    float o[16];
    layer3(W, H2, o);
    if (c == 31) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint64_t e = e0 + 8 * j + 4 * h + i;
          if (e < ne) out[e] = o[4 * j + i];
        }
    }
  }
}

// Backward: given go[e] = dL/d out[e], accumulate dL/dP over all edges. Each
// workgroup writes its sums to partial[blockIdx.x][kParams] (the forward is
// recomputed); k_mlp_reduce sums the partials in block order.
__global__ __launch_bounds__(kBlock) void k_mlp_backward(const float* __restrict__ P, const float* __restrict__ feat,
                                                         const float* __restrict__ go, uint64_t ne,
                                                         float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) short stage[kWavesPerBlock][32 * kStage];
  __shared__ float red[kParams];
  const int l = threadIdx.x & 63, h = l >> 5, c = l & 31, wid = threadIdx.x / kWave;
  for (int i = threadIdx.x; i < kParams; i += kBlock) red[i] = 0.0f;
  FwdWeights W;
  load_fwd(P, W);
  // W2 as B operand of dH1 = dH2 W2: B[k = n][col = k'] = W2[n][k'], k-chunk q of n
  s4 w2t[2][8];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kk = 32 * u + c, n0 = 8 * q + 4 * h;
      w2t[u][q] = pack4(P[oW2 + (n0 + 0) * kH + kk], P[oW2 + (n0 + 1) * kH + kk], P[oW2 + (n0 + 2) * kH + kk],
                        P[oW2 + (n0 + 3) * kH + kk]);
    }
  f16 gW2[2][2], gW1[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    gW1[t] = zero16();
#pragma unroll
    for (int u = 0; u < 2; ++u) gW2[t][u] = zero16();
  }
  float gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gw3[2] = {0.f, 0.f}, gb3 = 0.f;
  const uint64_t tiles = (ne + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.x * kWavesPerBlock + wid; tile < tiles;
       tile += (uint64_t)gridDim.x * kWavesPerBlock) {
    const uint64_t e0 = tile * 32;
    f16 H1[2], H2[2];
    forward_tile(feat, e0, ne, W, stage[wid], H1, H2);
    float g[16];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t e = e0 + 8 * j + 4 * h + i;
        g[4 * j + i] = e < ne ? go[e] : 0.0f;
      }
    // dH2 = go w3 [H2 > 0]; w3 / b3 / b2 gradients
    f16 dH2[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        dH2[t][r] = H2[t][r] > 0.0f ? g[r] * W.w3[t] : 0.0f;
        gw3[t] += g[r] * H2[t][r];
        gb2[t] += dH2[t][r];
      }
    if (c == 0)
#pragma unroll
      for (int r = 0; r < 16; ++r) gb3 += g[r];
    // dW2[n][k] += sum_m dH2[m][n] H1[m][k]: A = dH2^T, B = H1 straight from C layout
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const s4 a = pack4(dH2[t][4 * q], dH2[t][4 * q + 1], dH2[t][4 * q + 2], dH2[t][4 * q + 3]);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const s4 b = pack4(H1[u][4 * q], H1[u][4 * q + 1], H1[u][4 * q + 2], H1[u][4 * q + 3]);
          gW2[t][u] = mfma(a, b, gW2[t][u]);
        }
      }
    }
    // dH1 = (dH2 W2) [H1 > 0]: A = dH2 row-major via LDS
    short* sh = stage[wid];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) sh[(8 * j + 4 * h + i) * kStage + 32 * t + c] = bf16_bits(dH2[t][4 * j + i]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    f16 dH1[2] = {zero16(), zero16()};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const short* row = sh + c * kStage + 8 * q + 4 * h;
      const s4 a = s4{row[0], row[1], row[2], row[3]};
#pragma unroll
      for (int u = 0; u < 2; ++u) dH1[u] = mfma(a, w2t[u][q], dH1[u]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        dH1[u][r] = H1[u][r] > 0.0f ? dH1[u][r] : 0.0f;
        gb1[u] += dH1[u][r];
      }
    // dW1[n][k'] += sum_m dH1[m][n] F[m][k']: A = dH1^T (C layout), B = F (k' = c < 8)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float fb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t e = e0 + 8 * q + 4 * h + i;
        fb[i] = (c < kIn && e < ne) ? feat[e * kIn + c] : 0.0f;
      }
      const s4 b = pack4(fb[0], fb[1], fb[2], fb[3]);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const s4 a = pack4(dH1[t][4 * q], dH1[t][4 * q + 1], dH1[t][4 * q + 2], dH1[t][4 * q + 3]);
        gW1[t] = mfma(a, b, gW1[t]);
      }
    }
  }
  __syncthreads();
  // per-workgroup sums in LDS, the waves taking turns (fixed order)
  for (int w = 0; w < kWavesPerBlock; ++w) {
    if (wid == w) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int n = 32 * t + c;
        // gW2 tile (t, u): rows n' = 32t + 8j + 4h + i, column k = 32u + c
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              red[oW2 + (32 * t + 8 * j + 4 * h + i) * kH + 32 * u + c] += gW2[t][u][4 * j + i];
        // gW1 tile t: rows n' = 32t + 8j + 4h + i, column k' = c (< 8)
        if (c < kIn)
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[oW1 + (32 * t + 8 * j + 4 * h + i) * kIn + c] += gW1[t][4 * j + i];
        // column sums: lanes c and c + 32 hold the same column -> h order
        for (int hh = 0; hh < 2; ++hh) {
          if (h == hh) {
            red[oB1 + n] += gb1[t];
            red[oB2 + n] += gb2[t];
            red[oW3 + n] += gw3[t];
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      // gb3 lives in lanes 0 and 32 (rows of each half-wave): add them in turn
      if (l == 0) red[oB3] += gb3;
      __builtin_amdgcn_wave_barrier();
      if (l == 32) red[oB3] += gb3;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kParams; i += kBlock) partial[(size_t)blockIdx.x * kParams + i] = red[i];
}

// grad[i] = sum over blocks of partial[b][i] (fixed order).
__global__ __launch_bounds__(kBlock) void k_mlp_reduce(const float* __restrict__ partial, uint32_t nblocks,
                                                       float* __restrict__ grad) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)kParams; i += gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (uint32_t b = 0; b < nblocks; ++b) s += partial[(size_t)b * kParams + i];
    grad[i] = s;
  }
}

// Adam step on the FP32 master parameters.
__global__ __launch_bounds__(kBlock) void k_adam(float* __restrict__ P, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                 float b2, float eps, float c1, float c2, float wd) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kParams; i += gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * P[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    P[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  }
}

// go[e] = grad_g[e] * colscale[e] (dL/d out from dL/d g)
__global__ __launch_bounds__(kBlock) void k_scale_grad(const float* __restrict__ gg, const float* __restrict__ cs,
                                                       uint64_t ne, float* __restrict__ go) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (uint64_t)gridDim.x * blockDim.x)
    go[e] = gg[e] * cs[e];
}

}  // namespace mlp
}  // namespace msp
