// Mixed-precision iterative refinement: FP64 only for the outer residual
// r = b - A x and the accumulation x += d; the correction d comes from an FP32
// Krylov solve whose matrix values may be stored in BF16 / FP16.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "krylov.hpp"

namespace msp {

// Block-wide sums of ND doubles; result valid in thread 0.
template <int ND>
__device__ __forceinline__ void block_reduce_d(double (&acc)[ND]) {
  __shared__ double red[kWavesPerBlock][ND];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    double v = acc[d];
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (lane == 0) red[wid][d] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double t = 0.0;
      for (int w = 0; w < kWavesPerBlock; ++w) t += red[w][d];
      acc[d] = t;
    }
  }
  __syncthreads();
}

// One workgroup: out[d] = sum over nb blocks of partials[b][d], in block order
// of a fixed tree (deterministic).
template <int ND>
__global__ __launch_bounds__(kBlock) void k_finish_d(const double* __restrict__ partials, uint32_t nb,
                                                     double* __restrict__ out) {
  double a[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) a[d] = 0.0;
  for (uint32_t i = threadIdx.x; i < nb; i += kBlock) {
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] += partials[i * ND + d];
  }
  block_reduce_d<ND>(a);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) out[d] = a[d];
  }
}

// r = b - A x in FP64 (matrix values FP32, exact in FP64); block partials of
// <r,r>, <b,b> (finish with k_finish_d<2>).
template <typename Rows>
__global__ __launch_bounds__(kBlock) void k_residual64(Rows rows, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* __restrict__ r,
                                                       uint32_t n, double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double s = b[i];
    if constexpr (std::is_same<Rows, RowsCsr>::value) {
      // 8 entries at a time: values and columns, then the x gathers, then the
      // products in entry order (one entry at a time waited two memory round
      // trips per entry); the same sum, bit for bit
      constexpr int kB = 8;
      const uint32_t j0 = rows.row_ptr[i], j1 = rows.row_ptr[i + 1];
      for (uint32_t jb = j0; jb < j1; jb += kB) {
        float av[kB];
        uint32_t cv[kB];
        double xv[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const bool in = jb + u < j1;
          av[u] = in ? rows.val[jb + u] : 0.0f;
          cv[u] = in ? rows.col[jb + u] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) xv[u] = x[cv[u]];
#pragma unroll
        for (int u = 0; u < kB; ++u)
          if (jb + u < j1) s -= (double)av[u] * xv[u];
      }
    } else {
      rows.foreach(i, [&](float a, uint32_t c) { s -= (double)a * x[c]; });
    }
    r[i] = s;
    acc[0] += s * s;
    acc[1] += b[i] * b[i];
  }
  block_reduce_d<2>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 2] = acc[0];
    partials[blockIdx.x * 2 + 1] = acc[1];
  }
}

// The same with L lanes per row (rows averaging many entries): coalesced row
// loads, lane sums combined by a butterfly.
template <int L>
__global__ __launch_bounds__(kBlock) void k_residual64_lanes(RowsCsr rows, const double* __restrict__ x,
                                                             const double* __restrict__ b, double* __restrict__ r,
                                                             uint32_t n, double* __restrict__ partials) {
  constexpr uint32_t G = kBlock / L;
  const uint32_t lane = threadIdx.x % L, grp = threadIdx.x / L;
  double acc[2] = {0.0, 0.0};
  for (uint32_t i = blockIdx.x * G + grp; i < n; i += gridDim.x * G) {  // uniform within a lane group
    double s = 0.0;
    for (uint32_t j = rows.row_ptr[i] + lane; j < rows.row_ptr[i + 1]; j += L) s -= (double)rows.val[j] * x[rows.col[j]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) {
      s += b[i];
      r[i] = s;
      acc[0] += s * s;
      acc[1] += b[i] * b[i];
    }
  }
  block_reduce_d<2>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 2] = acc[0];
    partials[blockIdx.x * 2 + 1] = acc[1];
  }
}

// out = float(r * scale)
__global__ __launch_bounds__(kBlock) void k_scale_to_f32(const double* __restrict__ r, float* __restrict__ out,
                                                         double scale, uint32_t n) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) out[i] = (float)(r[i] * scale);
}

// x += scale * d (FP64 accumulation of the FP32 correction)
__global__ __launch_bounds__(kBlock) void k_axpy64(double* __restrict__ x, const float* __restrict__ d, double scale,
                                                   uint32_t n) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] += scale * (double)d[i];
}

__global__ __launch_bounds__(kBlock) void k_fill64(double* __restrict__ x, double v, uint32_t n) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = v;
}

}  // namespace msp
