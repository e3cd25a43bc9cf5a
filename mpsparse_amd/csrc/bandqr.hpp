// Direct sparse solve by Householder QR of a banded ordering (FP64).
//
// Replaces the reference's CPU factorisation SparseQRSolver
// (sparse_bridge/bridge.cpp:14-191: Accelerate SparseFactor(QR) + SparseSolve).
// The host orders the matrix into a band (reverse Cuthill-McKee on A + A^T for
// square matrices; rows sorted by leftmost column for tall ones), so that the
// ordered matrix has lower bandwidth kl and upper bandwidth ku. The GPU then:
//   k_bqr_fill     scatters the CSC entries into a row-major band
//                  RB[i][k] = A(i, i - kl + k), k in [0, kl + ku] (duplicates added);
//   k_bqr_factor   one workgroup walks the columns: reflector j spans rows
//                  j .. j + kl and updates columns j + 1 .. j + kl + ku, all in a
//                  sliding LDS window of (kl + 1) rows x (kl + ku + 1) columns
//                  (R's upper bandwidth grows to kl + ku); finished rows of R and
//                  reflectors stream out to HBM, new band rows stream in through
//                  a register prefetch one chunk ahead;
//   k_bqr_apply_qt y = Q^T b and
//   k_bqr_rsolve   x = R^-1 y: one thread per right-hand side, the recurrence
//                  window in LDS, reflectors / R rows staged per chunk.
// The column loop is sequential, so a factorisation costs n steps of
// O(kl (kl + ku)) work on one CU: a solver for banded / RCM-banded systems
// (1D and 2D grids, narrow-band operators), not for matrices whose ordering
// leaves a wide band (kl + ku + 1 > 256, kl > 127, or windows beyond the
// device's LDS are refused with MSP_ERR_UNSUPPORTED).
#pragma once

#include "common.hpp"

namespace msp {
namespace bqr {

constexpr int kMaxWC = 256;       // kl + ku + 1 (LDS use is checked against the device as well)
constexpr int kMaxKl = 127;       // the reflector (kl + 1 entries) is formed by one wave, two entries a lane
constexpr int kStage = 8;         // band rows per factor refill
constexpr int kChunk = 32;        // steps per refill in the solves
constexpr int kRhs = 64;          // right-hand sides per solve workgroup

// Scatter CSC entries into the ordered band (zero-filled by the caller). One
// thread per source column, entries in CSC order, so duplicates add in a fixed
// order. Entries outside the band (impossible for a correct ordering) are
// counted in bad.
__global__ __launch_bounds__(kBlock) void k_bqr_fill(const int64_t* __restrict__ col_starts,
                                                     const int64_t* __restrict__ row_idx,
                                                     const double* __restrict__ vals, int64_t ncols,
                                                     const int32_t* __restrict__ prow,
                                                     const int32_t* __restrict__ pcol, int kl, int wc,
                                                     double* __restrict__ RB, int* __restrict__ bad) {
  for (int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x; c < ncols; c += (int64_t)gridDim.x * kBlock) {
    const int64_t nc = pcol[c];
    for (int64_t e = col_starts[c]; e < col_starts[c + 1]; ++e) {
      const int64_t nr = prow[row_idx[e]];
      const int64_t k = nc - nr + kl;
      if (k < 0 || k >= wc) {
        atomicOr(bad, 1);
        continue;
      }
      RB[nr * wc + k] += vals[e];
    }
  }
}

// Workgroup barrier that orders LDS only: __syncthreads() also waits for the
// wave's outstanding global stores (the streamed-out rows of R and reflectors),
// which put an HBM write round trip into every column step. Nothing in these
// kernels reads back global memory written by another thread.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Householder QR of the band. LDS: window [kl+1][wc], stage [kStage][wc],
// reflector [kl+1]. Outputs: Rb[j][c] = R(j, j + c) (c < wc, j + c < n),
// Vb[j][i-1] = v_i of reflector j (v_0 = 1 implicit; 0 past the last row),
// tau[j]. H_j = I - tau_j v v^T; Q = H_0 H_1 ... H_{n-1}.
template <int T>
__global__ __launch_bounds__(T) void k_bqr_factor(const double* __restrict__ RB, int64_t m, int64_t n, int kl,
                                                  int ku, double* __restrict__ Rb, double* __restrict__ Vb,
                                                  double* __restrict__ tau_out) {
  extern __shared__ double smem[];
  const int WR = kl + 1, WC = kl + ku + 1;
  double* win = smem;                // [WR][WC]: row r in slot r % WR, column c in slot c % WC
  double* stage = win + WR * WC;     // [kStage][WC]: band rows entering next
  double* sv = stage + kStage * WC;  // [WR]: current reflector
  __shared__ double s_tau;
  const int t = threadIdx.x;
  for (int e = t; e < WR * WC; e += T) {  // rows 0..kl, columns 0..WC-1
    const int r = e / WC, c = e % WC, k = c - r + kl;
    win[e] = (r < m && c < n && k >= 0 && k < WC) ? RB[(int64_t)r * WC + k] : 0.0;
  }
  constexpr int MAXW = T == kWave ? kWave : kMaxWC;  // one-wave workgroups take only windows up to 64 wide
  constexpr int PER = (kStage * MAXW + T - 1) / T;
  double pre[PER];
  // band rows first .. first + kStage - 1 into registers (not waited for here)
  auto fetch = [&](int64_t first) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + q * T;
      const int64_t row = first + e / WC;
      pre[q] = (e < kStage * WC && row < m) ? RB[row * WC + e % WC] : 0.0;
    }
  };
  fetch(kl + 1);
  for (int64_t j = 0; j < n; ++j) {
    if (j % kStage == 0) {  // rows j + kl + 1 .. j + kl + kStage enter during the next kStage steps
      lds_barrier();
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = t + q * T;
        if (e < kStage * WC) stage[e] = pre[q];
      }
      fetch(j + kl + 1 + kStage);
      lds_barrier();
    }
    const int L = (int)(m - j < (int64_t)WR ? m - j : (int64_t)WR);
    const int W = (int)(n - 1 - j < (int64_t)(WC - 1) ? n - 1 - j : (int64_t)(WC - 1));
    const int cj = (int)(j % WC), rj = (int)(j % WR);
    // reflector of column j, rows j .. j + L - 1 (wave 0; two entries per lane)
    if (t < kWave) {
      const int i0 = t, i1 = t + kWave;
      const int s0 = (rj + i0) % WR, s1 = (rj + i1) % WR;
      const double a = i0 < L ? win[s0 * WC + cj] : 0.0;
      const double b = i1 < L ? win[s1 * WC + cj] : 0.0;
      const double ss = wave_sum_d((i0 >= 1 ? a * a : 0.0) + b * b);
      const double alpha = __shfl(a, 0, kWave);
      double beta = alpha, tau = 0.0, scal = 0.0;
      if (ss != 0.0) {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      if (i0 < L) {
        sv[i0] = i0 == 0 ? 1.0 : a * scal;
        win[s0 * WC + cj] = i0 == 0 ? beta : a * scal;
      }
      if (i1 < L) {
        sv[i1] = b * scal;
        win[s1 * WC + cj] = b * scal;
      }
      if (t == 0) s_tau = tau;
    }
    lds_barrier();
    const double tau = s_tau;
    if (tau != 0.0) {
      // rows j .. j + L - 1 sit in slots rj .. WR - 1, then 0 ..: two straight
      // runs (no wrap test in the loop, so the LDS loads can be issued ahead)
      const int n1 = L < WR - rj ? L : WR - rj;
      for (int c = t; c < W; c += T) {  // column j + 1 + c
        const int cs = cj + 1 + c < WC ? cj + 1 + c : cj + 1 + c - WC;
        double dot = 0.0;
#pragma unroll 4
        for (int i = 0; i < n1; ++i) dot += sv[i] * win[(rj + i) * WC + cs];
#pragma unroll 4
        for (int i = n1; i < L; ++i) dot += sv[i] * win[(i - n1) * WC + cs];
        const double f = tau * dot;
#pragma unroll 4
        for (int i = 0; i < n1; ++i) win[(rj + i) * WC + cs] -= f * sv[i];
#pragma unroll 4
        for (int i = n1; i < L; ++i) win[(i - n1) * WC + cs] -= f * sv[i];
      }
    }
    lds_barrier();
    // retire row j (a finished row of R) and column j (the reflector)
    for (int c = t; c < WC; c += T)
      if (j + c < n) Rb[j * WC + c] = win[rj * WC + (cj + c < WC ? cj + c : cj + c - WC)];
    for (int i = 1 + t; i < WR; i += T) Vb[j * kl + (i - 1)] = i < L ? win[((rj + i) % WR) * WC + cj] : 0.0;
    if (t == 0) tau_out[j] = tau;
    lds_barrier();
    // column j's slot becomes column j + WC (rows j+1 .. j+kl are fill: zero);
    // row j's slot becomes row j + kl + 1 (band row from the stage)
    for (int i = 1 + t; i < WR; i += T) win[((rj + i) % WR) * WC + cj] = 0.0;
    const int64_t e = j + kl + 1;
    const int srow = (int)(j % kStage);
    for (int c = t; c < WC; c += T)
      win[rj * WC + (cj + 1 + c < WC ? cj + 1 + c : cj + 1 + c - WC)] = e < m ? stage[srow * WC + c] : 0.0;
    lds_barrier();
  }
}

// Narrow bands (kl + 1 <= kSmallWR, kl + ku + 1 <= kSmallWC): the same
// factorisation by lane 0 alone with the window in registers (relative
// indexing: w[r][c] holds row j + r, column j + c and shifts by one each step),
// so a step is a dependent chain of a few flops instead of several LDS round
// trips and wave reductions. The other lanes only prefetch: each holds one
// band row of the block after next and parks it in an LDS stage for lane 0.
constexpr int kSmallWR = 4, kSmallWC = 8, kSmallD = 32;

// KL, KU >= 0: the band shape fixed at compile time (straight-line steps, the
// tridiagonal case); KL < 0: kl, ku from the arguments.
template <int KL, int KU>
__global__ __launch_bounds__(kWave) void k_bqr_factor_small(const double* __restrict__ RB, int64_t m, int64_t n,
                                                            int kl_arg, int ku_arg, double* __restrict__ Rb,
                                                            double* __restrict__ Vb, double* __restrict__ tau_out) {
  constexpr int R = KL >= 0 ? KL + 1 : kSmallWR, C = KL >= 0 ? KL + KU + 1 : kSmallWC, D = kSmallD;
  __shared__ double stage[D][C];
  const int kl = KL >= 0 ? KL : kl_arg, ku = KL >= 0 ? KU : ku_arg;
  const int WR = kl + 1, WC = kl + ku + 1;
  const int lane = threadIdx.x;
  // Rows at or past m and columns at or past n are zero in the window (the
  // band and the stage hold zeros there), so reflectors and updates need no
  // edge tests: zero rows add nothing, zero columns stay zero.
  double w[R][C];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = c - r + kl;
      w[r][c] = (r < WR && c < WC && r < m && c < n && k >= 0 && k < WC) ? RB[(int64_t)r * WC + k] : 0.0;
    }
  // lane l < D holds band row first + l (the row entering at step first - kl - 1 + l)
  double held[C];
  auto fetch = [&](int64_t first) {
    const int64_t row = first + lane;
#pragma unroll
    for (int c = 0; c < C; ++c) held[c] = (lane < D && c < WC && row < m) ? RB[row * WC + c] : 0.0;
  };
  fetch(kl + 1);
  for (int64_t j0 = 0; j0 < n; j0 += D) {
    if (lane < D) {
#pragma unroll
      for (int c = 0; c < C; ++c) stage[lane][c] = held[c];
    }
    __syncthreads();  // one wave: orders the stage writes before lane 0's reads
    fetch(j0 + D + kl + 1);
    if (lane == 0) {
      const int steps = (int)(n - j0 < (int64_t)D ? n - j0 : (int64_t)D);
      for (int d = 0; d < steps; ++d) {
        const int64_t j = j0 + d;
        double nrow[C];  // row j + kl + 1, read early
#pragma unroll
        for (int c = 0; c < C; ++c) nrow[c] = stage[d][c];
        const double alpha = w[0][0];
        double ss = 0.0;
#pragma unroll
        for (int i = 1; i < R; ++i) ss += w[i][0] * w[i][0];
        double beta = alpha, tau = 0.0, scal = 0.0;
        if (ss != 0.0) {
          beta = -copysign(sqrt(alpha * alpha + ss), alpha);
          tau = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        double v[R];
        v[0] = 1.0;
#pragma unroll
        for (int i = 1; i < R; ++i) v[i] = w[i][0] * scal;
#pragma unroll
        for (int c = 1; c < C; ++c) {
          double dot = w[0][c];
#pragma unroll
          for (int i = 1; i < R; ++i) dot += v[i] * w[i][c];
          const double f = tau * dot;
#pragma unroll
          for (int i = 0; i < R; ++i) w[i][c] -= f * v[i];
        }
        Rb[j * WC] = beta;
#pragma unroll
        for (int c = 1; c < C; ++c)
          if (c < WC && j + c < n) Rb[j * WC + c] = w[0][c];
#pragma unroll
        for (int i = 1; i < R; ++i)
          if (i < WR) Vb[j * kl + (i - 1)] = v[i];
        tau_out[j] = tau;
        // slide: row j + r + 1 / column j + c + 1 move to r / c; column j + WC
        // enters as fill (zero), row j + kl + 1 from the stage
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int c = 0; c < C; ++c) {
            double nv = 0.0;
            if (r + 1 < WR) nv = (c + 1 < C && c + 1 < WC) ? w[r + 1 < R ? r + 1 : 0][c + 1 < C ? c + 1 : 0] : 0.0;
            else if (r + 1 == WR) nv = c < WC ? nrow[c] : 0.0;
            w[r][c] = nv;
          }
      }
    }
    __syncthreads();  // lane 0 is done with the stage before it is refilled
  }
}

// y <- Q^T y for narrow bands: the window of rows j .. j + kl of this
// right-hand side in registers (shifted each step).
__global__ __launch_bounds__(kRhs) void k_bqr_apply_qt_small(const double* __restrict__ Vb,
                                                             const double* __restrict__ tau, int64_t m, int64_t n,
                                                             int kl, double* __restrict__ Y, int64_t nrhs) {
  constexpr int R = kSmallWR;
  __shared__ double vs[kChunk][R];
  __shared__ double ts[kChunk];
  const int WR = kl + 1;
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * kRhs + t;
  const bool act = col < nrhs;
  double w[R];
#pragma unroll
  for (int r = 0; r < R; ++r) w[r] = (act && r < WR && r < m) ? Y[(int64_t)r * nrhs + col] : 0.0;
  double pre[kChunk], cur[kChunk];
  auto fetch = [&](int64_t first) {
#pragma unroll
    for (int q = 0; q < kChunk; ++q) pre[q] = (act && first + q < m) ? Y[(first + q) * nrhs + col] : 0.0;
  };
  fetch(kl + 1);
  for (int64_t j0 = 0; j0 < n; j0 += kChunk) {
    lds_barrier();
    for (int e = t; e < kChunk * R; e += kRhs) {
      const int q = e / R, i = e % R;  // vs[q][i] = v_{i+1} of reflector j0 + q
      vs[q][i] = (j0 + q < n && i < kl) ? Vb[(j0 + q) * kl + i] : 0.0;
    }
    for (int e = t; e < kChunk; e += kRhs) ts[e] = j0 + e < n ? tau[j0 + e] : 0.0;
#pragma unroll
    for (int q = 0; q < kChunk; ++q) cur[q] = pre[q];
    fetch(j0 + kChunk + kl + 1);
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int64_t j = j0 + q;
      if (j >= n) break;
      const double tj = ts[q];
      double s = w[0];
#pragma unroll
      for (int i = 1; i < R; ++i) s += vs[q][i - 1] * w[i];  // v past the reflector's rows is 0
      s *= tj;
      w[0] -= s;
#pragma unroll
      for (int i = 1; i < R; ++i) w[i] -= s * vs[q][i - 1];
      if (act) Y[j * nrhs + col] = w[0];
      // shift: row j + 1 + r -> r; row j + kl + 1 enters at kl
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = (r + 1 < WR) ? w[r + 1 < R ? r + 1 : 0] : (r + 1 == WR ? cur[q] : 0.0);
    }
  }
  for (int64_t r = n; r < m && r < n + WR; ++r)
    if (act) Y[r * nrhs + col] = w[(int)(r - n)];
}

// x = R^-1 y for narrow bands (wc <= kSmallWC): x_{i+1} .. x_{i+wc-1} of this
// right-hand side in registers.
__global__ __launch_bounds__(kRhs) void k_bqr_rsolve_small(const double* __restrict__ Rb, int64_t n, int wc,
                                                           const double* __restrict__ Y, double* __restrict__ Xp,
                                                           int64_t nrhs) {
  constexpr int C = kSmallWC;
  __shared__ double rs[kChunk][C];
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * kRhs + t;
  const bool act = col < nrhs;
  double x[C];  // x[c] = x_{i+c}, c >= 1
#pragma unroll
  for (int c = 0; c < C; ++c) x[c] = 0.0;
  double pre[kChunk], cur[kChunk];
  auto fetch = [&](int64_t top) {
#pragma unroll
    for (int q = 0; q < kChunk; ++q) pre[q] = (act && top - q >= 0) ? Y[(top - q) * nrhs + col] : 0.0;
  };
  fetch(n - 1);
  for (int64_t top = n - 1; top >= 0; top -= kChunk) {
    lds_barrier();
    for (int e = t; e < kChunk * C; e += kRhs) {
      const int q = e / C, c = e % C;
      const int64_t i = top - q;
      rs[q][c] = (i >= 0 && c < wc && i + c < n) ? Rb[i * wc + c] : 0.0;  // zero past the matrix edge
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) cur[q] = pre[q];
    fetch(top - kChunk);
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int64_t i = top - q;
      if (i < 0) break;
      double s = cur[q];
#pragma unroll
      for (int c = 1; c < C; ++c) s -= rs[q][c] * x[c];
      const double xi = s / rs[q][0];
#pragma unroll
      for (int c = C - 1; c >= 2; --c) x[c] = x[c - 1];
      x[1] = xi;
      if (act) Xp[i * nrhs + col] = xi;
    }
  }
}

// Non-positive count of |R(i,i)| == 0 or non-finite (rank deficiency).
__global__ __launch_bounds__(kBlock) void k_bqr_check(const double* __restrict__ Rb, int64_t n, int wc,
                                                      int* __restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double d = Rb[i * wc];
    if (!(fabs(d) > 0.0) || !isfinite(d)) atomicOr(bad, 1);
  }
}

// Y[nr][k] = B[k * ldb + src_row[nr]] (ordered rows, right-hand sides
// interleaved so that one thread per right-hand side reads coalesced rows).
__global__ __launch_bounds__(kBlock) void k_bqr_gather_rhs(const double* __restrict__ B, int64_t ldb, int64_t m,
                                                           int64_t nrhs, const int32_t* __restrict__ src_row,
                                                           double* __restrict__ Y) {
  const int64_t total = m * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
    const int64_t nr = e / nrhs, k = e % nrhs;
    Y[e] = B[k * ldb + src_row[nr]];
  }
}

// X[k * ldx + c] = Xp[pcol[c]][k] (back to the caller's column order).
__global__ __launch_bounds__(kBlock) void k_bqr_scatter_x(const double* __restrict__ Xp, int64_t n, int64_t nrhs,
                                                          const int32_t* __restrict__ pcol, double* __restrict__ X,
                                                          int64_t ldx) {
  const int64_t total = n * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
    const int64_t k = e / n, c = e % n;
    X[k * ldx + c] = Xp[(int64_t)pcol[c] * nrhs + k];
  }
}

// Y <- Q^T Y (Y: [m][nrhs], in place). Thread k owns right-hand side
// blockIdx.x * kRhs + k; LDS: window [kl+1][kRhs], reflectors of the chunk
// [kChunk][kl], their tau [kChunk]. Rows entering the window are prefetched
// into registers one chunk ahead.
__global__ __launch_bounds__(kRhs) void k_bqr_apply_qt(const double* __restrict__ Vb, const double* __restrict__ tau,
                                                       int64_t m, int64_t n, int kl, double* __restrict__ Y,
                                                       int64_t nrhs) {
  extern __shared__ double smem[];
  const int WR = kl + 1;
  double* wb = smem;                 // [WR][kRhs]
  double* vs = wb + WR * kRhs;       // [kChunk][kl]
  double* ts = vs + kChunk * kl;     // [kChunk]
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * kRhs + t;
  const bool act = col < nrhs;
  for (int r = 0; r < WR; ++r) wb[r * kRhs + t] = (act && r < m) ? Y[(int64_t)r * nrhs + col] : 0.0;
  double pre[kChunk];
  auto fetch = [&](int64_t first) {  // rows first .. first + kChunk - 1 of this right-hand side
#pragma unroll
    for (int q = 0; q < kChunk; ++q) pre[q] = (act && first + q < m) ? Y[(first + q) * nrhs + col] : 0.0;
  };
  fetch(kl + 1);
  double cur[kChunk];
  for (int64_t j0 = 0; j0 < n; j0 += kChunk) {
    lds_barrier();
    for (int e = t; e < kChunk * kl; e += kRhs) {
      const int64_t jj = j0 + e / (kl > 0 ? kl : 1);
      vs[e] = jj < n ? Vb[j0 * kl + e] : 0.0;
    }
    for (int e = t; e < kChunk; e += kRhs) ts[e] = j0 + e < n ? tau[j0 + e] : 0.0;
#pragma unroll
    for (int q = 0; q < kChunk; ++q) cur[q] = pre[q];
    fetch(j0 + kChunk + kl + 1);
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int64_t j = j0 + q;
      if (j >= n) break;
      const int L = (int)(m - j < (int64_t)WR ? m - j : (int64_t)WR);
      const int rj = (int)(j % WR);
      const double tj = ts[q];
      const double* v = vs + q * kl;
      if (tj != 0.0) {
        // rows j + i in slots rj + i, wrapping once at WR (two straight runs)
        const int n1 = L < WR - rj ? L : WR - rj;
        double w = wb[rj * kRhs + t];
#pragma unroll 4
        for (int i = 1; i < n1; ++i) w += v[i - 1] * wb[(rj + i) * kRhs + t];
#pragma unroll 4
        for (int i = n1; i < L; ++i) w += v[i - 1] * wb[(i - n1) * kRhs + t];
        w *= tj;
        wb[rj * kRhs + t] -= w;
#pragma unroll 4
        for (int i = 1; i < n1; ++i) wb[(rj + i) * kRhs + t] -= w * v[i - 1];
#pragma unroll 4
        for (int i = n1; i < L; ++i) wb[(i - n1) * kRhs + t] -= w * v[i - 1];
      }
      if (act) Y[j * nrhs + col] = wb[rj * kRhs + t];  // row j of Q^T y is final
      wb[rj * kRhs + t] = cur[q];                       // row j + kl + 1 enters
    }
  }
  // rows n .. m-1 still in the window (tall matrices): write them back too
  for (int64_t r = n; r < m && r < n + WR; ++r)
    if (act) Y[r * nrhs + col] = wb[(int)(r % WR) * kRhs + t];
}

// Xp[i][k] = (Y[i][k] - sum_{c=1}^{wc-1} R(i, i+c) Xp[i+c][k]) / R(i,i), i = n-1 .. 0.
// LDS: x window [wc][kRhs], R rows of the chunk [kChunk][wc].
__global__ __launch_bounds__(kRhs) void k_bqr_rsolve(const double* __restrict__ Rb, int64_t n, int wc,
                                                     const double* __restrict__ Y, double* __restrict__ Xp,
                                                     int64_t nrhs) {
  extern __shared__ double smem[];
  double* wx = smem;                // [wc][kRhs]: x row i in slot i % wc
  double* rs = wx + wc * kRhs;      // [kChunk][wc]: R rows of the chunk, highest row first
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * kRhs + t;
  const bool act = col < nrhs;
  for (int s = 0; s < wc; ++s) wx[s * kRhs + t] = 0.0;
  double pre[kChunk], cur[kChunk];
  auto fetch = [&](int64_t top) {  // rows top, top-1, ..., top-kChunk+1
#pragma unroll
    for (int q = 0; q < kChunk; ++q) pre[q] = (act && top - q >= 0) ? Y[(top - q) * nrhs + col] : 0.0;
  };
  fetch(n - 1);
  for (int64_t top = n - 1; top >= 0; top -= kChunk) {
    lds_barrier();
    for (int e = t; e < kChunk * wc; e += kRhs) {
      const int64_t i = top - e / wc;
      rs[e] = i >= 0 ? Rb[i * wc + e % wc] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) cur[q] = pre[q];
    fetch(top - kChunk);
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int64_t i = top - q;
      if (i < 0) break;
      const double* r = rs + q * wc;
      const int lim = (int)(n - 1 - i < (int64_t)(wc - 1) ? n - 1 - i : (int64_t)(wc - 1));
      double s = cur[q];
      const int r0 = (int)(i % wc);
      const int c1 = lim < wc - 1 - r0 ? lim : wc - 1 - r0;  // x_{i+c} in slot r0 + c up to the wrap
#pragma unroll 4
      for (int c = 1; c <= c1; ++c) s -= r[c] * wx[(r0 + c) * kRhs + t];
#pragma unroll 4
      for (int c = c1 + 1; c <= lim; ++c) s -= r[c] * wx[(r0 + c - wc) * kRhs + t];
      const double x = s / r[0];
      wx[(int)(i % wc) * kRhs + t] = x;
      if (act) Xp[i * nrhs + col] = x;
    }
  }
}

}  // namespace bqr
}  // namespace msp
