// Factorized sparse approximate inverse (FSAI) support: M = G^T G with G lower
// triangular on the lower-triangle pattern of A (diagonal included).
//
// The reference has no preconditioner; this is the learned-preconditioner
// addition. G and G^T are kept as two CSR matrices so that applying M is two
// row-block SpMVs; G^T's values are gathered from G's through a permutation
// computed once by the radix sort.
#pragma once

#include "common.hpp"
#include "dia.hpp"

namespace msp {

constexpr int kFsaiMaxRow = 8;  // classical FSAI: largest lower row solved densely

// Lower-triangle entries per row (columns sorted within a row).
__global__ __launch_bounds__(kBlock) void k_lower_count(const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ col, uint32_t n,
                                                        uint32_t* __restrict__ cnt) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= n; r += gridDim.x * blockDim.x) {
    if (r == n) {
      cnt[r] = 0;
      continue;
    }
    uint32_t lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) {  // first entry with col > r
      uint32_t mid = lo + ((hi - lo) >> 1);
      if (col[mid] <= r) lo = mid + 1;
      else hi = mid;
    }
    cnt[r] = lo - row_ptr[r];
  }
}

// Lower pattern: Lcol, Lrow (row of each edge), Lsrc (position in A).
__global__ __launch_bounds__(kBlock) void k_lower_fill(const uint32_t* __restrict__ row_ptr,
                                                       const uint32_t* __restrict__ col, uint32_t n,
                                                       const uint32_t* __restrict__ Lrp, uint32_t* __restrict__ Lcol,
                                                       uint32_t* __restrict__ Lrow, uint32_t* __restrict__ Lsrc) {
  // a row's columns are loaded 8 at a time before they are stored (one load
  // then store per entry waited a memory round trip each)
  constexpr int kB = 8;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t a = row_ptr[r], o = Lrp[r], k = Lrp[r + 1] - o;
    for (uint32_t t0 = 0; t0 < k; t0 += kB) {
      uint32_t cv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) cv[u] = t0 + u < k ? col[a + t0 + u] : 0u;
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (t0 + u < k) {
          Lcol[o + t0 + u] = cv[u];
          Lrow[o + t0 + u] = r;
          Lsrc[o + t0 + u] = a + t0 + u;
        }
      }
    }
  }
}

// Transpose keys: (col << 32) | row, payload = edge index (as float bits).
__global__ __launch_bounds__(kBlock) void k_transpose_keys(const uint32_t* __restrict__ Lrow,
                                                           const uint32_t* __restrict__ Lcol, uint64_t ne,
                                                           uint64_t* __restrict__ keys, float* __restrict__ payload) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (uint64_t)gridDim.x * blockDim.x) {
    keys[e] = ((uint64_t)Lcol[e] << 32) | Lrow[e];
    payload[e] = __uint_as_float((uint32_t)e);
  }
}

// dst[k] = src[map[k]] (map carried as float bits from the sort payload)
__global__ __launch_bounds__(kBlock) void k_gather_by_payload(float* __restrict__ dst, const float* __restrict__ src,
                                                              const float* __restrict__ map, uint64_t cnt) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += (uint64_t)gridDim.x * blockDim.x)
    dst[k] = src[__float_as_uint(map[k])];
}

// a(r, c) by binary search in row r of A (0 if absent; duplicates summed).
__device__ __forceinline__ float csr_lookup(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                            const float* __restrict__ val, uint32_t r, uint32_t c) {
  uint32_t lo = row_ptr[r], hi = row_ptr[r + 1];
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo) >> 1);
    if (col[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  float s = 0.0f;
  for (uint32_t j = lo; j < row_ptr[r + 1] && col[j] == c; ++j) s += val[j];
  return s;
}

// Classical FSAI row by row: S = A[J, J] (J = lower pattern of row i, i last),
// solve S y = e_last by Cholesky, g = y / sqrt(y_last). Rows longer than
// kFsaiMaxRow, or with a non-positive pivot, fall back to the Jacobi row
// g = e_i / sqrt(a_ii) and are counted in `fallback`.
__global__ __launch_bounds__(kBlock) void k_fsai_classical(const uint32_t* __restrict__ row_ptr,
                                                           const uint32_t* __restrict__ col,
                                                           const float* __restrict__ val, uint32_t n,
                                                           const uint32_t* __restrict__ Lrp,
                                                           const uint32_t* __restrict__ Lcol, float* __restrict__ G,
                                                           uint32_t* __restrict__ fallback) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t o = Lrp[i], k = Lrp[i + 1] - o;
    bool ok = k >= 1 && k <= (uint32_t)kFsaiMaxRow && Lcol[o + k - 1] == i;
    double S[kFsaiMaxRow][kFsaiMaxRow];
    double y[kFsaiMaxRow];
    if (ok) {
      for (uint32_t p = 0; p < k; ++p)
        for (uint32_t q = 0; q <= p; ++q) {
          const float a = csr_lookup(row_ptr, col, val, Lcol[o + p], Lcol[o + q]);
          S[p][q] = a;
        }
      // Cholesky S = L L^T in place (lower triangle)
      for (uint32_t p = 0; p < k && ok; ++p) {
        double d = S[p][p];
        for (uint32_t t = 0; t < p; ++t) d -= S[p][t] * S[p][t];
        if (!(d > 0.0)) {
          ok = false;
          break;
        }
        d = sqrt(d);
        S[p][p] = d;
        for (uint32_t q = p + 1; q < k; ++q) {
          double s = S[q][p];
          for (uint32_t t = 0; t < p; ++t) s -= S[q][t] * S[p][t];
          S[q][p] = s / d;
        }
      }
      if (ok) {
        // L z = e_last, then L^T y = z
        for (uint32_t p = 0; p < k; ++p) {
          double s = (p == k - 1) ? 1.0 : 0.0;
          for (uint32_t t = 0; t < p; ++t) s -= S[p][t] * y[t];
          y[p] = s / S[p][p];
        }
        for (int p = (int)k - 1; p >= 0; --p) {
          double s = y[p];
          for (uint32_t t = p + 1; t < k; ++t) s -= S[t][p] * y[t];
          y[p] = s / S[p][p];
        }
        if (!(y[k - 1] > 0.0)) ok = false;
      }
    }
    if (ok) {
      const double sc = 1.0 / sqrt(y[k - 1]);
      for (uint32_t p = 0; p < k; ++p) G[o + p] = (float)(y[p] * sc);
    } else {
      atomicAdd(fallback, 1u);
      for (uint32_t p = 0; p < k; ++p) {
        const uint32_t c = Lcol[o + p];
        const float d = csr_lookup(row_ptr, col, val, i, i);
        G[o + p] = (c == i && d > 0.0f) ? 1.0f / sqrtf(d) : 0.0f;
      }
    }
  }
}

// Feature sets of the edge network.
enum FsaiFeatures : int {
  // [a_ij / sqrt(a_ii a_jj), log(a_jj / a_ii), nnz(row i) / 8, sum_j |a_ij| / a_ii,
  //  [j == i], s_low(i), s_all(i), t(i)] with the row couplings
  //  s_low(i) = sum_{j < i} a_ij^2 / (a_ii a_jj), s_all(i) = sum_{j != i} a_ij^2 / (a_ii a_jj)
  // and t(i) = 1 / sqrt(max(1 - s_low(i), 1/64)) (for an arrow-shaped A[J,J]
  // classical FSAI is g_ii sqrt(a_ii) = t, g_ij sqrt(a_jj) = -f0 t)
  kFeatCoupled8 = 0,
  // the edge's own five quantities a_ij, a_ii, a_jj, nnz(row i), sum_j |a_ij|,
  // made scale-free: [a_ij / sqrt(a_ii a_jj), log(a_jj / a_ii), nnz(row i) / 8,
  // sum_j |a_ij| / a_ii, [j == i], 0, 0, 0] (the diagonal flag is a function of
  // the first: an SPD matrix has |a_ij| < sqrt(a_ii a_jj) off the diagonal);
  // G's rows are then normalised so that diag(G A G^T) = 1 (k_fsai_rownorm)
  kFeatStated5 = 1,
};

// Edge features of the lower pattern, 8 floats per edge (FEAT: FsaiFeatures),
// and s[e] = 1 / sqrt(a_jj) (the column scaling of G).
template <int FEAT>
__global__ __launch_bounds__(kBlock) void k_fsai_features(const uint32_t* __restrict__ row_ptr,
                                                          const uint32_t* __restrict__ col,
                                                          const float* __restrict__ val, uint32_t n,
                                                          const uint32_t* __restrict__ Lrp,
                                                          const uint32_t* __restrict__ Lcol,
                                                          const uint32_t* __restrict__ Lsrc, const float* __restrict__ diag,
                                                          float* __restrict__ feat, float* __restrict__ colscale) {
  // A row's entries are read kB at a time: every column / value load of the
  // batch, then every diagonal gather, then the arithmetic in entry order (a
  // loop over single entries waited two memory round trips per entry). The
  // lower edges of row i are its first Lrp[i+1] - Lrp[i] entries (Lsrc, Lcol
  // restate that); an edge's 8 features are two 16-byte stores.
  constexpr int kB = 8;
  (void)Lcol;
  (void)Lsrc;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float aii = diag[i];
    const uint32_t a0 = row_ptr[i], a1 = row_ptr[i + 1];
    const uint32_t o = Lrp[i], k = Lrp[i + 1] - o;
    float asum = 0.0f, s_low = 0.0f, s_all = 0.0f;
    for (uint32_t j0 = a0; j0 < a1; j0 += kB) {
      uint32_t cv[kB];
      float av[kB], dv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const bool in = j0 + u < a1;
        cv[u] = in ? col[j0 + u] : i;
        av[u] = in ? val[j0 + u] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) dv[u] = diag[cv[u]];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (j0 + u < a1) {
          const float a = av[u];
          const uint32_t c = cv[u];
          asum += fabsf(a);
          const float dc = dv[u];
          if (c != i && aii > 0.0f && dc > 0.0f) {
            const float q = a * a / (aii * dc);
            s_all += q;
            if (c < i) s_low += q;
          }
        }
      }
    }
    const float rn = (float)(a1 - a0) * 0.125f;
    const float t = 1.0f / sqrtf(fmaxf(1.0f - s_low, 1.0f / 64.0f));
    for (uint32_t t0 = 0; t0 < k; t0 += kB) {
      uint32_t cv[kB];
      float av[kB], dv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const bool in = t0 + u < k;
        cv[u] = in ? col[a0 + t0 + u] : i;
        av[u] = in ? val[a0 + t0 + u] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) dv[u] = diag[cv[u]];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (t0 + u < k) {
          const uint32_t e = o + t0 + u, c = cv[u];
          const float ajj = dv[u], aij = av[u];
          const bool ok = aii > 0.0f && ajj > 0.0f;
          f4* f = reinterpret_cast<f4*>(feat + (size_t)e * 8);
          f[0] = f4{ok ? aij * rsqrtf(aii * ajj) : 0.0f, ok ? logf(ajj / aii) : 0.0f, rn, ok ? asum / aii : 0.0f};
          if (FEAT == kFeatCoupled8) f[1] = f4{(c == i) ? 1.0f : 0.0f, s_low, s_all, t};
          else f[1] = f4{(c == i) ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f};
          colscale[e] = ajj > 0.0f ? rsqrtf(ajj) : 0.0f;
        }
      }
    }
  }
}

// Diagonal of A (duplicates summed).
__global__ __launch_bounds__(kBlock) void k_diag(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                                 const float* __restrict__ val, uint32_t n, float* __restrict__ d) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (uint32_t j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
      if (col[j] == r) s += val[j];
    d[r] = s;
  }
}

// G values from the network output: g_e = out_e * colscale_e; counts diagonal
// entries that are not positive (the solver then falls back to plain CG).
// (Diagonal edges are found from the edge's row and column: reading feature 4
// touched every line of the 32-byte-per-edge feature array.)
__global__ __launch_bounds__(kBlock) void k_fsai_scale(const float* __restrict__ out, const float* __restrict__ colscale,
                                                       const uint32_t* __restrict__ Lrow,
                                                       const uint32_t* __restrict__ Lcol, uint64_t ne,
                                                       float* __restrict__ G, uint32_t* __restrict__ nonpos) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (uint64_t)gridDim.x * blockDim.x) {
    const float g = out[e] * colscale[e];
    G[e] = g;
    if (Lrow[e] == Lcol[e] && !(g > 0.0f)) atomicAdd(nonpos, 1u);
  }
}

// Row normalisation of G (the stated-5 network): with g~ = out * colscale the
// network's row, N_i = g~_i A[J,J] g~_i^T = (G~ A G~^T)_ii, and
// G_i = g~_i / sqrt(N_i), so that diag(G A G^T) = 1 -- the scaling classical
// FSAI has by construction (g = y / sqrt(y_i) gives g A[J,J] g^T = 1). The
// network then only sets each row's direction. v_p = (A[J,J] g~_i)_p is
// formed by merging row J_p of A with J (both in ascending column order;
// duplicates of A summed), so coupled lower neighbours count too.
__device__ __forceinline__ float rownorm_v(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                           const float* __restrict__ val, const uint32_t* __restrict__ Lcol,
                                           uint32_t o, uint32_t k, uint32_t jp, const float* __restrict__ gt,
                                           float gscale) {
  // gt[o + q] * gscale = g~ of edge q
  float v = 0.0f;
  uint32_t a = row_ptr[jp];
  const uint32_t a1 = row_ptr[jp + 1];
  uint32_t q = 0;
  while (a < a1 && q < k) {
    const uint32_t c = col[a], jq = Lcol[o + q];
    if (c < jq) {
      ++a;
    } else if (c > jq) {
      ++q;
    } else {
      v += val[a] * (gt[o + q] * gscale);
      ++a;
    }
  }
  return v;
}

constexpr int kRowNormRegs = 8;  // rows of G with at most this many entries are normalised from registers

// Forward: G = s_i g~ (s_i = N_i^{-1/2}; s_i = 1 when N_i is not positive and
// finite, counted in nonpos together with non-positive diagonal entries);
// rowscale[i] = s_i. DIA: A in diagonal storage (FP32 values, `Ad`), so each
// A(J_p, J_q) is one load from its diagonal's plane instead of a walk of row
// J_p (no dependent row_ptr -> col loads; 128^3: 169 -> 137 us); the terms
// are added in the same order (J ascending), a pair off A's diagonals or
// outside its pattern adding a stored zero.
template <bool DIA = false>
__global__ __launch_bounds__(kBlock) void k_fsai_rownorm(const uint32_t* __restrict__ row_ptr,
                                                         const uint32_t* __restrict__ col,
                                                         const float* __restrict__ val, uint32_t n,
                                                         const uint32_t* __restrict__ Lrp,
                                                         const uint32_t* __restrict__ Lcol,
                                                         const float* __restrict__ out,
                                                         const float* __restrict__ colscale, float* __restrict__ G,
                                                         float* __restrict__ rowscale, uint32_t* __restrict__ nonpos,
                                                         DiaView<ValF32> Ad = {}) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t o = Lrp[i], k = Lrp[i + 1] - o;
    float N = 0.0f;
    if (k <= (uint32_t)kRowNormRegs) {
      // the row's pattern and g~ in registers; each entry of row J_p of A is
      // matched against J by compares, so its loads do not wait on one
      // another (the merge below walks both lists one dependent load at a
      // time). The same products are added in the same order.
      uint32_t J[kRowNormRegs];
      float gt[kRowNormRegs];
#pragma unroll
      for (int p = 0; p < kRowNormRegs; ++p) {
        J[p] = (uint32_t)p < k ? Lcol[o + p] : 0xFFFFFFFFu;
        gt[p] = (uint32_t)p < k ? out[o + p] * colscale[o + p] : 0.0f;
      }
      if constexpr (DIA) {
        // every A(J_p, J_q) first (all loads in flight), then the sums
        float av[kRowNormRegs][kRowNormRegs];
#pragma unroll
        for (int p = 0; p < kRowNormRegs; ++p)
#pragma unroll
          for (int q = 0; q < kRowNormRegs; ++q) {
            av[p][q] = 0.0f;
            if ((uint32_t)p < k && (uint32_t)q < k) {
              const int32_t off = (int32_t)(J[q] - J[p]);
              int64_t at = -1;
              for (int d = 0; d < Ad.o.ndiag; ++d)
                if (Ad.o.off[d] == off)
                  at = (Ad.mirrored ? (int64_t)Ad.o.plane[d] * (int64_t)Ad.ld + Ad.o.shift[d]
                                    : (int64_t)d * (int64_t)Ad.ld) + (int64_t)J[p];
              if (at >= 0) av[p][q] = Ad.val[at];
            }
          }
#pragma unroll
        for (int p = 0; p < kRowNormRegs; ++p) {
          if ((uint32_t)p >= k) continue;
          G[o + p] = gt[p];
          float v = 0.0f;
#pragma unroll
          for (int q = 0; q < kRowNormRegs; ++q)
            if ((uint32_t)q < k) v += av[p][q] * gt[q];
          N += gt[p] * v;
        }
      } else {
      uint32_t ra[kRowNormRegs], rb[kRowNormRegs];
#pragma unroll
      for (int p = 0; p < kRowNormRegs; ++p) {
        ra[p] = (uint32_t)p < k ? row_ptr[J[p]] : 0u;
        rb[p] = (uint32_t)p < k ? row_ptr[J[p] + 1] : 0u;
      }
#pragma unroll
      for (int p = 0; p < kRowNormRegs; ++p) {
        if ((uint32_t)p >= k) continue;
        G[o + p] = gt[p];
        float v = 0.0f;
        for (uint32_t a0 = ra[p]; a0 < rb[p]; a0 += kRowNormRegs) {  // entries 8 at a time
          uint32_t cv[kRowNormRegs];
          float wv[kRowNormRegs];
#pragma unroll
          for (int u = 0; u < kRowNormRegs; ++u) {
            const bool in = a0 + u < rb[p];
            cv[u] = in ? col[a0 + u] : 0xFFFFFFFFu;
            wv[u] = in ? val[a0 + u] : 0.0f;
          }
#pragma unroll
          for (int u = 0; u < kRowNormRegs; ++u) {
            bool hit = false;
            float g = 0.0f;
#pragma unroll
            for (int q = 0; q < kRowNormRegs; ++q) {
              hit |= cv[u] == J[q];
              g = cv[u] == J[q] ? gt[q] : g;
            }
            v = hit ? v + wv[u] * g : v;
          }
        }
        N += gt[p] * v;
      }
      }
    } else {
      for (uint32_t p = 0; p < k; ++p) G[o + p] = out[o + p] * colscale[o + p];  // g~, scaled below
      for (uint32_t p = 0; p < k; ++p)
        N += G[o + p] * rownorm_v(row_ptr, col, val, Lcol, o, k, Lcol[o + p], G, 1.0f);
    }
    const bool ok = N > 0.0f && N < INFINITY;
    const float sc = ok ? rsqrtf(N) : 1.0f;
    bool diag_ok = false;
    for (uint32_t p = 0; p < k; ++p) {
      const float g = G[o + p] * sc;
      G[o + p] = g;
      if (Lcol[o + p] == i && g > 0.0f) diag_ok = true;
    }
    rowscale[i] = ok ? sc : 0.0f;
    if (!ok || !diag_ok) atomicAdd(nonpos, 1u);
  }
}

// Backward of k_fsai_rownorm followed by the column scaling: from gG = dL/dG
// per edge, dL/dout_e = colscale_e * dL/dg~_e with
//   dL/dg~_p = s gG_p - s^3 v_p sum_q gG_q g~_q   (v = A[J,J] g~, g~ = G / s);
// rows whose normalisation was skipped (rowscale 0) pass gG through.
__global__ __launch_bounds__(kBlock) void k_fsai_rownorm_grad(const uint32_t* __restrict__ row_ptr,
                                                              const uint32_t* __restrict__ col,
                                                              const float* __restrict__ val, uint32_t n,
                                                              const uint32_t* __restrict__ Lrp,
                                                              const uint32_t* __restrict__ Lcol,
                                                              const float* __restrict__ G,
                                                              const float* __restrict__ rowscale,
                                                              const float* __restrict__ colscale,
                                                              const float* __restrict__ gG, float* __restrict__ gout) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t o = Lrp[i], k = Lrp[i + 1] - o;
    const float s = rowscale[i];
    if (!(s > 0.0f)) {
      for (uint32_t p = 0; p < k; ++p) gout[o + p] = gG[o + p] * colscale[o + p];
      continue;
    }
    const float inv = 1.0f / s;
    float c = 0.0f;
    for (uint32_t p = 0; p < k; ++p) c += gG[o + p] * (G[o + p] * inv);
    const float s3c = s * s * s * c;
    for (uint32_t p = 0; p < k; ++p) {
      const float v = rownorm_v(row_ptr, col, val, Lcol, o, k, Lcol[o + p], G, inv);
      gout[o + p] = colscale[o + p] * (s * gG[o + p] - s3c * v);
    }
  }
}

// Probe-loss gradient per edge of G for L = sum_w ||G A G^T w - w||^2:
// with u = G^T w, y = A u, e = G y - w, h = A G^T e:
// dL/dG_ij = 2 (e_i y_j + w_i h_j).
__global__ __launch_bounds__(kBlock) void k_fsai_edge_grad(const uint32_t* __restrict__ Lrow,
                                                           const uint32_t* __restrict__ Lcol, uint64_t ne,
                                                           const float* __restrict__ e, const float* __restrict__ y,
                                                           const float* __restrict__ w, const float* __restrict__ h,
                                                           float scale, float* __restrict__ grad) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ne; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = Lrow[k], j = Lcol[k];
    grad[k] += scale * 2.0f * (e[i] * y[j] + w[i] * h[j]);
  }
}

// e = z - w; block partials of ||e||^2, ||w||^2 (finish with k_finish<2, FinStore>)
__global__ __launch_bounds__(kBlock) void k_probe_resid(const float* __restrict__ z, const float* __restrict__ w,
                                                        uint32_t n, float* __restrict__ e, GridRed rb) {
  CSum acc[2] = {csum_zero(), csum_zero()};
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float d = z[i] - w[i];
    e[i] = d;
    acc[0] = csum_add(acc[0], d * d);
    acc[1] = csum_add(acc[1], w[i] * w[i]);
  }
  block_partials<2>(acc, rb);
}

}  // namespace msp
