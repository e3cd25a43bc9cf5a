// Krylov iterations as a short chain of fused kernels per step, each with at
// most one deterministic grid reduction (block partials in the kernel, totals
// and scalar bookkeeping in the k_finish launched after it), all scalars kept
// on the device.
//
// CG (reference: iterative/BiCGSTAB/cgd.metal:11-90, spmv.cpp:528-659) per step,
// two launches:
//   the previous step's convergence decision and beta, taken by every
//   workgroup from that step's <r,r>, <r,z> sums (FinCgStep on a private copy
//   of the state, staged in a second state block); p = z + beta p_old for
//   every gathered column (XPNew, ping-pong p buffers), q = A p, block sums of
//   <p,q>                                (k_spmv / k_dia + EpiCgQ)
//   the decision committed; alpha = <r,z> / <p,q> from those sums in every
//   workgroup, x += alpha p, r -= alpha q, sums of <r,r>, <r,z>
//                                        (k_cg_update_xr)
// and a finisher (FinCgStepPending) only at the end of a batch of steps.
// FSAI-preconditioned CG (M = G^T G, so <r,z> = <G r, G r>), three launches
// a step:
//   z = G^T w with p = z + beta p_old formed in the epilogue, beta from the
//   ||w||^2 sums in every workgroup, which also take the previous step's
//   convergence decision from its <r,r> sums (EpiPcgDir)
//   q = A p, sums of <p,q>
//   the decision committed; alpha = ||w||^2 / <p,q>, x += alpha p (every other
//   step), r_new = r - alpha q in the gather of the next w = G r_new, sums of
//   <r_new,r_new> and ||w_new||^2 (XRNew / EpiCgXrW)
// with a finisher (FinCgStepPending) only at the end of a batch of steps.
// BiCGSTAB (reference: iterative/BiCGSTAB/bicgstab.metal:12-326, spmv.cpp:699-888):
//   v = A p, <r0,v> -> alpha             (k_spmv + FinBicgAlpha)
//   s = r - alpha v formed inside the second SpMV's gather, t = A s,
//   <t,s>, <t,t> -> omega                (k_spmv + XAxpy/EpiBicgT + FinBicgOmega)
//   x += alpha p + omega s, r = s - omega t, <r0,r>, <r,r> -> beta (k_bicg_update)
//   p = r + beta (p - omega v)           (k_bicg_p)
// Every kernel returns at once when the device `done` flag is set, so the host
// may enqueue whole batches of steps and poll the flag only between batches.
#pragma once

#include "bcsr.hpp"
#include "common.hpp"
#include "spmv.hpp"

namespace msp {

enum : int32_t { kRunning = 0, kConverged = 1, kBreakdown = 2, kMaxIter = 3, kStagnated = 4 };
enum : int32_t { kCritRelres = 0, kCritReference = 1 };
enum : int32_t { kPcNone = 0, kPcJacobi = 1, kPcExplicit = 2 };

struct Scalars {
  float rz;      // <r, z>, z = M r (z = r without preconditioner)
  float alpha;
  float beta;
  float pq;      // <p, A p>
  float rho;     // BiCGSTAB <r0, r>
  float sigma;   // BiCGSTAB <r0, v>
  float omega;
  float rr;      // ||r||^2 of the latest iterate
  float bb;      // ||b||^2
  float thresh;  // converged when the true ||r||^2 <= thresh
  float thresh_rec;  // recurrence target (tightened after a failed verification)
  float tol;
  float rr0;     // ||r||^2 at start
  float rr_best;  // smallest ||b - A x||^2 seen by a verification
  int32_t done;
  int32_t iter;
  int32_t maxiter;
  int32_t crit;
  int32_t use_verify;  // confirm recurrence convergence with a true residual
  int32_t verify;      // a verification is pending
  int32_t nverify;
  int32_t nfail;     // consecutive verifications without a 10% gain in ||r||
  int32_t collapse;  // fold xl into x at the next update (after a restart)
  int32_t restart;   // next preconditioned direction update uses beta = 0
  int32_t halt;      // done || verify: the step kernels do nothing (kept by k_finish)
  int32_t skip_g;    // explicit preconditioner: w = G r of the current residual is up to date
  int32_t pending;   // explicit preconditioner: the latest step's <r,r> sums await FinCgStep
  // CG: the iterate's update of the latest step is held (x += alpha_hold p,
  // p in direction buffer xpend_buf) and applied with the next step's, or by
  // k_cg_xflush before a verification and at the end of the solve
  int32_t xpend;
  int32_t xpend_buf;
  float alpha_hold;
};

// Solver state at the start of a solve: every field zero, then the settings.
__global__ void k_start_solve(Scalars* S, float tol, int32_t maxiter, int32_t crit, int32_t use_verify) {
  Scalars z{};
  z.tol = tol;
  z.maxiter = maxiter;
  z.crit = crit;
  z.use_verify = use_verify;
  *S = z;
}

// k_finish hook: after a finisher that owns the solver state, refresh the
// step kernels' gate.
__device__ __forceinline__ void refresh_halt(Scalars* S) { S->halt = (S->done != 0 || S->verify != 0) ? 1 : 0; }

// Recurrence says converged: either stop, or ask for a true-residual check.
__device__ __forceinline__ void on_recurrence_converged(Scalars* S) {
  if (S->use_verify && S->crit == kCritRelres) S->verify = 1;
  else S->done = kConverged;
}

__device__ __forceinline__ bool finite_pos(float v) { return v > 0.0f && isfinite(v); }

// Finishers (run in one thread after the grid reduction) --------------------
struct FinBicgAlpha {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return S->halt == 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[1]) const {
    S->sigma = tot[0];
    if (tot[0] == 0.0f || !isfinite(tot[0])) {
      S->done = kBreakdown;
      S->alpha = 0.0f;
    } else {
      S->alpha = S->rho / tot[0];
    }
  }
};

struct FinBicgOmega {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return S->halt == 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    S->omega = (tot[1] != 0.0f) ? tot[0] / tot[1] : 0.0f;
  }
};



__global__ void k_set_restart(Scalars* S) { S->restart = 1; }

// Where a step kernel takes the previous CG step's pending decision: the
// state S (read by every workgroup), the staging block S2 (written by
// workgroup 0, committed to S by the next kernel) and the <r,r>, <r,z> block
// partials of that step's x / r update (rr, nrr blocks, 2 values each).
struct CgDecide {
  const Scalars* S;
  Scalars* S2;
  GridRed rr;
  uint32_t nrr;
};

// CG direction formed on the fly: p = z + beta p_old, z = r (PC none) or
// D^-1 r (Jacobi); beta (0 on the first step and after a restart) from the
// decision the kernel's prologue takes (dec).
template <int PC>
struct XPNew {
  const float* __restrict__ r;
  const float* __restrict__ pold;
  const float* __restrict__ dinv;
  CgDecide dec;
  float beta;
  struct Raw {
    float r, p, d;
  };
  __device__ __forceinline__ Raw raw(uint32_t c) const {
    return Raw{r[c], pold[c], PC == kPcJacobi ? dinv[c] : 1.0f};
  }
  __device__ __forceinline__ float make(Raw w) const {
    float z = w.r;
    if (PC == kPcJacobi) z *= w.d;
    return z + beta * w.p;
  }
  __device__ __forceinline__ float operator()(uint32_t c) const {
    float z = r[c];
    if (PC == kPcJacobi) z *= dinv[c];
    return z + beta * pold[c];
  }
  __device__ __forceinline__ f4 load4(u4 c) const {
    if (quad_run(c)) {
      f4 z = *reinterpret_cast<const f4*>(r + c.x);
      if (PC == kPcJacobi) z *= *reinterpret_cast<const f4*>(dinv + c.x);
      return z + beta * *reinterpret_cast<const f4*>(pold + c.x);
    }
    return f4{(*this)(c.x), (*this)(c.y), (*this)(c.z), (*this)(c.w)};
  }
};

// The same direction for the BCSR kernel's 4-wide loads.
template <int PC>
struct X4PNew {
  const float* __restrict__ r;
  const float* __restrict__ pold;
  const float* __restrict__ dinv;
  CgDecide dec;
  float beta;
  __device__ __forceinline__ float operator()(uint32_t c) const {
    float z = r[c];
    if (PC == kPcJacobi) z *= dinv[c];
    return z + beta * pold[c];
  }
  __device__ __forceinline__ f4 load4(uint32_t c) const {
    f4 z = *reinterpret_cast<const f4*>(r + c);
    if (PC == kPcJacobi) z *= *reinterpret_cast<const f4*>(dinv + c);
    return z + beta * *reinterpret_cast<const f4*>(pold + c);
  }
};

// q = A p with p formed by XPNew / X4PNew: the owned rows also store p, and <p, q>.
template <class XL>
struct EpiCgQ {
  float* __restrict__ q;
  float* __restrict__ pnew;
  XL xl;
  typedef typename RawOf<XL>::type Pre;  // operands prefetched by the diagonal kernel
  static constexpr bool kOwnX = true;
  __device__ __forceinline__ Pre pre(uint32_t i) const { return xl.raw(i); }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum* acc, Pre w) const {
    const float pi = xl.make(w);
    pnew[i] = pi;
    q[i] = s;
    acc[0] = csum_add(acc[0], pi * s);
  }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum* acc) const {
    const float pi = xl(i);
    pnew[i] = pi;
    q[i] = s;
    acc[0] = csum_add(acc[0], pi * s);
  }
};

// Convergence test shared by the solvers; `rr_old` is ||r||^2 at the start of
// the step (the reference criterion tests it every 50 steps,
// iterative/BiCGSTAB/spmv.cpp:603-629).
__device__ __forceinline__ bool step_converged(const Scalars* S, int it, float rr_old, float rr_new) {
  if (S->crit == kCritReference) return (it % 50 == 0) && (rr_old < S->tol);
  return rr_new <= S->thresh_rec;
}

__device__ __forceinline__ uint32_t vec_chunks(uint32_t n) { return (n + 3u) / 4u; }


// Initialisation: x = 0 (or kept), r = b (or precomputed b - A x0), p = z
// (p = 0 for an explicit preconditioner, whose first direction update takes
// beta = 0 times this p), <b,b>, <r,r>, <r,z>.
__global__ __launch_bounds__(kBlock) void k_cg_init(float* __restrict__ x, float* __restrict__ xl,
                                                    float* __restrict__ r, float* __restrict__ p,
                                                    const float* __restrict__ b, const float* __restrict__ dinv,
                                                    uint32_t n, int use_x0, int pc, Scalars* S, GridRed rb) {
  CSum acc[3] = {csum_zero(), csum_zero(), csum_zero()};
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 bv = reinterpret_cast<const f4*>(b)[c];
    f4 rv = use_x0 ? reinterpret_cast<const f4*>(r)[c] : bv;
    f4 zv = rv;
    if (pc == kPcJacobi) zv = rv * reinterpret_cast<const f4*>(dinv)[c];
    if (!use_x0) {
      reinterpret_cast<f4*>(x)[c] = f4{0.f, 0.f, 0.f, 0.f};
      reinterpret_cast<f4*>(r)[c] = rv;
    }
    reinterpret_cast<f4*>(xl)[c] = f4{0.f, 0.f, 0.f, 0.f};
    reinterpret_cast<f4*>(p)[c] = pc != kPcExplicit ? zv : f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[0] = csum_add(acc[0], bv[k] * bv[k]);
      acc[1] = csum_add(acc[1], rv[k] * rv[k]);
      acc[2] = csum_add(acc[2], rv[k] * zv[k]);
    }
  }
  block_partials<3>(acc, rb);
}

// Totals of k_cg_init: <b,b>, <r,r>, <r,z>.
struct FinCgInit {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return true; }
  __device__ __forceinline__ void operator()(const float (&tot)[3]) const {
    S->bb = tot[0];
    S->rr = tot[1];
    S->rr0 = tot[1];
    S->rz = tot[2];
    S->rho = tot[1];
    S->thresh = S->tol * S->tol * tot[0];
    S->thresh_rec = S->thresh;
    S->iter = 0;
    S->beta = 0.0f;
    S->alpha = 0.0f;
    S->omega = 0.0f;
    S->skip_g = 1;
    S->pending = 0;
    int32_t d = kRunning;
    if (S->crit == kCritRelres && tot[1] <= S->thresh) d = kConverged;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) d = kBreakdown;
    if (S->maxiter <= 0 && d == kRunning) d = kMaxIter;
    S->done = d;
  }
};

// Total of nb block partials (value d of ND) computed by the whole workgroup
// in a fixed order (every workgroup gets the same bits); all threads get it.
template <int ND>
__device__ __forceinline__ float block_total(const CSum* __restrict__ partials, uint32_t nb, int d) {
  __shared__ float s_tot;
  CSum a[1] = {csum_zero()};
  merge_partials<1>(partials + d, nb, a, ND);
  block_reduce<1>(a);
  if (threadIdx.x == 0) s_tot = csum_val(a[0]);
  __syncthreads();
  const float t = s_tot;
  __syncthreads();
  return t;
}

// block_total of two partial sets (pa: na blocks, one value taken every sa
// CSums; pb: nb blocks, every sb) in one pass: the loads of both are issued
// before the first merge. Each total has the bits block_total gives it (and
// k_finish gives that value).
__device__ __forceinline__ void block_total2(const CSum* __restrict__ pa, uint32_t na, const CSum* __restrict__ pb,
                                             uint32_t nb, float& ta, float& tb, uint32_t sb = 1, uint32_t sa = 1) {
  __shared__ float s_tot[2];
  CSum a[2] = {csum_zero(), csum_zero()};
  const uint32_t nmax = na > nb ? na : nb;
  for (uint32_t base = 0; base < nmax; base += kPartialBatch * kBlock) {
    CSum va[kPartialBatch], vb[kPartialBatch];
#pragma unroll
    for (int j = 0; j < kPartialBatch; ++j) {
      const uint32_t i = base + j * kBlock + threadIdx.x;
      va[j] = i < na ? pa[(size_t)i * sa] : csum_zero();
      vb[j] = i < nb ? pb[(size_t)i * sb] : csum_zero();
    }
#pragma unroll
    for (int j = 0; j < kPartialBatch; ++j) {
      const uint32_t i = base + j * kBlock + threadIdx.x;
      a[0] = i < na ? csum_merge(a[0], va[j]) : a[0];
      a[1] = i < nb ? csum_merge(a[1], vb[j]) : a[1];
    }
  }
  block_reduce<2>(a);
  if (threadIdx.x == 0) {
    s_tot[0] = csum_val(a[0]);
    s_tot[1] = csum_val(a[1]);
  }
  __syncthreads();
  ta = s_tot[0];
  tb = s_tot[1];
  __syncthreads();
}

// The first kernel of an FSAI-preconditioned CG step, z = G^T w (w = G r
// from the previous step's update), with the new direction formed in its
// epilogue: p = z + beta p_old, written to the other direction buffer (p_old
// must not alias it). With M = G^T G, <r, z> = <G r, G r> = ||w||^2, which the
// kernel that produced w summed beside ||r||^2 (rr: nrr blocks of 2 partials);
// every workgroup reduces both in kernel_prologue, takes the previous step's
// pending decision on ||r||^2 in place of a finisher launch (FinCgStep on a
// private copy of the state, so all agree on the gate) and forms
// beta = ||w||^2 / <r,z>_old (0 on the first step, after a restart, or for a
// non-positive ||w||^2; the step's update then records the breakdown).
// Workgroup 0 stages the decided fields in S2, also when halted (every
// workgroup of this kernel reads S, so a later kernel commits S2 to it), and
// stages ||w||^2 as the step's <r,z> in S2->rz for alpha.
struct EpiPcgDir {
  float* __restrict__ pnew;
  const float* __restrict__ pold;
  const Scalars* S;
  Scalars* S2;
  GridRed rr;
  uint32_t nrr;
  float beta;
  typedef float Pre;
  __device__ __forceinline__ Pre pre(uint32_t i) const { return pold[i]; }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum*, Pre po) const { pnew[i] = s + beta * po; }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum* acc) const { row(i, s, acc, pre(i)); }
};

template <int PC>
__device__ __forceinline__ int take_pending_step(const Scalars* S, Scalars* S2, const float (&tot)[2], int halt,
                                                 float& beta);
__device__ __forceinline__ void commit_step_state(Scalars* S, const Scalars* S2);

__device__ __forceinline__ int kernel_prologue(XPlain&, EpiPcgDir& epi, int halt) {
  float rr, ww;
  block_total2(epi.rr.partials, epi.nrr, epi.rr.partials + 1, epi.nrr, rr, ww, 2, 2);
  const float tot[2] = {rr, 0.0f};
  float unused;
  halt = take_pending_step<kPcExplicit>(epi.S, epi.S2, tot, halt, unused);
  const Scalars* S = epi.S;
  epi.beta = (S->restart || S->rz == 0.0f || !finite_pos(ww)) ? 0.0f : ww / S->rz;
  if (blockIdx.x == 0 && threadIdx.x == 0) epi.S2->rz = ww;
  return halt;
}

// ||r||^2 and ||w||^2 as 2-value block partials of the layout the FSAI-PCG
// update writes, for a w = G r computed outside a step (the start of a solve,
// after a restart); gated like that product (skipped when *gate != 0).
__global__ __launch_bounds__(kBlock) void k_pcg_wnorm(const float* __restrict__ r, const float* __restrict__ w,
                                                      uint32_t n, GridRed rb, const int* __restrict__ gate) {
  if (*gate) return;
  CSum acc[2] = {csum_zero(), csum_zero()};
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float ri = r[i], wi = w[i];
    acc[0] = csum_add(acc[0], ri * ri);
    acc[1] = csum_add(acc[1], wi * wi);
  }
  block_partials<2>(acc, rb);
}

// The x / r update of an explicitly preconditioned CG step fused with the
// next step's first factor product w = G r_new: r_new = r - alpha q is formed
// in the gather (r_new goes to the other residual buffer, so every workgroup
// gathers the step's r), alpha = <r,z> / <p,q> with <r,z> staged in S2 by the
// step's first kernel (EpiPcgDir) and <p,q> from the block partials of
// q = A p (pq, npq blocks), reduced by every workgroup in kernel_prologue;
// alpha = 0 when either is not positive (the step's decision then records a
// breakdown). The kernel's gate is S2->halt (the state as decided by
// EpiPcgDir's prologue); workgroup 0 commits S2 to S, also when halted, then
// records pq, alpha, <r,z>, clears the restart flag, marks w current and the
// step's decision pending.
// The epilogue updates the owned rows (x, x_lo when COMP, r_new, w) and
// accumulates <r_new, r_new> and <w, w> (the next step's <r,z>).
struct XRNew {
  const float* __restrict__ r;
  const float* __restrict__ q;
  float alpha;
  struct Raw {
    float r, q;
  };
  __device__ __forceinline__ Raw raw(uint32_t c) const { return Raw{r[c], q[c]}; }
  // raw with 32-bit byte offsets (c < 2^30)
  __device__ __forceinline__ Raw raw32(uint32_t c) const { return Raw{ld_off32(r, c), ld_off32(q, c)}; }
  __device__ __forceinline__ float make(Raw w) const { return w.r - alpha * w.q; }
  __device__ __forceinline__ float operator()(uint32_t c) const { return r[c] - alpha * q[c]; }
  __device__ __forceinline__ f4 load4(u4 c) const {
    if (quad_run(c))
      return *reinterpret_cast<const f4*>(r + c.x) - alpha * *reinterpret_cast<const f4*>(q + c.x);
    return f4{(*this)(c.x), (*this)(c.y), (*this)(c.z), (*this)(c.w)};
  }
};

template <bool COMP>
// The iterate moves every other step, as in k_cg_update_xr: with a held
// update from the previous step (pend; direction pold, alpha_hold) the rows
// apply it and then this step's; otherwise the step holds its own (applyx
// false: x and xl neither loaded nor stored) unless it must fold xl into x.
struct EpiCgXrW {
  float* __restrict__ x;
  float* __restrict__ xl;
  const float* __restrict__ rold;
  float* __restrict__ rnew;
  const float* __restrict__ p;
  const float* __restrict__ q;
  float* __restrict__ w;
  Scalars* S;
  const Scalars* S2;
  GridRed pq;
  uint32_t npq;
  float alpha;
  int fold;
  const float* __restrict__ pold;  // the previous step's direction
  int32_t pbuf;                    // direction buffer index of p (recorded when holding)
  float ah;                        // alpha_hold of a pending update
  int pend, applyx;
  struct Pre {
    float x, xl, p, ph, r, q;
  };
  __device__ __forceinline__ Pre pre(uint32_t i) const {
    Pre v{0.0f, 0.0f, 0.0f, 0.0f, rold[i], q[i]};
    if (applyx) {
      v.x = x[i];
      if (COMP) v.xl = xl[i];
      v.p = p[i];
      if (pend) v.ph = pold[i];
    }
    return v;
  }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum* acc, Pre v) const {
    if (applyx) {
      float xi = v.x;
      if constexpr (COMP) {
        float li = v.xl;
        if (fold) {
          xi += li;
          li = 0.0f;
        }
        if (pend) comp_axpy(xi, li, ah, v.ph);
        comp_axpy(xi, li, alpha, v.p);
        xl[i] = li;
      } else {
        if (pend) xi += ah * v.ph;
        xi += alpha * v.p;
      }
      x[i] = xi;
    }
    const float ri = v.r - alpha * v.q;
    rnew[i] = ri;
    w[i] = s;
    acc[0] = csum_add(acc[0], ri * ri);
    acc[1] = csum_add(acc[1], s * s);
  }
  __device__ __forceinline__ void row(uint32_t i, float s, CSum* acc) const { row(i, s, acc, pre(i)); }
};

template <bool COMP>
__device__ __forceinline__ int kernel_prologue(XRNew& xl, EpiCgXrW<COMP>& epi, int halt) {
  const float pqv = block_total<1>(epi.pq.partials, epi.npq, 0), rz = epi.S2->rz;
  Scalars* S = epi.S;
  const bool ok = finite_pos(pqv) && finite_pos(rz);
  const float alpha = ok ? rz / pqv : 0.0f;
  xl.alpha = alpha;
  epi.alpha = alpha;
  epi.fold = epi.S2->collapse != 0;
  epi.pend = epi.S2->xpend != 0;
  // (S->alpha_hold is written below only by a step that holds, which reads none)
  epi.ah = epi.pend ? S->alpha_hold : 0.0f;
  epi.applyx = epi.fold || epi.pend;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    commit_step_state(S, epi.S2);
    if (!halt) {
      S->pq = ok ? pqv : -1.0f;
      S->alpha = alpha;
      S->rz = rz;
      S->restart = 0;
      S->skip_g = 1;
      S->pending = 1;
      S->xpend = epi.applyx ? 0 : 1;
      if (!epi.applyx) {
        S->alpha_hold = alpha;
        S->xpend_buf = epi.pbuf;
      }
    }
  }
  return halt;
}

// x += alpha p; r -= alpha q; <r,r>; <r, dinv r> (Jacobi) -> the next step's
// decision (pending).
// alpha = <r,z> / <p,q> with <p,q> from the SpMV's block partials (pq, npq
// blocks), computed identically by every workgroup; <r,z> and the gate come
// from S2, the state as decided by the step's first kernel, which block 0
// commits to S (also when halted) before it records pq and alpha; a
// non-positive <p,q> leaves x and r alone (the decision then stops with a
// breakdown). COMP: the iterate is x + xl with xl carrying the rounding
// error of x += alpha p (for targets near FP32's attainable accuracy);
// otherwise x alone, xl untouched. (The explicitly preconditioned step fuses
// this update into its factor product: XRNew / EpiCgXrW.)
// The iterate moves every other step: a step whose predecessor left its
// update held (S2->xpend; its direction is this step's pold, its alpha
// S->alpha_hold) applies that update and then its own, in that order, so x
// takes the same values; otherwise the step holds its own (records alpha,
// the buffer index pbuf of p) and touches neither x nor xl. A step that must
// fold xl into x (after a restart) applies at once. The recurrence never
// reads x, so r and every decision are unchanged; x and xl move through
// memory once per two steps (compensated: 32 -> 24 B per row and step).
template <int PC, bool COMP>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(float* __restrict__ x, float* __restrict__ xl,
                                                         float* __restrict__ r, const float* __restrict__ p,
                                                         const float* __restrict__ pold,
                                                         const float* __restrict__ q, const float* __restrict__ dinv,
                                                         uint32_t n, Scalars* S, const Scalars* S2, GridRed rb,
                                                         GridRed pq, uint32_t npq, int32_t pbuf) {
  // the gate is tested after the partials are in (their loads overlap its
  // latency); a halted step writes nothing but the commit
  const int32_t halt = S2->halt;
  const float pqv = block_total<1>(pq.partials, npq, 0), rz = S2->rz;
  const bool fold = S2->collapse != 0, pend = S2->xpend != 0;
  // (S->alpha_hold is written below only by a step that holds, which reads none)
  const float ah = pend ? S->alpha_hold : 0.0f;
  const bool applyx = fold || pend;
  if (blockIdx.x == 0 && threadIdx.x == 0) commit_step_state(S, S2);
  if (halt) return;
  const bool ok = finite_pos(pqv);
  const float alpha = ok ? rz / pqv : 0.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->pq = ok ? pqv : -1.0f;
    S->alpha = alpha;
    S->pending = 1;
    S->xpend = applyx ? 0 : 1;
    if (!applyx) {
      S->alpha_hold = alpha;
      S->xpend_buf = pbuf;
    }
  }
  CSum acc[2] = {csum_zero(), csum_zero()};
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 rv = reinterpret_cast<f4*>(r)[c];
    f4 qv = reinterpret_cast<const f4*>(q)[c];
    if (applyx) {
      f4 xv = reinterpret_cast<f4*>(x)[c];
      f4 pv = reinterpret_cast<const f4*>(p)[c];
      f4 hv = f4{0.f, 0.f, 0.f, 0.f};
      if (pend) hv = reinterpret_cast<const f4*>(pold)[c];
      if constexpr (COMP) {
        f4 xlv = reinterpret_cast<f4*>(xl)[c];
        if (fold) {
          xv += xlv;
          xlv = f4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float xk = xv[k], lk = xlv[k];
          if (pend) comp_axpy(xk, lk, ah, hv[k]);
          comp_axpy(xk, lk, alpha, pv[k]);
          xv[k] = xk;
          xlv[k] = lk;
        }
        reinterpret_cast<f4*>(xl)[c] = xlv;
      } else {
        if (pend) xv += ah * hv;
        xv += alpha * pv;
      }
      reinterpret_cast<f4*>(x)[c] = xv;
    }
    rv -= alpha * qv;
    reinterpret_cast<f4*>(r)[c] = rv;
    f4 zv = rv;
    if (PC == kPcJacobi) zv = rv * reinterpret_cast<const f4*>(dinv)[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[0] = csum_add(acc[0], rv[k] * rv[k]);
      acc[1] = csum_add(acc[1], rv[k] * zv[k]);
    }
  }
  block_partials<2>(acc, rb);
}

// A held update of the iterate applied ahead of a verification (VER: only when
// one is about to run: verify set, not done) or at the end of a solve; the
// flag is cleared by k_cg_xflush_clear once every workgroup has read it.
template <bool COMP, bool VER>
__global__ __launch_bounds__(kBlock) void k_cg_xflush(float* __restrict__ x, float* __restrict__ xl,
                                                      const float* __restrict__ p0, const float* __restrict__ p1,
                                                      uint32_t n, const Scalars* S) {
  if (!S->xpend || (VER && (S->verify == 0 || S->done != 0))) return;
  const float a = S->alpha_hold;
  const float* __restrict__ p = S->xpend_buf ? p1 : p0;
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 xv = reinterpret_cast<f4*>(x)[c];
    const f4 pv = reinterpret_cast<const f4*>(p)[c];
    if constexpr (COMP) {
      f4 xlv = reinterpret_cast<f4*>(xl)[c];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float xk = xv[k], lk = xlv[k];
        comp_axpy(xk, lk, a, pv[k]);
        xv[k] = xk;
        xlv[k] = lk;
      }
      reinterpret_cast<f4*>(xl)[c] = xlv;
    } else {
      xv += a * pv;
    }
    reinterpret_cast<f4*>(x)[c] = xv;
  }
}
template <bool VER>
__global__ void k_cg_xflush_clear(Scalars* S) {
  if (!S->xpend || (VER && (S->verify == 0 || S->done != 0))) return;
  S->xpend = 0;
}

// Totals of k_cg_update_xr: <r,r> and <r,z> -> iteration count, convergence
// (or a pending verification), maxiter, beta.
template <int PC>
struct FinCgStep {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return S->halt == 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    if (!finite_pos(S->pq)) {  // alpha was not applied
      S->done = kBreakdown;
      return;
    }
    const int it = S->iter;
    const float rr_old = S->rr;
    S->iter = it + 1;
    S->rr = tot[0];
    S->collapse = 0;
    const bool conv = step_converged(S, it, rr_old, tot[0]);
    if (!isfinite(tot[0])) {
      S->done = kBreakdown;
      return;
    }
    if (conv) on_recurrence_converged(S);
    if (S->done) return;
    if (!S->verify && it + 1 >= S->maxiter) {
      S->done = kMaxIter;
    } else if (PC != kPcExplicit) {
      if (S->rz == 0.0f) {
        if (!S->verify) S->done = kBreakdown;
      } else {
        S->beta = tot[1] / S->rz;
        S->rz = tot[1];
      }
    }
  }
};

// The fields FinCgStep changes, staged in S2 by the kernel that takes the
// decision and copied to S by the next kernel (one thread).
__device__ __forceinline__ void stage_step_state(Scalars* S2, const Scalars& L) {
  S2->iter = L.iter;
  S2->rr = L.rr;
  S2->collapse = L.collapse;
  S2->done = L.done;
  S2->verify = L.verify;
  S2->halt = L.halt;
  S2->pending = L.pending;
  S2->beta = L.beta;
  S2->rz = L.rz;
  S2->xpend = L.xpend;
}
__device__ __forceinline__ void commit_step_state(Scalars* S, const Scalars* S2) {
  const Scalars L = *S2;
  stage_step_state(S, L);
}

// The decision of a CG step (FinCgStep on the <r,r>, <r,z> sums of its x / r
// update) for a step whose decision is still pending; run at the end of a
// batch of steps, before the host reads the state.
template <int PC>
struct FinCgStepPending {
  Scalars* S;
  __device__ __forceinline__ bool active() const { return S->pending != 0 && S->halt == 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    FinCgStep<PC>{S}(tot);
    S->pending = 0;
  }
  __device__ __forceinline__ void after() const { refresh_halt(S); }
};

// The previous step's pending decision, taken by every workgroup of a step
// kernel on a private copy of the state; workgroup 0 stages the result in S2,
// also when halted. Returns the gate; beta is the direction update's.
template <int PC>
__device__ __forceinline__ int take_pending_step(const Scalars* S, Scalars* S2, const float (&tot)[2], int halt,
                                                 float& beta) {
  Scalars L = *S;
  if (!halt && L.pending) {
    FinCgStepPending<PC> f{&L};
    f(tot);
    f.after();
    halt = L.halt;
  }
  beta = L.beta;
  if (blockIdx.x == 0 && threadIdx.x == 0) stage_step_state(S2, L);
  return halt;
}

// The same for a plain or Jacobi CG step: both update sums in one pass.
template <int PC>
__device__ __forceinline__ int take_cg_decision(const CgDecide& d, int halt, float& beta) {
  float tot[2];
  block_total2(d.rr.partials, d.nrr, d.rr.partials + 1, d.nrr, tot[0], tot[1], 2, 2);
  return take_pending_step<PC>(d.S, d.S2, tot, halt, beta);
}

template <int PC>
__device__ __forceinline__ int kernel_prologue(XPNew<PC>& xl, EpiCgQ<XPNew<PC>>& epi, int halt) {
  float beta;
  halt = take_cg_decision<PC>(xl.dec, halt, beta);
  xl.beta = beta;
  epi.xl.beta = beta;
  return halt;
}
template <int PC>
__device__ __forceinline__ int kernel_prologue(X4PNew<PC>& xl, EpiCgQ<X4PNew<PC>>& epi, int halt) {
  float beta;
  halt = take_cg_decision<PC>(xl.dec, halt, beta);
  xl.beta = beta;
  epi.xl.beta = beta;
  return halt;
}

// pnew = z + beta pold with z = r or dinv*r (pnew may alias pold), after the
// previous step's pending decision and beta (take_cg_decision; staged in
// dec.S2, whose halt gates the SpMV that follows, committed by the update).
template <int PC>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(float* pnew, const float* pold, const float* __restrict__ r,
                                                        const float* __restrict__ dinv, uint32_t n, CgDecide dec) {
  float beta;
  const int halt = take_cg_decision<PC>(dec, dec.S->halt, beta);
  if (halt) return;
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 zv = reinterpret_cast<const f4*>(r)[c];
    if (PC == kPcJacobi) zv *= reinterpret_cast<const f4*>(dinv)[c];
    reinterpret_cast<f4*>(pnew)[c] = zv + beta * reinterpret_cast<const f4*>(pold)[c];
  }
}

// BiCGSTAB init: r0 = r, p = r (r = b, or b - A x0 precomputed), x = 0 unless kept.
__global__ __launch_bounds__(kBlock) void k_bicg_init(float* __restrict__ x, float* __restrict__ xl,
                                                      float* __restrict__ r, float* __restrict__ r0,
                                                      float* __restrict__ p, const float* __restrict__ b, uint32_t n,
                                                      int use_x0, Scalars* S, GridRed rb) {
  CSum acc[2] = {csum_zero(), csum_zero()};
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 bv = reinterpret_cast<const f4*>(b)[c];
    f4 rv = use_x0 ? reinterpret_cast<const f4*>(r)[c] : bv;
    if (!use_x0) {
      reinterpret_cast<f4*>(x)[c] = f4{0.f, 0.f, 0.f, 0.f};
      reinterpret_cast<f4*>(r)[c] = rv;
    }
    reinterpret_cast<f4*>(r0)[c] = rv;
    reinterpret_cast<f4*>(p)[c] = rv;
    reinterpret_cast<f4*>(xl)[c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[0] = csum_add(acc[0], bv[k] * bv[k]);
      acc[1] = csum_add(acc[1], rv[k] * rv[k]);
    }
  }
  block_partials<2>(acc, rb);
}

// Totals of k_bicg_init: <b,b>, <r,r>.
struct FinBicgInit {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return true; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    S->bb = tot[0];
    S->rr = tot[1];
    S->rr0 = tot[1];
    S->rho = tot[1];
    S->thresh = S->tol * S->tol * tot[0];
    S->thresh_rec = S->thresh;
    S->iter = 0;
    S->alpha = 0.0f;
    S->omega = 0.0f;
    S->beta = 0.0f;
    int32_t d = kRunning;
    if (S->crit == kCritRelres && tot[1] <= S->thresh) d = kConverged;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) d = kBreakdown;
    if (tot[1] == 0.0f) d = kConverged;
    if (S->maxiter <= 0 && d == kRunning) d = kMaxIter;
    S->done = d;
  }
};

// x += alpha p + omega s; r = s - omega t; <r0, r>, <r, r> -> bookkeeping, beta.
__global__ __launch_bounds__(kBlock) void k_bicg_update(float* __restrict__ x, float* __restrict__ xl,
                                                        float* __restrict__ r, const float* __restrict__ p,
                                                        const float* __restrict__ s, const float* __restrict__ t,
                                                        const float* __restrict__ r0, uint32_t n, Scalars* S,
                                                        GridRed rb) {
  if (S->halt) return;
  const float alpha = S->alpha, omega = S->omega;
  const bool fold = S->collapse != 0;
  CSum acc[2] = {csum_zero(), csum_zero()};
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 xv = reinterpret_cast<f4*>(x)[c];
    f4 xlv = reinterpret_cast<f4*>(xl)[c];
    if (fold) {
      xv += xlv;
      xlv = f4{0.f, 0.f, 0.f, 0.f};
    }
    f4 pv = reinterpret_cast<const f4*>(p)[c];
    f4 sv = reinterpret_cast<const f4*>(s)[c];
    f4 tv = reinterpret_cast<const f4*>(t)[c];
    f4 r0v = reinterpret_cast<const f4*>(r0)[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float xk = xv[k], lk = xlv[k];
      comp_axpy(xk, lk, alpha, pv[k]);
      comp_axpy(xk, lk, omega, sv[k]);
      xv[k] = xk;
      xlv[k] = lk;
    }
    f4 rv = sv - omega * tv;
    reinterpret_cast<f4*>(x)[c] = xv;
    reinterpret_cast<f4*>(xl)[c] = xlv;
    reinterpret_cast<f4*>(r)[c] = rv;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[0] = csum_add(acc[0], r0v[k] * rv[k]);
      acc[1] = csum_add(acc[1], rv[k] * rv[k]);
    }
  }
  block_partials<2>(acc, rb);
}

// Totals of k_bicg_update: <r0,r>, <r,r> -> bookkeeping, beta.
struct FinBicgStep {
  Scalars* S;
  __device__ __forceinline__ void after() const { refresh_halt(S); }
  __device__ __forceinline__ bool active() const { return S->halt == 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    const float alpha = S->alpha, omega = S->omega;
    const int it = S->iter;
    const float rr_old = S->rr;
    const float rho_new = tot[0];
    S->iter = it + 1;
    S->rr = tot[1];
    S->collapse = 0;
    if (!isfinite(tot[1])) {
      S->done = kBreakdown;
      return;
    }
    if (step_converged(S, it, rr_old, tot[1])) on_recurrence_converged(S);
    if (S->done) return;
    if (!S->verify && it + 1 >= S->maxiter) {
      S->done = kMaxIter;
    } else if (rho_new == 0.0f || !isfinite(rho_new) || !isfinite(omega) || S->rho == 0.0f) {
      if (!S->verify) S->done = kBreakdown;
      else S->beta = 0.0f;
    } else {
      S->beta = (omega != 0.0f) ? (rho_new / S->rho) * (alpha / omega) : 0.0f;
      S->rho = rho_new;
    }
  }
};

// p = r + beta (p - omega v)
__global__ __launch_bounds__(kBlock) void k_bicg_p(float* __restrict__ p, const float* __restrict__ r,
                                                   const float* __restrict__ v, uint32_t n, const Scalars* S) {
  if (S->halt) return;
  const float beta = S->beta, omega = S->omega;
  const uint32_t nc = vec_chunks(n);
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    f4 pv = reinterpret_cast<f4*>(p)[c];
    f4 rv = reinterpret_cast<const f4*>(r)[c];
    f4 vv = reinterpret_cast<const f4*>(v)[c];
    reinterpret_cast<f4*>(p)[c] = rv + beta * (pv - omega * vv);
  }
}

// Row access for row-per-thread kernels (true residual, diagonal).
struct RowsCsr {
  const uint32_t* __restrict__ row_ptr;
  const uint32_t* __restrict__ col;
  const float* __restrict__ val;
  template <typename F>
  __device__ __forceinline__ void foreach(uint32_t i, F&& f) const {
    for (uint32_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j) f(val[j], col[j]);
  }
};

// True-residual check after the recurrence reports convergence:
// r = b - A x for the FP32 iterate x = fl(x_hi + x_lo) that the solver returns,
// with error-free products (FMA) and TwoSum row accumulation (the residual is
// exact to about twice FP32 precision). On a restart the next update folds
// x_lo into x_hi so that x and the restarted residual describe the same point.
// Confirms convergence, or restarts the recurrence from the true residual
// (p = z; BiCGSTAB also r0 = r) -- keeping the old direction after replacing a
// residual that drifted by more than its own size breaks conjugacy. Stops as
// stagnated after two verifications in a row that do not improve the best true
// residual norm by 10% (the FP32 iterate has reached its attainable accuracy),
// or after 12 verifications.
// MODE 0: CG, 1: CG + Jacobi (z = dinv r), 2: BiCGSTAB, 3: PCG with an
// explicit preconditioner applied at the start of the next step (p is left
// alone; the restart flag makes that step use beta = 0).
template <int MODE, typename Rows>
__global__ __launch_bounds__(kBlock) void k_verify(Rows rows, const float* __restrict__ x,
                                                   const float* __restrict__ xl, const float* __restrict__ b,
                                                   float* __restrict__ r,
                                                   float* __restrict__ p, float* __restrict__ r0,
                                                   const float* __restrict__ dinv, uint32_t n, Scalars* S,
                                                   GridRed rb) {
  if (S->done || !S->verify) return;
  CSum acc[2] = {csum_zero(), csum_zero()};
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    CSum s{b[i], 0.0f};
    rows.foreach(i, [&](float a, uint32_t c) {
      const float xv = x[c] + xl[c];  // the FP32 iterate the solver would return
      const float pr = a * xv;
      const float e = __builtin_fmaf(a, xv, -pr);
      s = csum_add(s, -pr);
      s.c -= e;
    });
    const float ri = csum_val(s);
    const float zi = (MODE == 1) ? ri * dinv[i] : ri;
    r[i] = ri;
    if (MODE != 3) p[i] = zi;
    if (MODE == 2) r0[i] = ri;
    acc[0] = csum_add(acc[0], ri * ri);
    acc[1] = csum_add(acc[1], ri * zi);
  }
  block_partials<2>(acc, rb);
}

// The same check for CSR rows with many entries: L lanes per row (coalesced
// loads of the row), each lane a compensated partial sum, merged across the
// lanes (the row-per-thread form spent 5.5 ms on 144-entry rows at 1M rows).
template <int MODE, int L>
__global__ __launch_bounds__(kBlock) void k_verify_lanes(RowsCsr rows, const float* __restrict__ x,
                                                         const float* __restrict__ xl, const float* __restrict__ b,
                                                         float* __restrict__ r, float* __restrict__ p,
                                                         float* __restrict__ r0, const float* __restrict__ dinv,
                                                         uint32_t n, Scalars* S, GridRed rb) {
  if (S->done || !S->verify) return;
  constexpr uint32_t G = kBlock / L;
  const uint32_t lane = threadIdx.x % L, grp = threadIdx.x / L;
  CSum acc[2] = {csum_zero(), csum_zero()};
  for (uint32_t i = blockIdx.x * G + grp; i < n; i += gridDim.x * G) {  // uniform within a lane group
    CSum s = lane == 0 ? CSum{b[i], 0.0f} : csum_zero();
    const uint32_t j1 = rows.row_ptr[i + 1];
    // a lane's entries kB at a time: every column and value, then every
    // iterate load, then the sums in entry order (one entry per iteration
    // waited for its column's and then its iterate's load: 555 -> 420 us on
    // 144-entry rows at 1M rows)
    constexpr int kB = 4;
    for (uint32_t j0 = rows.row_ptr[i] + lane; j0 < j1; j0 += L * kB) {
      uint32_t cv[kB];
      float av[kB], xv[kB], lv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const uint32_t j = j0 + (uint32_t)u * L;
        cv[u] = j < j1 ? rows.col[j] : 0u;
        av[u] = j < j1 ? rows.val[j] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        xv[u] = x[cv[u]];
        lv[u] = xl[cv[u]];
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (j0 + (uint32_t)u * L < j1) {
          const float xi = xv[u] + lv[u];  // the FP32 iterate the solver would return
          const float pr = av[u] * xi;
          const float e = __builtin_fmaf(av[u], xi, -pr);
          s = csum_add(s, -pr);
          s.c -= e;
        }
      }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
      const CSum o{__shfl_xor(s.s, off, kWave), __shfl_xor(s.c, off, kWave)};
      s = csum_merge(s, o);
    }
    if (lane == 0) {
      const float ri = csum_val(s);
      const float zi = (MODE == 1) ? ri * dinv[i] : ri;
      r[i] = ri;
      if (MODE != 3) p[i] = zi;
      if (MODE == 2) r0[i] = ri;
      acc[0] = csum_add(acc[0], ri * ri);
      acc[1] = csum_add(acc[1], ri * zi);
    }
  }
  block_partials<2>(acc, rb);
}

// The same check for B x B block rows: a group of L = B lanes per row, lane l
// taking column l of each of the row's blocks (one coalesced block row per
// group), the lanes' compensated partial sums merged at the end (the
// row-per-thread form took 460 us on the 1M-row block-sparse system with
// B = 16, a quarter of a 1.9 ms BiCGSTAB solve).
template <int MODE, int L>
__global__ __launch_bounds__(kBlock) void k_verify_bcsr(BcsrView A, uint32_t m, const float* __restrict__ x,
                                                        const float* __restrict__ xl, const float* __restrict__ b,
                                                        float* __restrict__ r, float* __restrict__ p,
                                                        float* __restrict__ r0, const float* __restrict__ dinv,
                                                        uint32_t n, Scalars* S, GridRed rb) {
  if (S->done || !S->verify) return;
  constexpr uint32_t G = kBlock / L;
  const uint32_t lane = threadIdx.x % L, grp = threadIdx.x / L;
  CSum acc[2] = {csum_zero(), csum_zero()};
  for (uint32_t i = blockIdx.x * G + grp; i < n; i += gridDim.x * G) {  // uniform within a lane group
    CSum s = lane == 0 ? CSum{b[i], 0.0f} : csum_zero();
    const uint32_t br = i / L, ii = i % L;
    const uint32_t k0 = A.brow_ptr[br], k1 = A.brow_ptr[br + 1];
    // blocks kB at a time: every block column, then every value and iterate
    // load, then the sums in block order (one block per iteration waited for
    // its column's and then its iterate's load: 590 us, against 460 us for
    // the row-per-thread form; 358 us batched, 231 us with the larger grid)
    constexpr int kB = 8;
    for (uint32_t kb = k0; kb < k1; kb += kB) {
      uint32_t cv[kB];
      float av[kB], xv[kB], lv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) cv[u] = kb + u < k1 ? A.bcol[kb + u] * L + lane : 0u;
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const bool in = kb + u < k1 && cv[u] < m;
        const uint32_t c = in ? cv[u] : 0u;
        av[u] = in ? A.blocks[((size_t)(kb + u) * L + ii) * L + lane] : 0.0f;
        xv[u] = x[c];
        lv[u] = xl[c];
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (kb + u < k1 && cv[u] < m) {
          const float xi = xv[u] + lv[u];  // the FP32 iterate the solver would return
          const float pr = av[u] * xi;
          const float e = __builtin_fmaf(av[u], xi, -pr);
          s = csum_add(s, -pr);
          s.c -= e;
        }
      }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
      const CSum o{__shfl_xor(s.s, off, kWave), __shfl_xor(s.c, off, kWave)};
      s = csum_merge(s, o);
    }
    if (lane == 0) {
      const float ri = csum_val(s);
      const float zi = (MODE == 1) ? ri * dinv[i] : ri;
      r[i] = ri;
      if (MODE != 3) p[i] = zi;
      if (MODE == 2) r0[i] = ri;
      acc[0] = csum_add(acc[0], ri * ri);
      acc[1] = csum_add(acc[1], ri * zi);
    }
  }
  block_partials<2>(acc, rb);
}

// Totals of k_verify: true <r,r>, <r,z> -> converged / stagnated / restart.
template <int MODE>
struct FinVerify {
  Scalars* S;
  // explicit preconditioner: a restart replaced r, so w = G r is recomputed
  __device__ __forceinline__ void after() const {
    refresh_halt(S);
    if (MODE == 3) S->skip_g = (S->halt == 0 && S->restart != 0) ? 0 : 1;
  }
  __device__ __forceinline__ bool active() const { return S->done == 0 && S->verify != 0; }
  __device__ __forceinline__ void operator()(const float (&tot)[2]) const {
    const float best = S->nverify > 0 ? S->rr_best : tot[0] / 0.81f;
    S->verify = 0;
    S->nverify += 1;
    S->rr = tot[0];
    S->nfail = (tot[0] > 0.81f * best) ? S->nfail + 1 : 0;
    S->rr_best = fminf(best, tot[0]);
    S->beta = 0.0f;
    if (!isfinite(tot[0])) {
      S->done = kBreakdown;
    } else if (tot[0] <= S->thresh) {
      S->done = kConverged;
    } else if (S->nfail >= 2 || S->nverify >= 12) {
      S->done = kStagnated;
    } else if (S->iter >= S->maxiter) {
      S->done = kMaxIter;
    } else {
      // the recurrence drifted from the true residual: aim its next run at half
      // the target norm so that the returned iterate lands below tol
      S->thresh_rec = 0.25f * fminf(S->thresh_rec, S->thresh);
      S->collapse = 1;
      if (MODE == 2) S->rho = tot[0];
      else if (MODE == 3) S->restart = 1;
      else S->rz = tot[1];
    }
  }
};

// Whole CG / Jacobi-PCG solve in one workgroup for small systems (n <=
// kSmallN): 512 threads, thread t owns rows t, t + 512, ... with x, x_lo, r,
// p and the columns of its rows in registers (8 waves, so up to 256 registers
// each); the values of its rows and the direction p (or the iterate, during a
// verification) in LDS for the product's gathers; the scalar bookkeeping of
// FinCgInit / FinCgStep / FinVerify run by thread 0 (the step's on a register
// copy of the LDS state). On a 4096-row Laplacian a step is 3.8 us of LDS and
// barriers instead of four kernel launches (9 us while each step read A's
// rows from memory, one dependent row_ptr -> col -> p chain per row). Same
// recurrence, stopping rules, true-residual verification and restarts as the
// multi-kernel path (the reduction trees differ, so the bits do too).
constexpr int kSmallThreads = 512;
constexpr int kSmallRpt = 8;
// entries of a row held on chip for the step's product (longer rows read A
// from memory each step)
constexpr int kSmallRowCap = 7;
constexpr uint32_t kSmallN = (uint32_t)kSmallThreads * kSmallRpt;

template <int NDS>
__device__ __forceinline__ void small_total(CSum (&acc)[NDS], float (&tot)[NDS]) {
  __shared__ float s_tot[NDS];
  block_reduce<NDS, kSmallThreads / kWave>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < NDS; ++d) s_tot[d] = csum_val(acc[d]);
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < NDS; ++d) tot[d] = s_tot[d];
  __syncthreads();
}

template <int PC, bool COMP>
__global__ __launch_bounds__(kSmallThreads) void k_cg_small(RowsCsr A, const float* __restrict__ b, float* x_io,
                                                            const float* __restrict__ dinv, uint32_t n, int use_x0,
                                                            Scalars* Sg) {
  constexpr int RPT = kSmallRpt;
  __shared__ float sv[kSmallN];  // p, or the iterate during a verification
  __shared__ Scalars S;
  const int t = threadIdx.x;
  if (t == 0) S = *Sg;
  float x[RPT], xl[RPT], r[RPT], p[RPT], bv[RPT], dv[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const uint32_t i = t + k * kSmallThreads;
    const bool own = i < n;
    bv[k] = own ? b[i] : 0.0f;
    x[k] = (own && use_x0) ? x_io[i] : 0.0f;
    xl[k] = 0.0f;
    dv[k] = (own && PC == kPcJacobi) ? dinv[i] : 1.0f;
    if (own) sv[i] = x[k];
  }
  // the owned rows' entries on chip
  // (a missing entry holds value 0 and the row's own column, so the product
  // needs no branch: the row's sum gains +-0 terms, which change it only where
  // p_i is not finite; a row longer than kSmallRowCap holds kLongRow as its
  // first column and is read from memory)
  constexpr uint32_t kLongRow = 0xFFFFFFFFu;
  // a thread's kSmallRpt x kSmallRowCap values at a stride of one more word
  // (no bank conflicts; every offset fits a load's immediate)
  constexpr int kSv = kSmallRpt * kSmallRowCap + 1;
  __shared__ float sval[kSmallThreads * kSv];
  float* const my = sval + t * kSv;
  uint32_t rcol[RPT][kSmallRowCap];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const uint32_t i = t + k * kSmallThreads;
    const uint32_t a0 = i < n ? A.row_ptr[i] : 0u, a1 = i < n ? A.row_ptr[i + 1] : 0u;
    const uint32_t len = a1 - a0;
#pragma unroll
    for (int e = 0; e < kSmallRowCap; ++e) {
      const bool in = (uint32_t)e < len;
      rcol[k][e] = in ? A.col[a0 + e] : (i < n ? i : 0u);
      my[k * kSmallRowCap + e] = in ? A.val[a0 + e] : 0.0f;
    }
    if (len > (uint32_t)kSmallRowCap) rcol[k][0] = kLongRow;
  }
  __syncthreads();
  // r = b - A x0 (plain FP32 row sums, as the multi-kernel path's EpiResidual)
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const uint32_t i = t + k * kSmallThreads;
    float ax = 0.0f;
    if (i < n && use_x0) A.foreach(i, [&](float a, uint32_t c) { ax += a * sv[c]; });
    r[k] = bv[k] - ax;
  }
  __syncthreads();
  {
    CSum acc[3] = {csum_zero(), csum_zero(), csum_zero()};
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const uint32_t i = t + k * kSmallThreads;
      const float z = r[k] * dv[k];
      p[k] = z;
      if (i < n) sv[i] = z;
      acc[0] = csum_add(acc[0], bv[k] * bv[k]);
      acc[1] = csum_add(acc[1], r[k] * r[k]);
      acc[2] = csum_add(acc[2], r[k] * z);
    }
    float tot[3];
    small_total<3>(acc, tot);
    if (t == 0) FinCgInit{&S}(tot);
    __syncthreads();
  }
  while (true) {
    const int done = S.done, verify = S.verify;
    __syncthreads();
    if (done) break;
    if (!verify) {
      // q = A p, <p, q>; alpha; x, r update; <r, r>, <r, z>
      float q[RPT];
      CSum acc[1] = {csum_zero()};
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const uint32_t i = t + k * kSmallThreads;
        float s = 0.0f;
        if (rcol[k][0] != kLongRow) {  // the row's products in row order, then the padding
#pragma unroll
          for (int e = 0; e < kSmallRowCap; ++e) s += my[k * kSmallRowCap + e] * sv[rcol[k][e]];
        } else if (i < n) {
          A.foreach(i, [&](float a, uint32_t c) { s += a * sv[c]; });
        }
        q[k] = s;
        acc[0] = csum_add(acc[0], p[k] * s);
      }
      float tpq[1];
      small_total<1>(acc, tpq);
      const bool ok = finite_pos(tpq[0]);
      const float alpha = ok ? S.rz / tpq[0] : 0.0f;
      CSum acc2[2] = {csum_zero(), csum_zero()};
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        if constexpr (COMP) {
          comp_axpy(x[k], xl[k], alpha, p[k]);
        } else {
          x[k] += alpha * p[k];
        }
        r[k] -= alpha * q[k];
        acc2[0] = csum_add(acc2[0], r[k] * r[k]);
        acc2[1] = csum_add(acc2[1], r[k] * (r[k] * dv[k]));
      }
      float trr[2];
      small_total<2>(acc2, trr);
      if (t == 0) {  // on a register copy of the state (LDS accesses would chain)
        Scalars L = S;
        L.pq = ok ? tpq[0] : -1.0f;
        L.alpha = alpha;
        FinCgStep<PC>{&L}(trr);
        S = L;
      }
      __syncthreads();
      if (S.done == 0 && S.verify == 0) {
        const float beta = S.beta;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const uint32_t i = t + k * kSmallThreads;
          p[k] = r[k] * dv[k] + beta * p[k];
          if (i < n) sv[i] = p[k];
        }
      }
      __syncthreads();
    } else {
      // true residual of the FP32 iterate fl(x + x_lo), as k_verify
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const uint32_t i = t + k * kSmallThreads;
        if (i < n) sv[i] = x[k] + xl[k];
      }
      __syncthreads();
      float tr[RPT];
      CSum acc[2] = {csum_zero(), csum_zero()};
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const uint32_t i = t + k * kSmallThreads;
        CSum s{i < n ? b[i] : 0.0f, 0.0f};  // (b from memory: a rare path)
        if (i < n)
          A.foreach(i, [&](float a, uint32_t c) {
            const float xv = sv[c];
            const float pr = a * xv;
            const float e = __builtin_fmaf(a, xv, -pr);
            s = csum_add(s, -pr);
            s.c -= e;
          });
        tr[k] = i < n ? csum_val(s) : 0.0f;
        acc[0] = csum_add(acc[0], tr[k] * tr[k]);
        acc[1] = csum_add(acc[1], tr[k] * (tr[k] * dv[k]));
      }
      float tot[2];
      small_total<2>(acc, tot);
      if (t == 0) FinVerify<PC == kPcJacobi ? 1 : 0>{&S}(tot);
      __syncthreads();
      if (S.done == 0) {  // restart from the true residual with p = z
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const uint32_t i = t + k * kSmallThreads;
          x[k] = x[k] + xl[k];
          xl[k] = 0.0f;
          r[k] = tr[k];
          p[k] = tr[k] * dv[k];
          if (i < n) sv[i] = p[k];
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const uint32_t i = t + k * kSmallThreads;
    if (i < n) x_io[i] = x[k] + xl[k];
  }
  if (t == 0) {
    S.halt = 1;
    *Sg = S;
  }
}

// x = x + xl (collapse the compensated iterate to FP32)
__global__ __launch_bounds__(kBlock) void k_collapse(float* __restrict__ x, const float* __restrict__ xl, uint32_t n) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = x[i] + xl[i];
}

// Plain copy of n floats (vector-typed where aligned).
__global__ __launch_bounds__(kBlock) void k_copy(float* __restrict__ dst, const float* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace msp
