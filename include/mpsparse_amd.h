/*
 * mpsparse_amd -- C ABI of the MI355X sparse iterative-solver library.
 *
 * Plain C: opaque handles, raw pointers, sizes and status codes; no torch or
 * HIP types. Array arguments may be host memory or device memory of the
 * handle's GPU (the library detects which and stages host arrays itself).
 * `stream` is a hipStream_t passed as void* (NULL = the null stream). Calls
 * on one handle must not overlap (the handle owns one solver workspace).
 * On failure a function returns a nonzero msp_status and msp_last_error()
 * (thread-local) describes it.
 *
 * Each entry point names the reference interface it replaces
 * (SparseForge/mpsparse, file:line).
 */
#ifndef MPSPARSE_AMD_H
#define MPSPARSE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSP_ABI_VERSION 1

typedef enum msp_status {
  MSP_OK = 0,
  MSP_ERR_INVALID = 1,     /* bad argument (null pointer, size, dtype, ...) */
  MSP_ERR_KEY_RANGE = 2,   /* a COO key has row >= num_rows or col >= num_cols */
  MSP_ERR_HIP = 3,         /* HIP runtime failure */
  MSP_ERR_NO_DEVICE = 4,   /* no GPU visible */
  MSP_ERR_UNSUPPORTED = 5, /* option combination not implemented */
  MSP_ERR_SINGULAR = 6     /* zero / non-finite diagonal where one is required */
} msp_status;

typedef enum msp_value_type {
  MSP_VALUES_F32 = 0,
  MSP_VALUES_BF16 = 1,
  MSP_VALUES_F16 = 2
} msp_value_type;

/* Storage used by the SpMV of a CSR handle (msp_csr_set_format). AUTO keeps
 * the CSR arrays and adds diagonal storage when every entry lies on one of at
 * most 16 diagonals that are mostly full (stencil matrices). */
typedef enum msp_format {
  MSP_FORMAT_AUTO = 0,
  MSP_FORMAT_CSR = 1
} msp_format;

typedef enum msp_criterion {
  MSP_CRIT_RELRES = 0,   /* stop when ||r|| <= tol * ||b|| (checked every step) */
  MSP_CRIT_REFERENCE = 1 /* stop when sum(r^2) < tol, tested every 50 steps
                            (iterative/BiCGSTAB/spmv.cpp:17-18, 603-629) */
} msp_criterion;

typedef enum msp_precond {
  MSP_PRECOND_NONE = 0,
  MSP_PRECOND_JACOBI = 1
} msp_precond;

/* Input features of the learned-FSAI edge network (msp_fsai_select_features). */
typedef enum msp_fsai_feature_set {
  /* 8 features including the row couplings s_low, s_all and 1/sqrt(1 - s_low) */
  MSP_FSAI_FEATURES_COUPLED8 = 0,
  /* the edge's five quantities a_ij, a_ii, a_jj, nnz(row i), sum_j |a_ij|, made
   * scale-free (a_ij/sqrt(a_ii a_jj), log(a_jj/a_ii), nnz/8, sum|a|/a_ii,
   * [i == j]; three zero slots); G's rows are normalised to diag(G A G^T) = 1 */
  MSP_FSAI_FEATURES_STATED5 = 1
} msp_fsai_feature_set;

typedef enum msp_solve_status {
  MSP_SOLVE_CONVERGED = 1,
  MSP_SOLVE_BREAKDOWN = 2,
  MSP_SOLVE_MAXITER = 3,
  MSP_SOLVE_STAGNATED = 4 /* true residual stopped improving above tol */
} msp_solve_status;

typedef struct msp_solve_opts {
  double tol;            /* default 1e-6 */
  int32_t maxiter;       /* default 10000 */
  int32_t check_every;   /* steps enqueued between host polls of the device flag; default 16 */
  int32_t criterion;     /* msp_criterion */
  int32_t use_x0;        /* 0: start from x = 0 (reference behaviour); 1: x is the initial guess */
  int32_t precond;       /* msp_precond (CG only) */
  int32_t values;        /* msp_value_type used by the SpMV inside the loop */
  int32_t verify;        /* 1 (default): when the recurrence residual meets tol, recompute
                            r = b - A x with compensated FP32 sums and continue from it
                            unless it meets tol too (MSP_CRIT_RELRES only) */
  int32_t cg_form;       /* CG step form: 0 (default) timed on the first solve of a matrix,
                            1 direction formed in the SpMV gather (2 launches/step),
                            2 separate direction update (3 launches/step) */
  int32_t reserved[2];
} msp_solve_opts;

typedef struct msp_solve_info {
  int32_t iterations;
  int32_t status;        /* msp_solve_status */
  int32_t verifications; /* true-residual checks performed */
  int32_t reserved;
  double relres;         /* sqrt(rr / bb) at exit (the verified residual when verify = 1) */
  double rr;             /* ||r||^2 at exit */
  double bb;             /* ||b||^2 */
  double seconds;        /* device time of the iteration loop (HIP events) */
} msp_solve_info;

/* Mixed-precision iterative refinement (msp_csr_solve_refined). */
typedef struct msp_refine_opts {
  double tol;            /* outer target ||b - A x|| / ||b|| in FP64; default 1e-10 */
  double inner_tol;      /* relative tolerance of each FP32 correction solve; 0 (default):
                            1e-3, or the values' unit roundoff if larger (BF16: 2^-8) */
  int32_t max_outer;     /* default 30 */
  int32_t inner_maxiter; /* default 10000 */
  int32_t inner_solver;  /* 0: CG (default), 1: BiCGSTAB */
  int32_t values;        /* msp_value_type of the matrix inside the correction solve */
  int32_t precond;       /* msp_precond (CG) */
  int32_t use_x0;        /* 1: x holds an FP64 initial guess */
  int32_t reserved[4];
} msp_refine_opts;

typedef struct msp_refine_info {
  int32_t outer_iterations;
  int32_t inner_iterations; /* summed over all correction solves */
  int32_t status;           /* msp_solve_status */
  int32_t fallbacks;        /* correction solves repeated with FP32 values after a breakdown */
  double relres;            /* FP64 true relative residual at exit */
  double seconds;           /* wall time of the whole refinement */
} msp_refine_info;

/* Batched Jacobi-PCG (msp_batch_pcg). */
typedef struct msp_batch_opts {
  double tol;              /* relative residual per (system, rhs); default 1e-6 */
  int32_t maxiter;         /* per (system, rhs); default 10000 */
  int32_t nrhs;            /* right-hand sides per system, 1..4; default 4 */
  int32_t steps_per_graph; /* iterations per hipGraph launch (even); default 16 */
  int32_t tile_rows;       /* rows per tile, multiple of 64; 0 = autotune (default) */
  int32_t rows_per_thread; /* SpMM rows per thread, 1 or 2; 0 = autotune (default) */
  int32_t max_verify;      /* true-residual restarts per pair; default 3 */
  int32_t reserved[4];
} msp_batch_opts;

typedef struct msp_batch_info {
  int32_t converged_pairs;
  int32_t total_pairs;     /* num_systems * nrhs */
  int32_t max_iterations;  /* over all pairs */
  int32_t graph_launches;
  int32_t verifications;   /* batch-wide true-residual passes */
  int32_t tile_rows;       /* the schedule used */
  int32_t rows_per_thread;
  int32_t diagonal_planes;  /* product over batched diagonal storage with this many
                               value planes (every system on at most 8 diagonals);
                               0: the CSR product */
  int32_t symmetric_storage; /* 1: every system has 5 diagonals (-b, -a, 0, a, b) and
                                exactly symmetric values, so the product reads the lower
                                diagonals from the upper planes (3 planes streamed) */
  double seconds;          /* device time of the solve (events) */
  double setup_seconds;    /* host time of staging and tuning */
} msp_batch_info;

typedef struct msp_csr_s *msp_csr;
typedef struct msp_bcsr_s *msp_bcsr;
typedef struct msp_fsai_s *msp_fsai;
typedef struct msp_batch_s *msp_batch;

/* Library ----------------------------------------------------------------- */
int32_t msp_abi_version(void);
const char *msp_last_error(void);
void msp_default_opts(msp_solve_opts *opts);
/* Number of visible GPUs (0 when none); never fails. */
int32_t msp_device_count(void);
/* Temporaries of matrix builds (sort buffers, probe counters) come from a
 * library-owned scratch arena per (device, stream) that is kept for the next
 * call on that stream, and the blocks of destroyed matrices, preconditioners
 * and workspaces are kept in a per-device cache (at most 8 GiB per device)
 * for later allocations of about the same size. This returns the cached
 * blocks, and waits for each device with an idle arena and returns its memory
 * to the system (arenas in use by a call in progress are kept). A stream handle must not be destroyed and recreated while work that
 * used its arena is still running. No reference counterpart (the reference
 * allocates and frees Metal buffers per call, iterative/BiCGSTAB/spmv.cpp:62-228). */
msp_status msp_release_workspace(void);

/* Standalone COO -> CSR.
 * Replaces coo_csr.coo_csr(keys, values, row_ptr, col_ind, out_vals, num_rows, num_cols)
 * (spmv_kernel/coo_csr/coo_csr/coo_csr/coo_to_csr.cpp:251-283; sort semantics of
 * iterative/cgd/coo_csr.metal:17-193). keys[i] = (row << 32) | col. Writes
 * row_ptr[num_rows+1], col_ind[nnz], out_vals[nnz] (stable: duplicate keys keep
 * input order and are not summed). */
msp_status msp_coo_to_csr(const uint64_t *keys, const float *vals, uint64_t nnz, uint32_t num_rows,
                          uint32_t num_cols, uint32_t *row_ptr, uint32_t *col_ind, float *out_vals,
                          void *stream);

/* msp_coo_to_csr without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same results as
 * msp_coo_to_csr. Every pointer must be device memory; nnz is limited to 2^20
 * keys (larger inputs are refused with MSP_ERR_INVALID). The sort's temporaries
 * live in a fixed 32 MiB device array per GPU, held until the process ends
 * and shared by every call on that GPU: calls that overlap on one GPU race.
 * The call waits for its stream before it returns (it reads a status flag). */
msp_status msp_coo_to_csr_noalloc(const uint64_t *keys, const float *vals, uint64_t nnz, uint32_t num_rows,
                                  uint32_t num_cols, uint32_t *row_ptr, uint32_t *col_ind, float *out_vals,
                                  void *stream);

/* CSR matrix handle built on the GPU from packed COO keys.
 * Replaces spmv.csr_tensor(keys, values, row_ptr, col_ind, out_vals, num_rows, num_cols)
 * (iterative/BiCGSTAB/spmv.cpp:230-268, 62-228). */
msp_status msp_csr_create_coo(msp_csr *out, const uint64_t *keys, const float *vals, uint64_t nnz,
                              uint32_t num_rows, uint32_t num_cols, void *stream);

/* CSR handle from sorted CSR arrays (copied). Counterpart of the array
 * upload in bcsr.bcsr_tensor's constructor (iterative/BiCGSTAB/bcsr.cpp:357-422). */
msp_status msp_csr_create_csr(msp_csr *out, const uint32_t *row_ptr, const uint32_t *col_ind,
                              const float *vals, uint32_t num_rows, uint32_t num_cols, void *stream);

/* Copy the handle's CSR arrays out (any pointer may be NULL to skip it). The
 * reference constructors take these buffers but never fill them
 * (iterative/BiCGSTAB/spmv.cpp:246-248); the standalone converter does
 * (coo_to_csr.cpp:220-231). */
msp_status msp_csr_export(msp_csr A, uint32_t *row_ptr, uint32_t *col_ind, float *vals, void *stream);

msp_status msp_csr_shape(msp_csr A, uint32_t *num_rows, uint32_t *num_cols, uint64_t *nnz);

/* Build the reduced-precision value copy used when opts.values != F32: BF16 or
 * FP16 off-diagonal values, with the diagonal absorbing their rounding defects
 * (row sums kept) and stored as a 16-bit value plus an FP32 remainder. */
msp_status msp_csr_prepare_values(msp_csr A, int32_t value_type, void *stream);

/* Storage choice for the handle's SpMV (msp_format; default AUTO). The
 * reference always streams CSR (spmv.metal:11-49); AUTO adds diagonal storage
 * for stencil-like patterns. format_in_use (optional): MSP_FORMAT_CSR, 2
 * when diagonal storage is in use, 3 when it is and the values are exactly
 * symmetric (products whose value planes exceed 16 MB then read the lower
 * diagonals from the upper planes; results are bit-identical either way)
 * (known after the first SpMV or solve). */
msp_status msp_csr_set_format(msp_csr A, int32_t format);
msp_status msp_csr_get_format(msp_csr A, int32_t *format_in_use);

/* Release the handle (device buffers included). Replaces ~csr_tensor
 * (iterative/BiCGSTAB/spmv.cpp:354-366). */
msp_status msp_csr_destroy(msp_csr A);

/* y = A x. Replaces csr_tensor.mv(x, b) (iterative/BiCGSTAB/spmv.cpp:436-501;
 * kernel spmv.metal:11-49). */
msp_status msp_csr_spmv(msp_csr A, const float *x, float *y, int32_t value_type, void *stream);

/* CG on A x = b. Replaces csr_tensor.cgd(b, x) (iterative/BiCGSTAB/spmv.cpp:515-659)
 * and the cgd module's iter_solve(b, x) (iterative/cgd/spmv.cpp:607-747). */
msp_status msp_csr_cg(msp_csr A, const float *b, float *x, const msp_solve_opts *opts,
                      msp_solve_info *info, void *stream);

/* BiCGSTAB. Replaces csr_tensor.bicgstab(b, x) (iterative/BiCGSTAB/spmv.cpp:661-888). */
msp_status msp_csr_bicgstab(msp_csr A, const float *b, float *x, const msp_solve_opts *opts,
                            msp_solve_info *info, void *stream);

/* `iters` steps of x <- x + omega (b - A x). Replaces the richardson module's
 * iter_solve(b, x) (iterative/richardson/spmv.cpp:376-495: omega 0.01, 1000 steps,
 * x0 = 0) and the CUDA reference's loop (spmv_cuda_reference.cu:227-280). */
msp_status msp_csr_richardson(msp_csr A, const float *b, float *x, float omega, int32_t iters,
                              int32_t use_x0, int32_t value_type, void *stream);

/* Mixed-precision solve: FP64 outer residual r = b - A x and FP64 accumulation
 * x += d; each correction d solves A d = r / ||r|| in FP32 (values optionally
 * BF16/FP16, compensated FP32 dots). b and x are FP64 (host or device). The
 * reference has no counterpart (it is FP32 only); this is the precision path
 * layered on msp_csr_cg / msp_csr_bicgstab. */
void msp_default_refine_opts(msp_refine_opts *opts);
msp_status msp_csr_solve_refined(msp_csr A, const double *b, double *x, const msp_refine_opts *opts,
                                 msp_refine_info *info, void *stream);

/* FSAI preconditioner ------------------------------------------------------
 * M = G^T G with G lower triangular on the lower-triangle pattern of A
 * (diagonal included); applied inside PCG as two SpMVs. No counterpart in the
 * reference (which has no preconditioner): this is the learned-preconditioner
 * addition. G's values come from classical FSAI, from the caller, or from an
 * edge network's output o_e as g_e = o_e / sqrt(a_jj). A handle borrows A: A
 * must outlive it. */
msp_status msp_fsai_create(msp_fsai *out, msp_csr A, void *stream);
msp_status msp_fsai_destroy(msp_fsai F);
msp_status msp_fsai_num_edges(msp_fsai F, uint64_t *num_edges);
/* Edge list of the pattern (row, col), in G's CSR order. */
msp_status msp_fsai_edges(msp_fsai F, uint32_t *rows, uint32_t *cols, void *stream);
/* 8 features per edge: a_ij/sqrt(a_ii a_jj), log(a_jj/a_ii), nnz(row i)/8,
 * sum_j|a_ij|/a_ii, [i == j], s_low(i), s_all(i), 1/sqrt(max(1 - s_low(i), 1/64))
 * with s_low(i) = sum_{j<i} a_ij^2/(a_ii a_jj) and s_all(i) the same over
 * j != i; and the column scaling 1/sqrt(a_jj). */
msp_status msp_fsai_features(msp_fsai F, float *feat, float *colscale, void *stream);
/* Feature set of the edge network (msp_fsai_feature_set; a new handle has
 * MSP_FSAI_FEATURES_COUPLED8). Recomputes the features that msp_fsai_features
 * exports. With MSP_FSAI_FEATURES_STATED5, G from a network output (and the
 * training gradient) passes through the row normalisation
 * g_i <- g_i / sqrt(g_i A[J,J] g_i^T); a row whose quadratic form is not
 * positive keeps its scale and counts as a non-positive diagonal entry. */
msp_status msp_fsai_select_features(msp_fsai F, int32_t feature_set, void *stream);
msp_status msp_fsai_get_features(msp_fsai F, int32_t *feature_set);
/* Classical FSAI: per row, A[J,J] y = e_last, g = y/sqrt(y_last); rows that
 * cannot be solved become Jacobi rows (counted in fallback_rows). */
msp_status msp_fsai_classical(msp_fsai F, int32_t *fallback_rows, void *stream);
msp_status msp_fsai_set_values(msp_fsai F, const float *g, int32_t *nonpositive_diag, void *stream);
msp_status msp_fsai_set_network_output(msp_fsai F, const float *out, int32_t *nonpositive_diag, void *stream);
msp_status msp_fsai_export_values(msp_fsai F, float *g, void *stream);
/* z = G^T G r */
msp_status msp_fsai_apply(msp_fsai F, const float *r, float *z, void *stream);
/* loss = sum_k ||G A G^T w_k - w_k||^2 over k probe vectors (W: k rows of n);
 * adds dloss/dg_e into grad[num_edges] (device). */
msp_status msp_fsai_probe_loss_grad(msp_fsai F, const float *W, int32_t k, float *grad, double *loss,
                                    void *stream);
/* FSAI-preconditioned CG; falls back to plain CG (info->reserved = 1) when G
 * has a non-positive diagonal entry. */
msp_status msp_csr_pcg_fsai(msp_csr A, msp_fsai F, const float *b, float *x, const msp_solve_opts *opts,
                            msp_solve_info *info, void *stream);
/* msp_csr_solve_refined with FSAI-PCG corrections (opts->values selects A's
 * values inside them; inner_solver and precond are ignored). A correction that
 * ends without reaching inner_tol is redone with plain CG (info->fallbacks). */
msp_status msp_csr_solve_refined_fsai(msp_csr A, msp_fsai F, const double *b, double *x, const msp_refine_opts *opts,
                                      msp_refine_info *info, void *stream);

/* Edge MLP of the learned FSAI: 8 inputs (msp_fsai_features) -> 64 -> 64 -> 1,
 * ReLU, BF16 MFMA with FP32 accumulation. Parameters are msp_mlp_num_params()
 * FP32 values: W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3. All
 * arrays are device memory; feat is 16-byte aligned. No reference counterpart. */
int32_t msp_mlp_num_params(void);
msp_status msp_mlp_forward(const float *params, const float *feat, uint64_t num_edges, float *out, void *stream);
/* grad_params = d(sum_e grad_out[e] * out[e]) / d params (overwritten). */
msp_status msp_mlp_backward(const float *params, const float *feat, const float *grad_out, uint64_t num_edges,
                            float *grad_params, void *stream);
msp_status msp_mlp_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
                        float eps, int32_t step, float weight_decay, void *stream);
/* G from the network applied to F's features. */
msp_status msp_fsai_apply_network(msp_fsai F, const float *params, int32_t *nonpositive_diag, void *stream);
/* forward + probe loss (msp_fsai_probe_loss_grad) + backward: grad_params overwritten. */
msp_status msp_fsai_train_step(msp_fsai F, const float *params, const float *W, int32_t k, float *grad_params,
                               double *loss, void *stream);

/* Batched Jacobi-PCG ------------------------------------------------------
 * Many independent SPD systems stored as the diagonal blocks of one CSR matrix
 * A: system s owns rows and columns [sys_off[s], sys_off[s+1]) (host or
 * device array of num_systems + 1 entries, from 0 to num_rows). Entries
 * coupling two systems and rows without a positive diagonal are rejected. The
 * handle borrows A (A must outlive it). The SpMM tile schedule is autotuned on
 * the first solve; iterations run as hipGraph launches of steps_per_graph
 * PCG steps over all still-active systems. The reference has no batched
 * solver (one csr_tensor per system, iterative/BiCGSTAB/spmv.cpp:892-905);
 * this is the batching addition.
 * b, x: num_rows x nrhs, row-major (rhs fastest), host or device. The
 * iterate is carried as a compensated FP32 pair; msp_batch_pcg returns it
 * rounded to FP32 (a pair whose FP32-rounded solution cannot meet tol stops
 * as stagnated), msp_batch_pcg_x64 returns the pair's exact sum in FP64.
 * pair_status / pair_iterations / pair_relres: optional host arrays of
 * num_systems * nrhs (msp_solve_status, iterations, true relative residual of
 * the returned x). Calls on one handle must not overlap. */
void msp_default_batch_opts(msp_batch_opts *opts);
msp_status msp_batch_create(msp_batch *out, msp_csr A, const uint32_t *sys_off, uint32_t num_systems, void *stream);
msp_status msp_batch_pcg(msp_batch B, const float *b, float *x, const msp_batch_opts *opts, msp_batch_info *info,
                         int32_t *pair_status, int32_t *pair_iterations, double *pair_relres, void *stream);
msp_status msp_batch_pcg_x64(msp_batch B, const float *b, double *x, const msp_batch_opts *opts,
                             msp_batch_info *info, int32_t *pair_status, int32_t *pair_iterations,
                             double *pair_relres, void *stream);
msp_status msp_batch_destroy(msp_batch B);

/* Block CSR ---------------------------------------------------------------
 * Square B x B blocks (B = 4, 8 or 16), row-major inside a block (the layout of
 * scipy's BSR `data`), block rows of B rows; the last block row/column may be
 * partial (num_rows / num_cols need not be multiples of B). */

/* BCSR handle from block arrays (copied): brow_ptr[ceil(num_rows/B)+1],
 * bcol[nblk], blocks[nblk*B*B]. Replaces bcsr.bcsr_tensor(keys, values, row_ptr,
 * col_ind, num_rows, num_cols, block_size) (iterative/BiCGSTAB/bcsr.cpp:357-422,
 * binding bcsr.cpp:953-965), which is fixed to B = 4. */
msp_status msp_bcsr_create(msp_bcsr *out, const uint32_t *brow_ptr, const uint32_t *bcol, const float *blocks,
                           uint32_t num_rows, uint32_t num_cols, uint32_t block, void *stream);

/* BCSR handle from packed COO keys, converted on the GPU (COO -> CSR -> BCSR;
 * entries with equal (row, col) are summed). Replaces the reference's CPU
 * preprocessing coo_to_metal_bcsr (iterative/BiCGSTAB/bcsr_benchmark.py:36-66). */
msp_status msp_bcsr_create_coo(msp_bcsr *out, const uint64_t *keys, const float *vals, uint64_t nnz,
                               uint32_t num_rows, uint32_t num_cols, uint32_t block, void *stream);

msp_status msp_bcsr_shape(msp_bcsr A, uint32_t *num_rows, uint32_t *num_cols, uint32_t *block,
                          uint64_t *num_blocks);

/* Copy the block arrays out (NULL skips an array). */
msp_status msp_bcsr_export(msp_bcsr A, uint32_t *brow_ptr, uint32_t *bcol, float *blocks, void *stream);

msp_status msp_bcsr_destroy(msp_bcsr A);

/* y = A x (bcsr_spmv.metal:30-69 `spmv_op`; bcsr.cpp:438-452). */
msp_status msp_bcsr_spmv(msp_bcsr A, const float *x, float *y, void *stream);

/* CG / BiCGSTAB / Richardson on BCSR. Replace bcsr_tensor.cgd / .bicgstab
 * (iterative/BiCGSTAB/bcsr.cpp:577-721, 723-949). */
msp_status msp_bcsr_cg(msp_bcsr A, const float *b, float *x, const msp_solve_opts *opts, msp_solve_info *info,
                       void *stream);
msp_status msp_bcsr_bicgstab(msp_bcsr A, const float *b, float *x, const msp_solve_opts *opts,
                             msp_solve_info *info, void *stream);
msp_status msp_bcsr_richardson(msp_bcsr A, const float *b, float *x, float omega, int32_t iters, int32_t use_x0,
                               void *stream);

/* Direct sparse solve -----------------------------------------------------
 * Replaces SparseQRSolver(col_starts, row_indices, values, n_rows, n_cols) and
 * its solve(B) (sparse_bridge/bridge.cpp:14-191, Accelerate sparse QR on the
 * CPU). The matrix comes as CSC with int64 indices and float64 values
 * (bridge.cpp:24-57); duplicates are added. The host orders it into a band
 * (reverse Cuthill-McKee on A + A^T for square matrices, or the identity when
 * that gives a narrower band; rows sorted by leftmost column for tall ones) and
 * the GPU factors the band by Householder QR in FP64. num_rows >= num_cols;
 * the solve returns the least-squares solution (the solution for square
 * nonsingular A). Orderings with kl + ku + 1 > 256, kl > 127, or a band window
 * beyond the device's LDS are refused with MSP_ERR_UNSUPPORTED (tall matrices
 * need kl >= num_rows - num_cols), and a zero or non-finite diagonal of R with
 * MSP_ERR_SINGULAR. Arrays may be host or device memory. */
typedef struct msp_qr_s *msp_qr;
msp_status msp_qr_create_csc(msp_qr *out, const int64_t *col_starts, const int64_t *row_indices,
                             const double *values, int64_t num_rows, int64_t num_cols, void *stream);
/* X[:, k] = argmin ||A x - B[:, k]|| for k < nrhs; B column-major num_rows x nrhs
 * with column stride ldb >= num_rows, X num_cols x nrhs with stride ldx >= num_cols
 * (bridge.cpp:92-158; one right-hand side is nrhs = 1). */
msp_status msp_qr_solve(msp_qr F, const double *B, int64_t ldb, int64_t nrhs, double *X, int64_t ldx,
                        void *stream);
/* Lower / upper bandwidth of the ordered matrix. */
msp_status msp_qr_bandwidth(msp_qr F, int32_t *lower, int32_t *upper);
msp_status msp_qr_destroy(msp_qr F);

#ifdef __cplusplus
}
#endif

#endif /* MPSPARSE_AMD_H */
