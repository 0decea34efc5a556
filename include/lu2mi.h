/*
 * lu2mi.h -- C ABI of the MI355X-native lu2 LU library (liblu2mi.so).
 *
 * The library factors lu2's augmented matrix on one GPU: row-major,
 * leading dimension lda >= n+1, right-hand side b stored in column n
 * (lu2.c:28 RDIM, lu2.c:70-85).  The stored factor form is lu2's
 * (lu2.c:524-581, lu2.c:1088-1120, lu2tlib.c:1041-1128): per nb-block the
 * pivot rows are scaled by 1/d (unit-upper U), the diagonal block holds
 * D^-1 L11 below the diagonal and d on it, rows below keep L21 unscaled,
 * and earlier L columns are never re-swapped.  Column n ends up as the
 * right-hand side of U x = y and x comes from the unit-upper back
 * substitution (lu2.c:110-116).
 *
 * All functions return LU2MI_OK (0) or a negative error code;
 * lu2mi_last_error() gives a message.  Pointers named d* are device
 * pointers (from lu2mi_malloc or any hipMalloc on the context's device).
 * Calls are synchronous unless stated otherwise.  A context may be used by
 * one host thread at a time.
 */
#ifndef LU2MI_H
#define LU2MI_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define LU2MI_OK               0
#define LU2MI_ERR_ARG         -1
#define LU2MI_ERR_HIP         -2
#define LU2MI_ERR_NOMEM       -3
#define LU2MI_ERR_DEVICE      -4   /* device-side failure (barrier time-out) */
#define LU2MI_ERR_UNSUPPORTED -5
#define LU2MI_ERR_KAT         -6   /* known-answer test failed */
#define LU2MI_ERR_NOCONV      -7   /* iterative refinement did not converge */

/* Pivot rules.  COMPAT reproduces lu2's leaf search, which reads column 0
 * of each 2-column leaf for both columns (lu2tlib.c:42-52 via
 * lu2tlib.c:1050), so odd column j pivots on argmax_{r>=j}|a[r][j-1]|.
 * GEPP is standard partial pivoting.  Ties go to the lowest row. */
#define LU2MI_PIV_COMPAT 0
#define LU2MI_PIV_GEPP   1

/* options for lu2mi_set_option */
#define LU2MI_OPT_PIVMODE  1   /* LU2MI_PIV_*; default COMPAT (the reference's rule) */
#define LU2MI_OPT_LEAF_W   2   /* panel leaf width: 16 or 32 (default 16) */
#define LU2MI_OPT_TIMING   3   /* 1: record HIP events around trailing updates */
#define LU2MI_OPT_LOOKAHEAD 4  /* 1: factor block i+1's panel on a second stream while
                                  block i's trailing update runs (default 1; 0 = serial
                                  schedule, bit-identical results) */
#define LU2MI_OPT_SUBPANEL 5   /* widest part of a panel (a power of two, 16..256) factored
                                  by one persistent launch with 16-wide leaves and grid
                                  barriers (default 0 = one launch per leaf and per GEMM,
                                  which measured faster; results identical either way) */
#define LU2MI_OPT_TRAIL_KERNEL 6  /* FP64 trailing-update and U12 DGEMM tile: 1 = operands
                                     loaded straight into LDS (default), 0 = staged through
                                     registers; the two sum each product's K terms in
                                     different orders */
#define LU2MI_OPT_DGEMM_TRAILING 7 /* 1: lu2mi_dgemm_dev (and lu2mi_sgemm_dev) run their
                                     product with the trailing-update kernel (to measure or
                                     test that kernel alone);
                                     2: the same, with A taken as lower triangular (the U12
                                     product's form: entries above the diagonal are not read) */
#define LU2MI_OPT_STRICT_SYNC 8   /* 1: the panel leaf and the back substitution publish their
                                     cross-workgroup data with agent-scope release/acquire
                                     (an L2 writeback and invalidate per publication); default
                                     0: device-coherent stores completed in order (results
                                     identical either way) */
#define LU2MI_OPT_WIDE16 9        /* 1: 16-bit trailing updates in 256 x 256 tiles (128 x 64 per
                                     wave) where they give every CU a workgroup; 0 (default):
                                     128 x 128 tiles (results identical either way) */
#define LU2MI_OPT_BACKSUB128 10   /* 1: back substitutions (the solve and the refinement's
                                     re-solves) in 128-row blocks, half as many cross-workgroup
                                     hand-overs; 0 (default): 64-row blocks.  The two sum each
                                     row's terms in different orders */

typedef struct lu2mi_ctx lu2mi_ctx;

int         lu2mi_init(int device, lu2mi_ctx **ctx);
int         lu2mi_finalize(lu2mi_ctx *ctx);
const char *lu2mi_last_error(void);
int         lu2mi_set_option(lu2mi_ctx *ctx, int opt, int value);
int         lu2mi_synchronize(lu2mi_ctx *ctx);
int         lu2mi_device_info(lu2mi_ctx *ctx, int *num_cu, char *arch, int arch_len);

/* device memory helpers */
int lu2mi_malloc(lu2mi_ctx *ctx, void **dptr, size_t bytes);
int lu2mi_free(lu2mi_ctx *ctx, void *dptr);
int lu2mi_memcpy_h2d(lu2mi_ctx *ctx, void *ddst, const void *hsrc, size_t bytes);
int lu2mi_memcpy_d2h(lu2mi_ctx *ctx, void *hdst, const void *dsrc, size_t bytes);
int lu2mi_memcpy_d2d(lu2mi_ctx *ctx, void *ddst, const void *dsrc, size_t bytes);

/* lu2.c:70-85 randomsetmat on the device: a[i][j] = drand48()-0.5 in row
 * order with n+1 draws per row (b last), bit-identical to glibc drand48
 * after srand48(seed).  Columns n+1..lda-1 are set to 0. */
int lu2mi_randomsetmat(lu2mi_ctx *ctx, int n, int seed, double *dA, long lda);

/* a[i][i] += sum_{j<n} |a[i][j]| for every row: lu2's random matrix made
 * diagonally dominant, the test matrix of the 16-bit solves (no counterpart
 * in the reference).  Each row sum is one device reduction; the result is
 * within the rounding of that sum of the sequential one. */
int lu2mi_make_diag_dominant(lu2mi_ctx *ctx, int n, double *dA, long lda);

/* lu2.c:1122-1144 lumcolumn(n, a, b, nb, awork, pv, 1), GPU-resident:
 * factors dA in place (lu2's stored form, pivot rule from
 * LU2MI_OPT_PIVMODE) and writes x to dx (may be NULL).  dpiv (may be NULL)
 * receives the full global pivot sequence, dpiv[j] = row exchanged with
 * row j at step j; the reference keeps only the last block's (lu2.c:1128).
 * Requires nb a power of two, 32 <= nb, n % nb == 0 (reference INSTALL:54-58). */
int lu2mi_lumcolumn_dev(lu2mi_ctx *ctx, int n, double *dA, long lda, int nb,
                        int *dpiv, double *dx);

/* lu2lib.c:2079-2135 mydgemm on device pointers: row-major
 * C[m x n] = alpha*A[m x k]*B[k x n] + beta*C, FP64 MFMA. */
int lu2mi_dgemm_dev(lu2mi_ctx *ctx, int m, int n, int k, double alpha,
                    const double *dA, long lda, const double *dB, long ldb,
                    double beta, double *dC, long ldc);

/* FP64 residual on the device against the unfactored augmented matrix
 * dA0 (b in column n).  out[0]=||Ax-b||_oo out[1]=||A||_oo out[2]=||A||_1
 * out[3]=||x||_oo out[4]=||x||_1 out[5]=||b||_oo
 * out[6]=HPL-2 scaled residual ||Ax-b||_oo/(eps(||A||_oo||x||_oo+||b||_oo)n),
 * eps=2^-53; out[7..9] = lu2_mpi.c:2170-2176 resid1..3 (eps=2e-16). */
int lu2mi_residual_dev(lu2mi_ctx *ctx, int n, const double *dA0, long lda,
                       const double *dx, double *out);

/* lu2mi_residual_dev without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  The ten
 * figures of lu2mi_residual_dev, bit for bit, go to dout (device memory, 10
 * doubles); the work is ordered on the context's stream and the call does not
 * wait for it.  The partial results live in a fixed array of
 * 3 * LU2MI_RESID_NMAX doubles (3 MiB) that each device holds until the
 * process ends, so n above LU2MI_RESID_NMAX is rejected (LU2MI_ERR_ARG);
 * calls that overlap on one device share that array and race. */
#define LU2MI_RESID_NMAX 131072
int lu2mi_residual_dev_noalloc(lu2mi_ctx *ctx, int n, const double *dA0, long lda,
                               const double *dx, double *dout);

/* ---- mixed precision ---------------------------------------------------- */
#define LU2MI_PREC_FP64 0
#define LU2MI_PREC_FP32 1
#define LU2MI_PREC_FP16 2   /* FP32 factors, trailing updates on FP16 operands */
#define LU2MI_PREC_BF16 3   /* FP32 factors, trailing updates on BF16 operands */
#define LU2MI_PREC_FP16X2 4 /* FP32 factors, trailing updates on FP16 operands split
                               hi + lo (hi*hi + hi*lo + lo*hi, ~22 bits) */

typedef struct {
    int    iterations;          /* refinement steps taken */
    int    gmres_used;          /* 1 once classical refinement stalled and GMRES-IR took over */
    int    gmres_inner;         /* total GMRES inner iterations */
    double resid_hpl2_initial;  /* HPL-2 scaled residual of the low-precision solution */
    double resid_hpl2;          /* ... after refinement (FP64, against the original matrix) */
    double t_factor_ms;         /* conversion + low-precision factorization */
    double t_refine_ms;         /* residuals, re-solves and updates */
} lu2mi_ir_info;

/* Solve lu2's augmented system (dA: FP64, b in column n, not modified) by an
 * LU of an FP32 copy (same pivot rule and factor form as
 * lu2mi_lumcolumn_dev) whose trailing updates run on the matrix cores with
 * FP32 accumulation on FP32 operands (prec FP32) or on operands rounded to
 * FP16 / BF16 under a power-of-two scale per 128-row block of L21 and per
 * 128-column block of U12 (prec FP16 / BF16; nb a multiple of 64), and
 * FP64 iterative refinement -- classical, switching to GMRES-IR preconditioned
 * by the same factors when it stalls -- until the HPL-2 scaled residual is at
 * most target or maxiter steps are done.  x is written to dx (FP64).  Returns
 * LU2MI_ERR_NOCONV when the target is not reached (dx then holds the last
 * iterate). */
int lu2mi_solve_mixed(lu2mi_ctx *ctx, int n, const double *dA, long lda, int nb, int prec,
                      double *dx, int maxiter, double target, lu2mi_ir_info *info);

/* The 16-bit trailing update on its own: C -= L U, all FP32 row-major on the
 * device (L: m x k, U: k x n), with L and U rounded to FP16 (fmt 1) or BF16
 * (fmt 2), or split into FP16 hi + lo (fmt 3), under per-128-block
 * power-of-two scales, FP32 accumulation.  k a multiple of 64. */
int lu2mi_gemm16_dev(lu2mi_ctx *ctx, int fmt, int m, int n, int k, const float *dL, long ldl,
                     const float *dU, long ldu, float *dC, long ldc);

/* FP32 row-major GEMM on device pointers, C = alpha*A*B + beta*C (FP32 MFMA). */
int lu2mi_sgemm_dev(lu2mi_ctx *ctx, int m, int n, int k, float alpha, const float *dA, long lda,
                    const float *dB, long ldb, float beta, float *dC, long ldc);

/* ---- batched uneven solve ----------------------------------------------
 * No counterpart in the reference, which solves one system per run
 * (lu2.c:1122-1144): many independent systems, each lu2's augmented
 * row-major [A | B] (1 <= n <= 512, 1 <= nrhs <= 8, B in columns
 * n..n+nrhs-1, leading dimension lda >= n+nrhs) at element offset a_off[s]
 * of one FP64 device pool.  Each is factored with partial pivoting (LU of a
 * working copy: dA is not modified) and solved; X[s] (n x nrhs, row-major)
 * is written at offset x_off[s] of the solution pool.  One persistent
 * launch serves all systems, longest first, from a queue; solves after the
 * first replay a captured HIP graph.  prec FP64: plain FP64.  prec FP32:
 * FP32 factors, FP64 classical refinement against dA until the HPL-2 scaled
 * residual is at most target or maxiter steps are done. */
#define LU2MI_BATCH_EAGER 1     /* flag: launch directly instead of the graph */

typedef struct lu2mi_batch lu2mi_batch;
typedef struct {
    double max_resid_hpl2;      /* max over systems and right-hand sides (NaN if any is NaN) */
    int    max_iterations;      /* refinement steps, 0 in the FP64 form */
    int    n_failed;            /* systems above target (FP32) or above 16 (FP64) */
    double t_ms;                /* device time of the launch (events around it) */
    int    nt, wp_small, wp_large, split;  /* launch shape in use: threads per system,
                                              panel width for n <= split and above */
} lu2mi_batch_info;

int lu2mi_batch_create(lu2mi_ctx *ctx, int nsys, const int *n, const int *nrhs, const long *a_off,
                       const long *lda, const long *x_off, int prec, int maxiter, double target,
                       lu2mi_batch **plan);
int lu2mi_batch_destroy(lu2mi_batch *plan);
/* per-system drand48 streams: system s is drawn row by row after
 * srand48(base_seed + s), entries drand48()-0.5 (lu2.c:70-85's rule with
 * nrhs B columns); padding columns are zeroed */
int lu2mi_batch_generate(lu2mi_batch *plan, double *dA, long base_seed);
/* autotuner: times every launch shape (threads per system 256/512, panel
 * width 16/32 per size class) on this batch and keeps the fastest */
int lu2mi_batch_tune(lu2mi_batch *plan, const double *dA, double *dX, int reps);
int lu2mi_batch_set_config(lu2mi_batch *plan, int nt, int wp_small, int wp_large, int split);
int lu2mi_batch_get_config(lu2mi_batch *plan, int *out4);
int lu2mi_batch_solve(lu2mi_batch *plan, const double *dA, double *dX, int flags, lu2mi_batch_info *info);
/* per-system HPL-2 residuals, refinement steps and pivot sequences (piv:
 * the systems' sequences concatenated in system order, piv[j] = row
 * exchanged with row j at step j); any pointer may be NULL */
int lu2mi_batch_results(lu2mi_batch *plan, double *resid, int *iters, int *piv);

/* Timing of the last lu2mi_lumcolumn_dev with LU2MI_OPT_TIMING on:
 * out[0]=total ms of trailing-update GEMMs (HIP events on the library
 * stream), out[1]=number of launches, out[2]=their algorithmic flops,
 * out[3]=wall ms of the whole call. */
int lu2mi_get_timing(lu2mi_ctx *ctx, double *out, int nout);

/* Per-subsystem timing of the last lu2mi_lumcolumn_dev with LU2MI_OPT_TIMING
 * on (lu2's print_timers groups, timerlib.c:47-70): HIP events around each
 * subsystem's launches on the stream they run on (with the lookahead schedule
 * the panel's stream runs beside the main one, so the times overlap).
 * out[4c..4c+3] for subsystem c = LU2MI_T_*: GPU ms, algorithmic flops,
 * algorithmic bytes, number of launch groups; out[4*LU2MI_T_COUNT] = wall ms
 * of the call, out[4*LU2MI_T_COUNT+1] = the device clock in Hz.
 *   SWAP    block-level row exchanges with the 1/d row scale (lu2.c:524-556):
 *           32 nb (n+1-i-nb) bytes per block (both rows of each exchange
 *           read and written)
 *   PANEL   the recursive panel (lu2.c:1088-1120, leaves lu2tlib.c:1041-1137):
 *           R nb^2 flops per block (R = n-i rows: the panel LU and the
 *           inverse of its unit-lower block)
 *   U12     the TRSM of the U row block (lu2.c:485-521): nb^2 (n+1-i-nb) flops
 *   TRAIL   the trailing update (lu2.c:312-343): 2 M N nb flops
 *   BACKSUB back substitution (lu2.c:110-116): n^2 flops, 4 n^2 bytes of U */
#define LU2MI_T_SWAP    0
#define LU2MI_T_PANEL   1
#define LU2MI_T_U12     2
#define LU2MI_T_TRAIL   3
#define LU2MI_T_BACKSUB 4
#define LU2MI_T_COUNT   5
int lu2mi_get_subsystem_timing(lu2mi_ctx *ctx, double *out, int nout);

/* The last product made with host memory through the reference's seams
 * (mygdrdgemm, gdrblas_dgemm, gdr_dgemm_, mydgemm, matmul_for_small_nk): bytes
 * uploaded, bytes downloaded, wall ms, row chunks.  The operands pass through
 * pinned bounce buffers of the seams' default context in chunks whose
 * packing, transfers and products overlap. */
int lu2mi_host_gemm_stats(double *out, int nout);

#ifdef __cplusplus
}
#endif
#endif
