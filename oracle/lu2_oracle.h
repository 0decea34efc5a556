/*
 * lu2_oracle.h -- CPU oracle for the lu2 single-process LU + solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * library is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (liblu2mi.so) never
 * links, loads or calls it.
 *
 * Matrix convention (lu2.c:28, lu2.c:70-85): row-major, leading dimension
 * lda >= n+1, right-hand side b stored as column n of the augmented matrix.
 *
 * Pivot modes:
 *   LU2O_PIV_COMPAT (0): lu2's leaf rule.  The 2-column leaf of the
 *       column-major panel searches column 0 of the leaf for both of its
 *       columns (lu2tlib.c:42-52 findpivot, called from lu2tlib.c:1050), so
 *       the pivot of every odd column j is argmax_{r>=j} |a[r][j-1]|.
 *   LU2O_PIV_GEPP   (1): standard partial pivoting, argmax_{r>=j} |a[r][j]|.
 * Ties resolve to the lowest row index (BLAS idamax semantics).
 *
 * Pivot output: piv[j] = global row exchanged with row j at step j (0-based).
 */
#ifndef LU2_ORACLE_H
#define LU2_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define LU2O_PIV_COMPAT 0
#define LU2O_PIV_GEPP   1

/* drand48 restatement (POSIX LCG, X0=(seed<<16)|0x330E, a=0x5DEECE66D, c=0xB) */
void   lu2o_srand48(long seed, unsigned long long *state);
double lu2o_drand48(unsigned long long *state);
/* lu2.c:70-85 randomsetmat: a[i][j]=drand48()-0.5 row by row, b=a[i][n] last */
void   lu2o_randomsetmat(int n, int seed, double *a, long lda);
/* rows [r0, r0+nrows) of the same matrix, sequence run from the start */
void   lu2o_randomrows(int n, int seed, long r0, int nrows, double *a, long lda);

/* Row-major C = alpha*A*B + beta*C (lu2lib.c:2079 mydgemm semantics). */
void   lu2o_dgemm(int m, int n, int k, double alpha, const double *a, long lda,
                  const double *b, long ldb, double beta, double *c, long ldc);

/* lu2.c:1122-1144 lumcolumn: blocked recursive LU of the augmented matrix
 * (factor form of lu2 kept in place), then backward_sub into x. */
int    lu2o_lumcolumn(int n, double *a, long lda, int nb, int pivmode,
                      int *piv, double *x);
/* Textbook unblocked GEPP producing the same stored factor form (lu2.c:120-156
 * with the leaf pivot rule of lu2tlib.c:1050 when pivmode is COMPAT). */
int    lu2o_gepp_unblocked(int n, double *a, long lda, int nb, int pivmode,
                           int *piv, double *x);
/* lu2tlib.c:1130-1137 cm_column_decomposition_recursive on the m-column panel
 * at column i of a row-major matrix (rows [i, n)); pv[k] = global pivot row. */
int    lu2o_cm_panel(int n, double *a, long lda, int m, int *pv, int i, int pivmode);
/* lu2.c:110-116 backward_sub on the unit-upper factor; x <- column n. */
void   lu2o_backward_sub(int n, const double *a, long lda, double *x);

/* lu2.c:45-60 showresult: Emax = max_i |(A x)_i - b_i| (A, b unfactored). */
double lu2o_emax(int n, const double *a, long lda, const double *x);
/* out[0]=||Ax-b||_oo out[1]=||A||_oo out[2]=||A||_1 out[3]=||x||_oo
 * out[4]=||x||_1 out[5]=||b||_oo
 * out[6]=HPL-2 scaled residual ||Ax-b||_oo/(eps(||A||_oo||x||_oo+||b||_oo)n),
 *        eps=2^-53
 * out[7..9]=lu2_mpi.c:2170-2176 resid1..3 with eps=2e-16 */
void   lu2o_residuals(int n, const double *a, long lda, const double *x,
                      double *out);

/* Re-solve with the stored factors for a new right-hand side r (length n),
 * following the block structure of lumcolumn: per block apply the pivots and
 * 1/d scaling, solve with the unit-lower diagonal block, update the rows
 * below with L21, then unit-upper back substitution.  r is overwritten with
 * the solution.  Used to pin iterative refinement. */
void   lu2o_factor_solve(int n, const double *a, long lda, int nb,
                         const int *piv, double *r);

/* n x n row-major GEMM known-answer test of gdrdgemm.c:17-58: with A = I and
 * B[i][j] = i*nmat+j, returns the number of entries with A*B != B. */
long   lu2o_kat_identity(int nmat);

/* Batched systems (lu2 has no batched form; the restated per-system rules).
 * lu2o_random_aug: rows x cols entries drawn row by row after
 *   srand48(seed), drand48()-0.5 (lu2.c:70-85's rule, cols-rows right-hand
 *   sides); columns [cols, lda) set to 0.
 * lu2o_gepp_solve: textbook GEPP on [A | B] (n x (n+nrhs)), whole-row swaps,
 *   multipliers a*(1/d) below the diagonal, then x = U^-1 y (x: n x nrhs
 *   row-major).  Pivot: first max |a|, NaN never wins, no comparable entry
 *   keeps row j.  Returns 0, or -1 on bad arguments. */
/* a[i][i] += sum_{j<n} |a[i][j]| (sequential sum; the 16-bit solves' matrix) */
void   lu2o_diag_dominant(int n, double *a, long lda);
void   lu2o_random_aug(int rows, int cols, long seed, double *a, long lda);
int    lu2o_gepp_solve(int n, int nrhs, double *a, long lda, int *piv, double *x);

#ifdef __cplusplus
}
#endif
#endif
