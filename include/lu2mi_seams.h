/*
 * lu2mi_seams.h -- the reference's own link-time entry points, exported by
 * liblu2mi.so so that lu2's single-process objects link against the GPU
 * library unchanged.  Every symbol keeps the reference's name, argument
 * order and meaning; matrices are host memory, row-major unless stated, and
 * every call is synchronous (the library keeps no host pointer).
 *
 * GRAPE-DR's CBLAS enum arguments are passed as plain int with the CBLAS
 * values below; lu2's callers pass CBLAS enums, which have the same ABI.
 */
#ifndef LU2MI_SEAMS_H
#define LU2MI_SEAMS_H
#ifdef __cplusplus
extern "C" {
#endif

#define LU2MI_CBLAS_ROWMAJOR 101
#define LU2MI_CBLAS_COLMAJOR 102
#define LU2MI_CBLAS_NOTRANS  111
#define LU2MI_CBLAS_TRANS    112
#define LU2MI_NMAT           2048   /* gdrdgemm.h:10 NMAT */

/* ---- gdrdgemm.h (gdrdgemm.h:10-49) ------------------------------------ */

/* gdrdgemm.c:155-199: row-major C[m x n] = alpha*A[m x k]*B[k x n] + beta*C
 * with leading dimensions na, nb, nc.  Runs on the GPU (FP64 MFMA). */
void mygdrdgemm(int m, int n, int k, double alpha, double *a, int na, double *b, int nb,
                double beta, double *c, int nc);

/* gdrdgemm.c:64-120: CBLAS-shaped DGEMM.  Row- and column-major, NoTrans and
 * Trans operands (transposed operands are repacked on the host). */
void gdrblas_dgemm(int order, int transa, int transb, int M, int N, int K, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta,
                   double *C, int ldc);

/* gdrdgemm.c:17-58: known-answer test I*B == B on 2048^2 (B[i][j]=2048i+j)
 * with device re-initialisation between attempts.  The reference retries
 * forever; this library stops after 8 attempts with exit(-1), the
 * reference's fatal-error behaviour (lu2_mp_nompi.c:55-59). */
void gdr_check_and_restart(double a[][LU2MI_NMAT], double b[][LU2MI_NMAT],
                           double c[][LU2MI_NMAT]);

/* ---- GRAPE-DR device library (prototypes implicit in the reference;
 *      call sites gdrdgemm.c:38,55-56,105-110,171,185-186, lu2lib.c:2162) -- */

/* Fortran ABI, column-major, arguments by pointer.  *doneflag=1 on success;
 * left 0 for transposed operands so the caller takes its CBLAS fallback
 * (gdrdgemm.c:113-116,191-194). */
void gdr_dgemm_(int *nota, int *notb, int *m, int *n, int *k, double *alpha, double *beta,
                int *lda, int *ldb, int *ldc, const double *A, const double *B, double *C,
                int *doneflag);
void gdr_init(void);                      /* (re)initialise the device context */
void gdr_free(void);                      /* release the device context */
void gdrdgemm_set_procname(int procid);   /* gdrdgemm.c:171 */
void gdrsetboardid(int boardid);          /* lu2.c:1232, lu2lib.c:26-29 */
void gdrsetnboards(int nboards);          /* lu2.c:1233 */
void gdrsetforceswapab(void);             /* lu2lib.c:2162 */
void gdrresetforceswapab(void);           /* lu2lib.c:2171 (stub in the reference) */
void gdrsetskipsendjmat(void);            /* lu2lib.c:2172, MPI build only */
void gdrresetskipsendjmat(void);          /* lu2lib.c:2173, MPI build only */
void gdrdgemm_set_stress_factor(int x);   /* lu2lib.c:2176, MPI build only */
void set_matmul_msg_level(int level);     /* lu2lib.c:2175, MPI build only */

/* ---- lu2lib.h (lu2lib.h:1-23) ------------------------------------------ */

/* lu2lib.c:2079-2135: same contract as mygdrdgemm; every shape on the GPU. */
void mydgemm(int m, int n, int k, double alpha, double *a, int na, double *b, int nb,
             double beta, double *c, int nc);

/* lu2lib.c:1984-2075: C[m x n] -= A[m x kk]*B[kk x n] with leading
 * dimensions n1, n2, n3 (the reference declares them as VLAs a[][n1] ...). */
void matmul_for_small_nk(int n1, double *a, int n2, double *b, int n3, double *c, int m,
                         int kk, int n);

/* lu2lib.c:2138-2168: device self-test (the 2048^2 known-answer test) and
 * warm-up.  The reference's nb >= 2048 restriction (lu2lib.c:2142-2145) is
 * gone; a and awork are not touched. */
void reset_gdr(int m, double *a, int nb, double *awork, int n);

/* ---- lu2tlib.h (lu2tlib.h:2-18) ---------------------------------------- */

/* lu2tlib.c:1130-1137 (recursion lu2tlib.c:1095-1128): factor the panel held
 * column-major in a (a[c][r] = column c, row r, leading dimension n; rows
 * [i, n) of m columns) with lu2's recursive panel and pivot rule, on the GPU;
 * pv[0..m) receives the global pivot rows.  m must be 16 or 32 (lu2.c:1082
 * calls it with 16). */
void cm_column_decomposition_recursive(int n, double *a, int m, int *pv, int i);

/* ---- lu2.c:1122 (the driver's LU + solve) ------------------------------ */

/* lu2.c:1122-1144 lumcolumn(n, a, b, nb, awork, pv, recursive) on a host
 * matrix a[n][n+16] whose column n holds b: factored on the GPU in lu2's
 * stored form, x written to b[], the last block's pivots to pv[0..nb).
 * Pivot rule COMPAT (the reference's) unless LU2MI_PIVOT=gepp is set.
 * recursive must be 1, as the driver passes (lu2.c:1292).  lu2.c defines
 * its own lumcolumn: link it with -Dlumcolumn=lu2_host_lumcolumn. */
void lumcolumn(int n, double *a, double *b, int nb, double *awork, int *pv, int recursive);

#ifdef __cplusplus
}
#endif
#endif
