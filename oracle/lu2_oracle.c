/*
 * lu2_oracle.c -- CPU restatement of lu2's single-process LU + solve.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the GPU library
 * (liblu2mi.so).  It is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, and by nothing else.
 *
 * Each function restates the reference algorithm it cites (file:line into
 * telmin/lu2).  The reference itself needs an external CBLAS that this image
 * does not provide, so it is not built; this restatement is pinned instead by
 *   - glibc's drand48 (generator),
 *   - LAPACK dgetrf through SciPy (gepp pivots, x),
 *   - a pure-Python unblocked elimination on small n (compat pivots),
 *   - lu2o_gepp_unblocked (textbook elimination: the blocked recursion must
 *     give the same pivots and the same stored factors to rounding),
 *   - the gdrdgemm.c:17-58 identity known-answer test and closed-form checks
 *     of the triangular inverse and back substitution.
 *
 * Stored factor form (lu2.c:524-581, lu2.c:1088-1120, lu2tlib.c:1041-1128):
 * for the nb-block starting at column i,
 *   rows i..i+nb-1 : left of the diagonal L[r][c]/d_r, diagonal d_r,
 *                    right of the diagonal the unit-upper U (row scaled by 1/d_r)
 *   rows >= i+nb   : L21 unscaled, in the row order at the end of block i
 *                    (later blocks never re-swap earlier L columns).
 * Column n carries b through the same swaps, scalings and updates, so after
 * the last block it holds the right-hand side of U x = y.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "lu2_oracle.h"

#define A_(r, c) a[(long)(r) * lda + (c)]

/* ------------------------------------------------------------------ */
/* drand48: POSIX 48-bit LCG.  srand48 keeps the low 32 bits of the seed
 * in the high bits of X0 and sets the low 16 bits to 0x330E; drand48
 * returns X/2^48 (exact in binary64).  Used by lu2.c:72-84.            */
void lu2o_srand48(long seed, unsigned long long *state)
{
    unsigned long long s = (unsigned long long)seed & 0xffffffffULL;
    *state = (s << 16) | 0x330EULL;
}

double lu2o_drand48(unsigned long long *state)
{
    *state = (0x5DEECE66DULL * *state + 0xBULL) & 0xffffffffffffULL;
    return (double)*state * (1.0 / 281474976710656.0);
}

/* lu2.c:70-85 randomsetmat: row i gets n draws for A, then one draw for b. */
void lu2o_randomsetmat(int n, int seed, double *a, long lda)
{
    unsigned long long st;
    lu2o_srand48(seed, &st);
    for (long r = 0; r < n; r++) {
        for (long c = 0; c < n; c++) A_(r, c) = lu2o_drand48(&st) - 0.5;
        A_(r, n) = lu2o_drand48(&st) - 0.5;
    }
}

/* Rows [r0, r0+nrows) of the lu2.c:70-85 matrix, by running the sequence
 * sequentially past the first r0*(n+1) draws. */
void lu2o_randomrows(int n, int seed, long r0, int nrows, double *a, long lda)
{
    unsigned long long st;
    lu2o_srand48(seed, &st);
    for (long e = 0; e < r0 * (long)(n + 1); e++) (void)lu2o_drand48(&st);
    for (long r = 0; r < nrows; r++) {
        for (long c = 0; c <= n; c++) A_(r, c) = lu2o_drand48(&st) - 0.5;
        for (long c = n + 1; c < lda; c++) A_(r, c) = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Row-major GEMM, C = alpha*A*B + beta*C.  Stands in for cblas_dgemm as
 * called through mydgemm (lu2lib.c:2079-2135) and lu2tlib.c:1026.  Register
 * blocked 4x8 with OpenMP over output tiles.                           */
typedef double v4d __attribute__((vector_size(32)));
typedef double v4du __attribute__((vector_size(32), aligned(8)));

static inline void mk4x8(int kc, const double *ap, long lda, const double *bp,
                         long ldb, double *cp, long ldc, double alpha)
{
    v4d c00 = {0}, c01 = {0}, c10 = {0}, c11 = {0};
    v4d c20 = {0}, c21 = {0}, c30 = {0}, c31 = {0};
    const double *a0 = ap, *a1 = ap + lda, *a2 = ap + 2 * lda, *a3 = ap + 3 * lda;
    for (int p = 0; p < kc; p++) {
        const double *brow = bp + (long)p * ldb;
        v4d b0 = *(const v4du *)brow;
        v4d b1 = *(const v4du *)(brow + 4);
        c00 += a0[p] * b0; c01 += a0[p] * b1;
        c10 += a1[p] * b0; c11 += a1[p] * b1;
        c20 += a2[p] * b0; c21 += a2[p] * b1;
        c30 += a3[p] * b0; c31 += a3[p] * b1;
    }
#define STORE(row, v0, v1)                                         \
    do {                                                           \
        double *cr = cp + (long)(row) * ldc;                       \
        v4d t0 = *(v4du *)cr, t1 = *(v4du *)(cr + 4);              \
        t0 += alpha * v0; t1 += alpha * v1;                        \
        *(v4du *)cr = t0; *(v4du *)(cr + 4) = t1;                  \
    } while (0)
    STORE(0, c00, c01); STORE(1, c10, c11); STORE(2, c20, c21); STORE(3, c30, c31);
#undef STORE
}

static void gemm_tile(int m, int n, int k, double alpha, const double *a, long lda,
                      const double *b, long ldb, double *c, long ldc)
{
    const int KC = 128;
    for (int k0 = 0; k0 < k; k0 += KC) {
        int kc = k - k0 < KC ? k - k0 : KC;
        int i = 0;
        for (; i + 4 <= m; i += 4) {
            int j = 0;
            for (; j + 8 <= n; j += 8)
                mk4x8(kc, a + (long)i * lda + k0, lda, b + (long)k0 * ldb + j, ldb,
                      c + (long)i * ldc + j, ldc, alpha);
            for (; j < n; j++)
                for (int r = i; r < i + 4; r++) {
                    double s = 0.0;
                    for (int p = 0; p < kc; p++)
                        s += a[(long)r * lda + k0 + p] * b[(long)(k0 + p) * ldb + j];
                    c[(long)r * ldc + j] += alpha * s;
                }
        }
        for (; i < m; i++)
            for (int j = 0; j < n; j++) {
                double s = 0.0;
                for (int p = 0; p < kc; p++)
                    s += a[(long)i * lda + k0 + p] * b[(long)(k0 + p) * ldb + j];
                c[(long)i * ldc + j] += alpha * s;
            }
    }
}

void lu2o_dgemm(int m, int n, int k, double alpha, const double *a, long lda,
                const double *b, long ldb, double beta, double *c, long ldc)
{
    if (m <= 0 || n <= 0) return;
    if (beta != 1.0) {
#pragma omp parallel for schedule(static) if ((long)m * n > 65536)
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++)
                c[(long)i * ldc + j] = (beta == 0.0) ? 0.0 : beta * c[(long)i * ldc + j];
    }
    if (k <= 0 || alpha == 0.0) return;
    const int TM = 64, TN = 256;
    int tm = (m + TM - 1) / TM, tn = (n + TN - 1) / TN;
    long ntiles = (long)tm * tn;
#pragma omp parallel for schedule(dynamic, 1) if ((double)m * n * k > 2.0e6)
    for (long t = 0; t < ntiles; t++) {
        int i0 = (int)(t / tn) * TM, j0 = (int)(t % tn) * TN;
        int mm = m - i0 < TM ? m - i0 : TM, nn = n - j0 < TN ? n - j0 : TN;
        gemm_tile(mm, nn, k, alpha, a + (long)i0 * lda, lda, b + j0, ldb,
                  c + (long)i0 * ldc + j0, ldc);
    }
}

/* C -= A*B on sub-blocks of the matrix itself (lu2.c:312-343 mmmulandsub,
 * lu2tlib.c:1003-1038). */
static void mmmulandsub(double *a, long lda, int m0, int m1, int c0, int c1, int r0, int r1)
{
    if (r1 <= r0 || c1 <= c0 || m1 <= m0) return;
    lu2o_dgemm(r1 - r0, c1 - c0, m1 - m0, -1.0, &A_(r0, m0), lda, &A_(m0, c0), lda,
               1.0, &A_(r0, c0), lda);
}

/* ------------------------------------------------------------------ */
static inline void swap_rows(double *a, long lda, int r1, int r2, int c0, int c1)
{
    if (r1 == r2) return;
    double *p = &A_(r1, 0), *q = &A_(r2, 0);
    for (int c = c0; c < c1; c++) { double t = p[c]; p[c] = q[c]; q[c] = t; }
}

static inline void scale_row(double *a, long lda, int r, double s, int c0, int c1)
{
    double *p = &A_(r, 0);
    for (int c = c0; c < c1; c++) p[c] *= s;
}

static inline int idamax_col(const double *a, long lda, int col, int r0, int n)
{
    /* first index of max |a[r][col]|, r in [r0, n) (cblas_idamax semantics) */
    int p = r0;
    double amax = fabs(A_(r0, col));
    for (int r = r0 + 1; r < n; r++) {
        double v = fabs(A_(r, col));
        if (v > amax) { amax = v; p = r; }
    }
    return p;
}

/* lu2tlib.c:1041-1066 column_decomposition, leaf of m<=2 columns at [i,i+m).
 * The pivot search reads column 0 of the leaf for every ip (lu2tlib.c:47 via
 * findpivot(n,a,ii) with a pointing at the leaf) in COMPAT mode; GEPP reads
 * column ii. */
static void cm_leaf(int n, double *a, long lda, int m, int *pv, int i, int pivmode)
{
    for (int ip = 0; ip < m; ip++) {
        int ii = i + ip;
        int col = (pivmode == LU2O_PIV_COMPAT) ? i : ii;
        int p = idamax_col(a, lda, col, ii, n);
        pv[ip] = p;
        swap_rows(a, lda, p, ii, i, i + m);
        double ainv = 1.0 / A_(ii, ii);
        scale_row(a, lda, ii, ainv, i, ii);
        scale_row(a, lda, ii, ainv, ii + 1, i + m);
        /* vvmulandsub: a[r][k] -= a[r][ii]*a[ii][k], k in (ii,i+m), r in (ii,n) */
#pragma omp parallel for schedule(static) if (n - ii > 4096)
        for (int r = ii + 1; r < n; r++) {
            double l = A_(r, ii);
            for (int k = ii + 1; k < i + m; k++) A_(r, k) -= l * A_(ii, k);
        }
    }
}

/* lu2tlib.c:1068-1094 process_right_part (column-major panel version):
 * block [i,i+m), right columns [i+m,iend). */
static void cm_right(int n, double *a, long lda, int m, const int *pv, int i, int iend)
{
    for (int ii = i; ii < i + m; ii++) swap_rows(a, lda, pv[ii - i], ii, i + m, iend);
    for (int ii = i; ii < i + m; ii++) scale_row(a, lda, ii, 1.0 / A_(ii, ii), i + m, iend);
    for (int ii = i; ii < i + m; ii++)
        for (int k = i + m; k < iend; k++) {
            double s = A_(ii, k);
            for (int j = ii + 1; j < i + m; j++) A_(j, k) -= A_(j, ii) * s;
        }
    mmmulandsub(a, lda, i, i + m, i + m, iend, i + m, n);
}

/* lu2tlib.c:1095-1128 column_decomposition_recursive (16 -> 8 -> 4 -> 2). */
static void cm_recursive(int n, double *a, long lda, int m, int *pv, int i, int pivmode)
{
    if (m <= 2) { cm_leaf(n, a, lda, m, pv, i, pivmode); return; }
    int h = m / 2;
    cm_recursive(n, a, lda, h, pv, i, pivmode);
    cm_right(n, a, lda, h, pv, i, i + m);
    cm_recursive(n, a, lda, h, pv + h, i + h, pivmode);
    for (int ii = i + h; ii < i + m; ii++) swap_rows(a, lda, pv[ii - i], ii, i, i + h);
    for (int ii = i + h; ii < i + m; ii++) scale_row(a, lda, ii, 1.0 / A_(ii, ii), i, i + h);
}

/* lu2tlib.c:1130-1137 cm_column_decomposition_recursive: the panel of m
 * columns starting at column i (rows [i, n)); restated on row-major storage
 * (the column-major copy of lu2.c:1072-1085 is pure data movement). */
int lu2o_cm_panel(int n, double *a, long lda, int m, int *pv, int i, int pivmode)
{
    if (m < 2 || (m & (m - 1)) || i < 0 || i + m > n || lda < i + m) return -1;
    cm_recursive(n, a, lda, m, pv, i, pivmode);
    return 0;
}

/* lu2.c:370-465 solve_triangle_for_unit_mat_recursive: b := L^-1 b for the
 * unit-lower part of the m x m block l (b = I on entry at the top level). */
static void tri_inv_rec(int m, const double *l, long ldl, double *b, long ldb)
{
    if (m < 16) {               /* lu2.c:370-381 */
        for (int ii = 0; ii < m; ii++)
            for (int j = ii + 1; j < m; j++)
                for (int k = 0; k < m; k++)
                    b[(long)j * ldb + k] -= l[(long)j * ldl + ii] * b[(long)ii * ldb + k];
        return;
    }
    int h = m / 2;
    tri_inv_rec(h, l, ldl, b, ldb);
    lu2o_dgemm(h, h, h, -1.0, l + (long)h * ldl, ldl, b, ldb, 1.0, b + (long)h * ldb, ldb);
    double *bw = (double *)calloc((size_t)h * h, sizeof(double));
    double *bw2 = (double *)malloc((size_t)h * h * sizeof(double));
    for (int j = 0; j < h; j++) bw[(long)j * h + j] = 1.0;
    tri_inv_rec(h, l + (long)h * ldl + h, ldl, bw, h);
    for (int r = 0; r < h; r++)
        for (int c = 0; c < h; c++) bw2[(long)r * h + c] = b[(long)(r + h) * ldb + c];
    lu2o_dgemm(h, h, h, 1.0, bw, h, bw2, h, 0.0, b + (long)h * ldb, ldb);
    for (int j = 0; j < h; j++)
        for (int k = 0; k <= j; k++) b[(long)(h + j) * ldb + h + k] = bw[(long)j * h + k];
    free(bw); free(bw2);
}

/* lu2.c:485-521 solve_triangle: U12 = (D^-1 L11)^-1 * rows, via an explicit
 * inverse and one GEMM. */
static void solve_triangle(int n, double *a, long lda, int m, int i, int iend)
{
    int nc = iend - i - m;
    if (nc <= 0) return;
    double *b = (double *)calloc((size_t)m * m, sizeof(double));
    double *w = (double *)malloc((size_t)m * nc * sizeof(double));
    for (int j = 0; j < m; j++) b[(long)j * m + j] = 1.0;
    tri_inv_rec(m, &A_(i, i), lda, b, m);
    for (int j = 0; j < m; j++) memcpy(w + (long)j * nc, &A_(i + j, i + m), sizeof(double) * nc);
    lu2o_dgemm(m, nc, m, 1.0, b, m, w, nc, 0.0, &A_(i, i + m), lda);
    free(b); free(w);
    (void)n;
}

/* lu2.c:524-581 process_right_part (row-major): swap + 1/d scale of the
 * block's rows over [i+m,iend), TRSM, then the trailing update. */
static void process_right_part(int n, double *a, long lda, int m, const int *pv, int i, int iend)
{
    int c0 = i + m;
    if (iend > c0) {
        /* column chunks are independent: parallel over columns as lu2.c:537-554 */
        int nt = omp_get_max_threads();
        int w = (iend - c0 + nt - 1) / nt;
#pragma omp parallel for schedule(static) if (iend - c0 > 2048)
        for (int t = 0; t < nt; t++) {
            int cs = c0 + t * w, ce = cs + w > iend ? iend : cs + w;
            if (cs >= ce) continue;
            for (int ii = i; ii < i + m; ii++) {
                swap_rows(a, lda, pv[ii - i], ii, cs, ce);
                scale_row(a, lda, ii, 1.0 / A_(ii, ii), cs, ce);
            }
        }
    }
    solve_triangle(n, a, lda, m, i, iend);
    mmmulandsub(a, lda, i, i + m, i + m, iend, i + m, n);
}

/* lu2.c:1088-1120 column_decomposition_recursive (row-major), leaf of <=16
 * columns through the column-major recursion (lu2.c:1072-1085; the transposed
 * copy is pure data movement and is not restated). */
static void column_decomposition_recursive(int n, double *a, long lda, int m, int *pv,
                                           int i, int pivmode)
{
    if (m <= 16) { cm_recursive(n, a, lda, m, pv, i, pivmode); return; }
    int h = m / 2;
    column_decomposition_recursive(n, a, lda, h, pv, i, pivmode);
    process_right_part(n, a, lda, h, pv, i, i + m);
    column_decomposition_recursive(n, a, lda, h, pv + h, i + h, pivmode);
    for (int ii = i + h; ii < i + m; ii++) swap_rows(a, lda, pv[ii - i], ii, i, i + h);
    for (int ii = i + h; ii < i + m; ii++) scale_row(a, lda, ii, 1.0 / A_(ii, ii), i, i + h);
}

/* lu2.c:110-116 backward_sub; U unit upper, x from column n. */
void lu2o_backward_sub(int n, const double *a, long lda, double *x)
{
    for (int r = 0; r < n; r++) x[r] = A_(r, n);
    for (int j = n - 2; j >= 0; j--) {
        double s = x[j];
        for (int k = j + 1; k < n; k++) s -= x[k] * A_(j, k);
        x[j] = s;
    }
}

static int check_args(int n, long lda, int nb)
{
    if (n <= 0 || lda < n + 1 || nb < 2) return -1;
    if (nb & (nb - 1)) return -2;       /* power of two (reference INSTALL:54-58) */
    if (n % nb) return -3;              /* n multiple of nb */
    return 0;
}

/* lu2.c:1122-1144 lumcolumn with recursive=1 (the driver's call, lu2.c:1292). */
int lu2o_lumcolumn(int n, double *a, long lda, int nb, int pivmode, int *piv, double *x)
{
    int rc = check_args(n, lda, nb);
    if (rc) return rc;
    for (int i = 0; i < n; i += nb) {
        column_decomposition_recursive(n, a, lda, nb, piv + i, i, pivmode);
        process_right_part(n, a, lda, nb, piv + i, i, n + 1);
    }
    if (x) lu2o_backward_sub(n, a, lda, x);
    return 0;
}

/* Unblocked elimination in lu2's stored form (lu2.c:120-156 lu(), with the
 * leaf rule of lu2tlib.c:1050 for COMPAT).  Row exchanges cover columns from
 * the start of the current nb-block through n, which is exactly the set the
 * blocked code swaps. */
int lu2o_gepp_unblocked(int n, double *a, long lda, int nb, int pivmode, int *piv, double *x)
{
    int rc = check_args(n, lda, nb);
    if (rc) return rc;
    for (int j = 0; j < n; j++) {
        int i0 = (j / nb) * nb;
        int col = (pivmode == LU2O_PIV_COMPAT && (j & 1)) ? j - 1 : j;
        int p = idamax_col(a, lda, col, j, n);
        piv[j] = p;
        swap_rows(a, lda, p, j, i0, n + 1);
        double s = 1.0 / A_(j, j);
        scale_row(a, lda, j, s, i0, j);
        scale_row(a, lda, j, s, j + 1, n + 1);
#pragma omp parallel for schedule(static) if ((long)(n - j) * (n - j) > 65536)
        for (int r = j + 1; r < n; r++) {
            double l = A_(r, j);
            double *ar = &A_(r, 0);
            const double *aj = &A_(j, 0);
            for (int k = j + 1; k <= n; k++) ar[k] -= l * aj[k];
        }
    }
    if (x) lu2o_backward_sub(n, a, lda, x);
    return 0;
}

/* ------------------------------------------------------------------ */
/* lu2.c:45-60 showresult */
double lu2o_emax(int n, const double *a, long lda, const double *x)
{
    double emax = 0.0;
#pragma omp parallel for reduction(max : emax) schedule(static)
    for (int r = 0; r < n; r++) {
        double s = 0.0;
        for (int c = 0; c < n; c++) s += A_(r, c) * x[c];
        double e = fabs(s - A_(r, n));
        if (e > emax) emax = e;
    }
    return emax;
}

void lu2o_residuals(int n, const double *a, long lda, const double *x, double *out)
{
    double rmax = 0.0, ainf = 0.0, bmax = 0.0;
    double *colsum = (double *)calloc((size_t)n, sizeof(double));
#pragma omp parallel
    {
        double *cs = (double *)calloc((size_t)n, sizeof(double));
#pragma omp for reduction(max : rmax, ainf, bmax) schedule(static)
        for (int r = 0; r < n; r++) {
            double s = 0.0, rs = 0.0;
            for (int c = 0; c < n; c++) {
                s += A_(r, c) * x[c];
                rs += fabs(A_(r, c));
                cs[c] += fabs(A_(r, c));
            }
            double e = fabs(s - A_(r, n));
            if (e > rmax) rmax = e;
            if (rs > ainf) ainf = rs;
            if (fabs(A_(r, n)) > bmax) bmax = fabs(A_(r, n));
        }
#pragma omp critical
        for (int c = 0; c < n; c++) colsum[c] += cs[c];
        free(cs);
    }
    double a1 = 0.0, xinf = 0.0, x1 = 0.0;
    for (int c = 0; c < n; c++) {
        if (colsum[c] > a1) a1 = colsum[c];
        if (fabs(x[c]) > xinf) xinf = fabs(x[c]);
        x1 += fabs(x[c]);
    }
    free(colsum);
    const double eps = ldexp(1.0, -53);
    const double eps_lu2 = 2e-16;        /* lu2_mpi.c:2172 */
    out[0] = rmax; out[1] = ainf; out[2] = a1; out[3] = xinf; out[4] = x1; out[5] = bmax;
    out[6] = rmax / (eps * (ainf * xinf + bmax) * (double)n);
    out[7] = rmax / (eps_lu2 * a1 * (double)n);
    out[8] = rmax / (eps_lu2 * a1 * x1);
    out[9] = rmax / (eps_lu2 * ainf * xinf * (double)n);
}

/* Re-solve with lu2's stored factors (see the file header for the form). */
void lu2o_factor_solve(int n, const double *a, long lda, int nb, const int *piv, double *r)
{
    for (int i = 0; i < n; i += nb) {
        int ie = i + nb < n ? i + nb : n;
        for (int ii = i; ii < ie; ii++) {
            int p = piv[ii];
            if (p != ii) { double t = r[ii]; r[ii] = r[p]; r[p] = t; }
            r[ii] *= 1.0 / A_(ii, ii);
        }
        for (int j = i + 1; j < ie; j++) {
            double s = r[j];
            for (int c = i; c < j; c++) s -= A_(j, c) * r[c];
            r[j] = s;
        }
#pragma omp parallel for schedule(static) if (n - ie > 4096)
        for (int j = ie; j < n; j++) {
            double s = r[j];
            for (int c = i; c < ie; c++) s -= A_(j, c) * r[c];
            r[j] = s;
        }
    }
    for (int j = n - 2; j >= 0; j--) {
        double s = r[j];
        for (int k = j + 1; k < n; k++) s -= r[k] * A_(j, k);
        r[j] = s;
    }
}

/* gdrdgemm.c:17-58 gdr_check_and_restart: identity times B must return B. */
long lu2o_kat_identity(int nmat)
{
    size_t sz = (size_t)nmat * nmat;
    double *ia = (double *)calloc(sz, sizeof(double));
    double *b = (double *)malloc(sz * sizeof(double));
    double *c = (double *)calloc(sz, sizeof(double));
    for (long i = 0; i < nmat; i++) {
        ia[i * nmat + i] = 1.0;
        for (long j = 0; j < nmat; j++) b[i * nmat + j] = (double)(i * nmat + j);
    }
    lu2o_dgemm(nmat, nmat, nmat, 1.0, ia, nmat, b, nmat, 0.0, c, nmat);
    long err = 0;
    for (size_t t = 0; t < sz; t++) err += (b[t] != c[t]);
    free(ia); free(b); free(c);
    return err;
}

/* ------------------------------------------------------------------ */
/* Batched systems.  lu2 has no batched form; these restate, per system,
 * the generator and the solve that the batched GPU form runs.           */

/* rows x cols entries drawn row by row after srand48(seed), value
 * drand48()-0.5: lu2.c:70-85's rule for an augmented matrix with
 * cols - rows right-hand sides.  Columns [cols, lda) are set to 0. */
void lu2o_random_aug(int rows, int cols, long seed, double *a, long lda)
{
    unsigned long long st;
    lu2o_srand48(seed, &st);
    for (long r = 0; r < rows; r++) {
        for (long c = 0; c < cols; c++) A_(r, c) = lu2o_drand48(&st) - 0.5;
        for (long c = cols; c < lda; c++) A_(r, c) = 0.0;
    }
}

/* Textbook Gaussian elimination with partial pivoting on the augmented
 * row-major [A | B] (n rows, n + nrhs columns): at step j the pivot is the
 * first row of max |a[r][j]|, r >= j (a NaN never wins; a column with no
 * comparable entry keeps row j), whole rows are swapped (LAPACK form), the
 * multipliers a[r][j] * (1/a[j][j]) are stored below the diagonal and the
 * rest of each row r > j, B included, is updated.  Then x = U^-1 y per
 * right-hand side, x row-major n x nrhs.  piv[j] = row exchanged with row j. */
int lu2o_gepp_solve(int n, int nrhs, double *a, long lda, int *piv, double *x)
{
    if (n <= 0 || nrhs <= 0 || lda < (long)n + nrhs) return -1;
    const int nc = n + nrhs;
    for (int j = 0; j < n; j++) {
        int p = -1;
        double best = -1.0;
        for (int r = j; r < n; r++) {
            double v = fabs(A_(r, j));
            if (v > best) { best = v; p = r; }
        }
        if (p < 0) p = j;
        piv[j] = p;
        if (p != j)
            for (int c = 0; c < nc; c++) { double t = A_(j, c); A_(j, c) = A_(p, c); A_(p, c) = t; }
        double inv = 1.0 / A_(j, j);
        for (int r = j + 1; r < n; r++) {
            double l = A_(r, j) * inv;
            A_(r, j) = l;
            for (int c = j + 1; c < nc; c++) A_(r, c) -= l * A_(j, c);
        }
    }
    for (int q = 0; q < nrhs; q++) {
        for (int i = n - 1; i >= 0; i--) {
            double s = A_(i, n + q);
            for (int k = i + 1; k < n; k++) s -= A_(i, k) * x[(long)k * nrhs + q];
            x[(long)i * nrhs + q] = s / A_(i, i);
        }
    }
    return 0;
}

/* a[i][i] += sum_{j<n} |a[i][j]| (sequential sum): lu2's random matrix made
 * diagonally dominant for the 16-bit solves. */
void lu2o_diag_dominant(int n, double *a, long lda)
{
    for (long r = 0; r < n; r++) {
        double s = 0.0;
        for (long c = 0; c < n; c++) s += fabs(A_(r, c));
        A_(r, r) += s;
    }
}
