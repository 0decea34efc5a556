// lu2mi_batched.hip -- thousands of independent dense systems of uneven size
// (n <= 512, 1-8 right-hand sides each), LU with partial pivoting and solve,
// in one persistent launch that is captured once in a HIP graph and replayed.
//
// lu2 solves one system (lu2.c:1122-1144 lumcolumn).  Small systems cannot
// fill an MI355X one at a time, so here one workgroup owns a whole system
// and pulls the next one from a queue sorted by decreasing n (longest
// first): the 256 CUs stay busy while the sizes vary, and the launch is a
// single kernel whatever the mix.
//
// Per system (caller's augmented row-major [A | B], FP64, not modified):
//   copy (FP64 form) or convert (FP32 form) [A | B] into the working pool
//   right-looking blocked LU with partial pivoting, panel width WP (16/32,
//   chosen per size class by the autotuner):
//     the m x WP panel in LDS: per column a workgroup argmax of |a| (lowest
//       row on ties, NaN never wins), row swap, L scaled by 1/d, rank-1
//       update of the rest of the panel
//     the panel's swaps on every other column (whole rows, LAPACK form)
//     X = L11^-1 formed in LDS (one row-elimination step per column), then
//     U12 = X A12 on the matrix cores, in place
//     A22 -= L21 U12 on the matrix cores: L21 from LDS, U12 and A22 from L2
//   y (the B columns) -> x = U^-1 y, blocked, in FP64 arithmetic
//   FP32 form: classical iterative refinement in FP64 against the original
//     FP64 [A | B]: r = b - A x, d = U^-1 L^-1 P r, x += d, until the HPL-2
//     scaled residual is at most the target or maxiter steps are done
//   both forms record each system's HPL-2 scaled residual
//   ||Ax-b||_oo / (eps (||A||_oo ||x||_oo + ||b||_oo) n), eps = 2^-53,
//   the maximum over its right-hand sides, and its pivot sequence.
#include "lu2mi_internal.h"
#include "lu2mi_dev.h"
#include <limits.h>
#include <algorithm>
#include <numeric>
#include <vector>

namespace {
constexpr int BNMAX = 512;   // largest system
constexpr int BRHSMAX = 8;   // most right-hand sides per system
constexpr int BSZ = 32;      // diagonal block of the triangular solves
constexpr double EPS53 = 1.1102230246251565e-16;

struct BatchSys {
    long a_off, lda;   // caller's FP64 [A | B], row-major
    long x_off;        // caller's X, n x nrhs row-major
    long w_off;        // working copy in the working precision, n x ldw
    long p_off;        // pivot pool
    int n, nrhs, ldw, wp, id, pad;
};

template <typename T>
struct BatchArgs {
    const BatchSys *sys;   // in scheduling order
    int nsys;
    const double *A;
    double *X;
    T *Wk;
    int *piv;
    double *resid;         // by caller's index
    int *iters;
    unsigned *queue;
    int maxiter;
    double target;
    int vstride;           // doubles per LDS vector (nmax * max nrhs)
};

__device__ __forceinline__ bool better(double v, int r, double bv, int br)
{
    return (v > bv) || (v == bv && r < br);
}
__device__ __forceinline__ double nanmax(double a, double b)
{
    return (a != a) ? a : ((b != b) ? b : (a > b ? a : b));
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_nanmax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ---------------------------------------------------------------------------
// A22 -= L21 * U12: rows [j0+w, n), columns [j0+w, nc).  Each wave takes
// 32x32 tiles of A22 (2x2 matrix-core blocks); the K = w operand columns of
// L21 come from the LDS panel, U12 rows from the pivot rows just written.
template <typename T, int NT, int WP>
__device__ __forceinline__ void batch_update(T *W, int ldw, int j0, int w, int M2, int N2, const T *P)
{
    typedef Mfma<T> F;
    constexpr int PP = WP + 1, NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int tm_n = (M2 + 31) >> 5, tn_n = (N2 + 31) >> 5, ntile = tm_n * tn_n;
    T *C0 = W + (long)(j0 + w) * ldw + j0 + w;
    const T *B0 = W + (long)j0 * ldw + j0 + w;
    // two tiles per wave at a time (tile and tile + NW), so that the C and B
    // loads of both are in flight together: each tile alone is one load
    // latency, a few matrix-core steps and a store
    for (int tile = wv; tile < ntile; tile += 2 * NW) {
        typename F::acc_t acc[2][2][2];   // [pair][bi][bj]
        int r0[2], c0[2];
        bool on[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int tl = tile + u * NW;
            on[u] = tl < ntile;
            const int tm = on[u] ? tl / tn_n : 0, tn = on[u] ? tl - tm * tn_n : 0;
            r0[u] = tm * 32;
            c0[u] = tn * 32;
#pragma unroll
            for (int bi = 0; bi < 2; bi++)
#pragma unroll
                for (int bj = 0; bj < 2; bj++)
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int row = r0[u] + bi * 16 + F::out_row(lane, rr), col = c0[u] + bj * 16 + l16;
                        acc[u][bi][bj][rr] = (on[u] && row < M2 && col < N2) ? C0[(long)row * ldw + col] : T(0);
                    }
        }
#pragma unroll
        for (int k0 = 0; k0 < WP; k0 += 4) {
            if (k0 < w) {
                const int k = k0 + lk;
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    T af[2], bf[2];
#pragma unroll
                    for (int bi = 0; bi < 2; bi++) {
                        const int row = r0[u] + bi * 16 + l16;
                        af[bi] = (on[u] && k < w && row < M2) ? -P[(w + row) * PP + k] : T(0);
                    }
#pragma unroll
                    for (int bj = 0; bj < 2; bj++) {
                        const int col = c0[u] + bj * 16 + l16;
                        bf[bj] = (on[u] && k < w && col < N2) ? B0[(long)k * ldw + col] : T(0);
                    }
#pragma unroll
                    for (int bi = 0; bi < 2; bi++)
#pragma unroll
                        for (int bj = 0; bj < 2; bj++) acc[u][bi][bj] = F::step(af[bi], bf[bj], acc[u][bi][bj]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int bi = 0; bi < 2; bi++)
#pragma unroll
                for (int bj = 0; bj < 2; bj++)
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int row = r0[u] + bi * 16 + F::out_row(lane, rr), col = c0[u] + bj * 16 + l16;
                        if (on[u] && row < M2 && col < N2) C0[(long)row * ldw + col] = acc[u][bi][bj][rr];
                    }
    }
}

// U12 = X A12 with X = L11^-1 (in LDS, pitch WP+1), in place on the w pivot
// rows: each wave takes 32-column tiles; a tile's w rows are all read (the
// K loop) before any of them is written, and tiles are disjoint.
template <typename T, int NT, int WP>
__device__ __forceinline__ void batch_trsm(T *W, int ldw, int j0, int w, int N2, const T *X)
{
    typedef Mfma<T> F;
    constexpr int PX = WP + 1, NW = NT / 64, MB = WP / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    T *B0 = W + (long)j0 * ldw + j0 + w;
    const int tn_n = (N2 + 31) >> 5;
    for (int tn = wv; tn < tn_n; tn += NW) {
        const int c0 = tn * 32;
        typename F::acc_t acc[MB][2];
#pragma unroll
        for (int bi = 0; bi < MB; bi++)
#pragma unroll
            for (int bj = 0; bj < 2; bj++) acc[bi][bj] = (typename F::acc_t){T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k0 = 0; k0 < WP; k0 += 4) {
            if (k0 < w) {
                const int k = k0 + lk;
                T af[MB], bf[2];
#pragma unroll
                for (int bi = 0; bi < MB; bi++) {
                    const int row = bi * 16 + l16;
                    af[bi] = (k < w && row < w) ? X[row * PX + k] : T(0);
                }
#pragma unroll
                for (int bj = 0; bj < 2; bj++) {
                    const int col = c0 + bj * 16 + l16;
                    bf[bj] = (k < w && col < N2) ? B0[(long)k * ldw + col] : T(0);
                }
#pragma unroll
                for (int bi = 0; bi < MB; bi++)
#pragma unroll
                    for (int bj = 0; bj < 2; bj++) acc[bi][bj] = F::step(af[bi], bf[bj], acc[bi][bj]);
            }
        }
#pragma unroll
        for (int bi = 0; bi < MB; bi++)
#pragma unroll
            for (int bj = 0; bj < 2; bj++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int row = bi * 16 + F::out_row(lane, rr), col = c0 + bj * 16 + l16;
                    if (row < w && col < N2) B0[(long)row * ldw + col] = acc[bi][bj][rr];
                }
    }
}

// ---------------------------------------------------------------------------
// Blocked right-looking LU with partial pivoting of W (n rows, nc = n + nrhs
// columns).  piv[j] = row exchanged with row j at step j; perm[i] = original
// row now at row i.
template <typename T, int NT, int WP>
__device__ __forceinline__ void batch_factor(T *__restrict__ W, int ldw, int n, int nc, T *P, T *Xi, int *piv,
                                             int *perm, double *red_v, int *red_i)
{
    constexpr int PP = WP + 1, NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int j0 = 0; j0 < n; j0 += WP) {
        const int w = min(WP, n - j0), m = n - j0;
        for (int idx = t; idx < m * w; idx += NT) {
            const int r = idx / w, c = idx - r * w;
            P[r * PP + c] = W[(long)(j0 + r) * ldw + j0 + c];
        }
        __syncthreads();
        for (int jj = 0; jj < w; jj++) {
            double bv = -1.0;
            int br = INT_MAX;
            for (int r = jj + t; r < m; r += NT) {
                const double v = fabs((double)P[r * PP + jj]);
                if (better(v, r, bv, br)) { bv = v; br = r; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(bv, o, 64);
                const int orow = __shfl_xor(br, o, 64);
                if (better(ov, orow, bv, br)) { bv = ov; br = orow; }
            }
            if (lane == 0) { red_v[wv] = bv; red_i[wv] = br; }
            __syncthreads();
            bv = red_v[0];
            br = red_i[0];
#pragma unroll
            for (int k = 1; k < NW; k++)
                if (better(red_v[k], red_i[k], bv, br)) { bv = red_v[k]; br = red_i[k]; }
            const int p = (br == INT_MAX) ? jj : br;   // a column of NaN keeps the current row
            if (t == 0) {
                piv[j0 + jj] = j0 + p;
                const int tmp = perm[j0 + jj];
                perm[j0 + jj] = perm[j0 + p];
                perm[j0 + p] = tmp;
            }
            if (p != jj)
                for (int c = t; c < w; c += NT) {
                    const T x = P[jj * PP + c];
                    P[jj * PP + c] = P[p * PP + c];
                    P[p * PP + c] = x;
                }
            __syncthreads();
            const T inv = T(1) / P[jj * PP + jj];
            for (int r = jj + 1 + t; r < m; r += NT) {
                const T l = P[r * PP + jj] * inv;
                P[r * PP + jj] = l;
                for (int c = jj + 1; c < w; c++) P[r * PP + c] -= l * P[jj * PP + c];
            }
            __syncthreads();
        }
        for (int idx = t; idx < m * w; idx += NT) {
            const int r = idx / w, c = idx - r * w;
            W[(long)(j0 + r) * ldw + j0 + c] = P[r * PP + c];
        }
        // X = L11^-1 for the TRSM, in Xi: start from I, then row k (final once
        // step k begins) eliminates column k from the rows below it
        for (int idx = t; idx < WP * WP; idx += NT) {
            const int i = idx / WP, j = idx - i * WP;
            Xi[i * PP + j] = (i == j) ? T(1) : T(0);
        }
        // the panel's swaps on all other columns, in step order
        const int no = nc - w;
        for (int cc = t; cc < no; cc += NT) {
            const int col = cc < j0 ? cc : cc + w;
            for (int jj = 0; jj < w; jj++) {
                const int p = piv[j0 + jj];
                if (p != j0 + jj) {
                    T *x = W + (long)(j0 + jj) * ldw + col;
                    T *y = W + (long)p * ldw + col;
                    const T v = *x;
                    *x = *y;
                    *y = v;
                }
            }
        }
        __syncthreads();
        for (int k = 0; k + 1 < w; k++) {
            const int cnt = (w - k - 1) * (k + 1);
            for (int idx = t; idx < cnt; idx += NT) {
                const int i = k + 1 + idx / (k + 1), j = idx % (k + 1);
                Xi[i * PP + j] -= P[i * PP + k] * Xi[k * PP + j];
            }
            __syncthreads();
        }
        // U12 = L11^-1 A12 on the matrix cores
        const int N2 = nc - j0 - w;
        batch_trsm<T, NT, WP>(W, ldw, j0, w, N2, Xi);
        __syncthreads();
        if (m - w > 0 && N2 > 0) batch_update<T, NT, WP>(W, ldw, j0, w, m - w, N2, P);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Triangular solves on LDS vectors v[i*nr + c] (FP64) with the factors in W.
// The BSZ x BSZ diagonal block is staged in LDS and solved by wave 0 with
// shuffles; the off-block update is a row-parallel GEMV.
template <typename T, int NT>
__device__ __forceinline__ void batch_back(const T *W, int ldw, int n, int nr, double *v, double *blk)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int ib = ((n - 1) / BSZ) * BSZ; ib >= 0; ib -= BSZ) {
        const int bs = min(BSZ, n - ib);
        for (int idx = t; idx < bs * bs; idx += NT) {
            const int r = idx / bs, c = idx - r * bs;
            blk[r * (BSZ + 1) + c] = (double)W[(long)(ib + r) * ldw + ib + c];
        }
        __syncthreads();
        if (wv == 0) {
            double vi[BRHSMAX];
#pragma unroll
            for (int c = 0; c < BRHSMAX; c++) vi[c] = (c < nr && lane < bs) ? v[(ib + lane) * nr + c] : 0.0;
            for (int k = bs - 1; k >= 0; k--) {
                const double d = blk[k * (BSZ + 1) + k];
                const double u = (lane < k) ? blk[lane * (BSZ + 1) + k] : 0.0;
#pragma unroll
                for (int c = 0; c < BRHSMAX; c++)
                    if (c < nr) {
                        const double xk = __shfl(vi[c], k, 64) / d;
                        if (lane == k) vi[c] = xk;
                        else if (lane < k) vi[c] -= u * xk;
                    }
            }
            if (lane < bs)
                for (int c = 0; c < nr; c++) v[(ib + lane) * nr + c] = vi[c];
        }
        __syncthreads();
        for (int i = t; i < ib; i += NT) {
            const T *row = W + (long)i * ldw + ib;
            double s[BRHSMAX];
#pragma unroll
            for (int c = 0; c < BRHSMAX; c++) s[c] = 0.0;
            for (int k = 0; k < bs; k++) {
                const double u = (double)row[k];
#pragma unroll
                for (int c = 0; c < BRHSMAX; c++)
                    if (c < nr) s[c] = fma(u, v[(ib + k) * nr + c], s[c]);
            }
            for (int c = 0; c < nr; c++) v[i * nr + c] -= s[c];
        }
        __syncthreads();
    }
}

template <typename T, int NT>
__device__ __forceinline__ void batch_fwd(const T *W, int ldw, int n, int nr, double *v, double *blk)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int ib = 0; ib < n; ib += BSZ) {
        const int bs = min(BSZ, n - ib), ie = ib + bs;
        for (int idx = t; idx < bs * bs; idx += NT) {
            const int r = idx / bs, c = idx - r * bs;
            blk[r * (BSZ + 1) + c] = (double)W[(long)(ib + r) * ldw + ib + c];
        }
        __syncthreads();
        if (wv == 0) {
            double vi[BRHSMAX];
#pragma unroll
            for (int c = 0; c < BRHSMAX; c++) vi[c] = (c < nr && lane < bs) ? v[(ib + lane) * nr + c] : 0.0;
            for (int k = 0; k < bs; k++) {
                const double l = (lane > k && lane < bs) ? blk[lane * (BSZ + 1) + k] : 0.0;
#pragma unroll
                for (int c = 0; c < BRHSMAX; c++)
                    if (c < nr) {
                        const double xk = __shfl(vi[c], k, 64);
                        if (lane > k) vi[c] -= l * xk;
                    }
            }
            if (lane < bs)
                for (int c = 0; c < nr; c++) v[(ib + lane) * nr + c] = vi[c];
        }
        __syncthreads();
        for (int i = ie + t; i < n; i += NT) {
            const T *row = W + (long)i * ldw + ib;
            double s[BRHSMAX];
#pragma unroll
            for (int c = 0; c < BRHSMAX; c++) s[c] = 0.0;
            for (int k = 0; k < bs; k++) {
                const double l = (double)row[k];
#pragma unroll
                for (int c = 0; c < BRHSMAX; c++)
                    if (c < nr) s[c] = fma(l, v[(ib + k) * nr + c], s[c]);
            }
            for (int c = 0; c < nr; c++) v[i * nr + c] -= s[c];
        }
        __syncthreads();
    }
}

// r = b - A x against the caller's FP64 [A | B]; returns the HPL-2 scaled
// residual, the maximum over the nr right-hand sides.  One wave per row.
template <int NT>
__device__ __forceinline__ double batch_resid(const double *A0, long lda, int n, int nr, const double *x, double *r,
                              double *red)
{
    constexpr int NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // a wave takes 8 rows at a time, 8 lanes per row (lane sl of row sub
    // reads elements sl, sl + 8, ...: 64-byte pieces of 8 rows per load), so
    // a row's sums need 3 shuffle steps instead of a 6-step wave reduction
    const int sub = lane >> 3, sl = lane & 7;
    double amax = 0.0;
    for (int i0 = 8 * wv; i0 < n; i0 += 8 * NW) {
        const int i = i0 + sub;
        const double *row = A0 + (long)i * lda;
        double s[BRHSMAX];
#pragma unroll
        for (int c = 0; c < BRHSMAX; c++) s[c] = 0.0;
        double asum = 0.0;
        if (i < n)
            for (int j = sl; j < n; j += 8) {
                const double a = row[j];
                asum += fabs(a);
#pragma unroll
                for (int c = 0; c < BRHSMAX; c++)
                    if (c < nr) s[c] = fma(a, x[j * nr + c], s[c]);
            }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            asum += __shfl_xor(asum, o, 64);
#pragma unroll
            for (int c = 0; c < BRHSMAX; c++)
                if (c < nr) s[c] += __shfl_xor(s[c], o, 64);
        }
        amax = nanmax(amax, asum);
        if (i < n && sl < nr) {
            double sc = s[0];
#pragma unroll
            for (int c = 1; c < BRHSMAX; c++)
                if (sl == c) sc = s[c];
            r[i * nr + sl] = row[n + sl] - sc;
        }
    }
    amax = wave_nanmax(amax);   // lanes of different rows hold different maxima
    // per column maxima of |r|, |x|, |b|; the row-sum maximum
    double mr[BRHSMAX], mx[BRHSMAX], mb[BRHSMAX];
#pragma unroll
    for (int c = 0; c < BRHSMAX; c++) mr[c] = mx[c] = mb[c] = 0.0;
    __syncthreads();
    for (int i = t; i < n; i += NT) {
#pragma unroll
        for (int c = 0; c < BRHSMAX; c++)
            if (c < nr) {
                mr[c] = nanmax(mr[c], fabs(r[i * nr + c]));
                mx[c] = nanmax(mx[c], fabs(x[i * nr + c]));
                mb[c] = nanmax(mb[c], fabs(A0[(long)i * lda + n + c]));
            }
    }
    constexpr int RW = 3 * BRHSMAX + 1;
#pragma unroll
    for (int c = 0; c < BRHSMAX; c++) {
        mr[c] = wave_nanmax(mr[c]);
        mx[c] = wave_nanmax(mx[c]);
        mb[c] = wave_nanmax(mb[c]);
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < BRHSMAX; c++) {
            red[wv * RW + c] = mr[c];
            red[wv * RW + BRHSMAX + c] = mx[c];
            red[wv * RW + 2 * BRHSMAX + c] = mb[c];
        }
        red[wv * RW + 3 * BRHSMAX] = amax;
    }
    __syncthreads();
    double an = red[3 * BRHSMAX];
    for (int k = 1; k < NW; k++) an = nanmax(an, red[k * RW + 3 * BRHSMAX]);
    double res = 0.0;
    for (int c = 0; c < nr; c++) {
        double a = red[c], b = red[BRHSMAX + c], bb = red[2 * BRHSMAX + c];
        for (int k = 1; k < NW; k++) {
            a = nanmax(a, red[k * RW + c]);
            b = nanmax(b, red[k * RW + BRHSMAX + c]);
            bb = nanmax(bb, red[k * RW + 2 * BRHSMAX + c]);
        }
        res = nanmax(res, a / (EPS53 * (an * b + bb) * (double)n));
    }
    __syncthreads();   // red is reused by the next call
    return res;
}

template <typename T, int NT>
__device__ __forceinline__ void batch_one(const BatchArgs<T> &a, const BatchSys &s, unsigned char *smem, double *lt,
                                          int *piv, int *perm, double *red_v, int *red_i)
{
    const int t = threadIdx.x;
    const int n = s.n, nr = s.nrhs, nc = n + nr, ldw = s.ldw;
    const double *A0 = a.A + s.a_off;
    T *W = a.Wk + s.w_off;
    // a wave per row, lanes along it; a row's loads are all issued before its
    // stores (nc <= BNMAX + BRHSMAX = 9 x 64)
    for (int i = t >> 6; i < n; i += NT / 64) {
        T v[(BNMAX + BRHSMAX + 63) / 64];
#pragma unroll
        for (int k = 0; k < (BNMAX + BRHSMAX + 63) / 64; k++) {
            const int j = (t & 63) + 64 * k;
            v[k] = (j < nc) ? (T)A0[(long)i * s.lda + j] : T(0);
        }
#pragma unroll
        for (int k = 0; k < (BNMAX + BRHSMAX + 63) / 64; k++) {
            const int j = (t & 63) + 64 * k;
            if (j < nc) W[(long)i * ldw + j] = v[k];
        }
    }
    for (int i = t; i < n; i += NT) perm[i] = i;
    __syncthreads();
    if (s.wp == 16) batch_factor<T, NT, 16>(W, ldw, n, nc, (T *)smem, (T *)lt, piv, perm, red_v, red_i);
    else batch_factor<T, NT, 32>(W, ldw, n, nc, (T *)smem, (T *)lt, piv, perm, red_v, red_i);

    double *xv = (double *)smem, *rv = xv + a.vstride, *cv = rv + a.vstride, *blk = cv + a.vstride;
    for (int idx = t; idx < n * nr; idx += NT) {
        const int i = idx / nr, c = idx - i * nr;
        xv[idx] = (double)W[(long)i * ldw + n + c];
    }
    __syncthreads();
    batch_back<T, NT>(W, ldw, n, nr, xv, blk);
    int it = 0;
    double res = batch_resid<NT>(A0, s.lda, n, nr, xv, rv, red_v);
    if (sizeof(T) < sizeof(double)) {
        while (!(res <= a.target) && it < a.maxiter) {
            for (int idx = t; idx < n * nr; idx += NT) {
                const int i = idx / nr, c = idx - i * nr;
                cv[idx] = rv[perm[i] * nr + c];
            }
            __syncthreads();
            batch_fwd<T, NT>(W, ldw, n, nr, cv, blk);
            batch_back<T, NT>(W, ldw, n, nr, cv, blk);
            for (int idx = t; idx < n * nr; idx += NT) xv[idx] += cv[idx];
            __syncthreads();
            it++;
            res = batch_resid<NT>(A0, s.lda, n, nr, xv, rv, red_v);
        }
    }
    double *X = a.X + s.x_off;
    for (int idx = t; idx < n * nr; idx += NT) X[idx] = xv[idx];
    for (int i = t; i < n; i += NT) a.piv[s.p_off + i] = piv[i];
    if (t == 0) {
        a.resid[s.id] = res;
        a.iters[s.id] = it;
    }
    __syncthreads();
}
}  // namespace

template <typename T, int NT>
__global__ __launch_bounds__(NT) void batch_lu_kernel(BatchArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_piv[BNMAX], s_perm[BNMAX];
    __shared__ double s_lt[32 * 33];   // inverse of the current panel's L11
    __shared__ double red_v[(NT / 64) * (3 * BRHSMAX + 1)];
    __shared__ int red_i[NT / 64];
    __shared__ int s_q;
    for (;;) {
        if (threadIdx.x == 0) s_q = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        if (q >= a.nsys) break;   // every wave of the workgroup reads the same q
        batch_one<T, NT>(a, a.sys[q], smem, s_lt, s_piv, s_perm, red_v, red_i);
    }
}

__global__ void batch_reset_kernel(unsigned *queue)
{
    if (threadIdx.x == 0) *queue = 0u;
}

// out[0] = max scaled residual (NaN if any), out[1] = max iterations,
// out[2] = systems above fail_above (or NaN)
__global__ __launch_bounds__(256) void batch_summary_kernel(const double *resid, const int *iters, int nsys,
                                                           double fail_above, double *out)
{
    __shared__ double rm[4], im[4], fc[4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double mres = 0.0, mit = 0.0, nf = 0.0;
    for (int i = t; i < nsys; i += 256) {
        const double r = resid[i];
        mres = nanmax(mres, r);
        mit = fmax(mit, (double)iters[i]);
        nf += (r <= fail_above) ? 0.0 : 1.0;
    }
    mres = wave_nanmax(mres);
    mit = wave_nanmax(mit);
    nf = wave_sum(nf);
    if (lane == 0) { rm[wv] = mres; im[wv] = mit; fc[wv] = nf; }
    __syncthreads();
    if (t == 0) {
        for (int k = 1; k < 4; k++) { rm[0] = nanmax(rm[0], rm[k]); im[0] = fmax(im[0], im[k]); fc[0] += fc[k]; }
        out[0] = rm[0];
        out[1] = im[0];
        out[2] = fc[0];
    }
}

// per-system drand48 streams: system s draws its n x (n+nrhs) entries in row
// order after srand48(base + s), value drand48() - 0.5 (the lu2.c:70-85 rule
// with nrhs columns of B); padding columns are set to 0
__global__ __launch_bounds__(256) void batch_generate_kernel(const BatchSys *sys, double *A, long base)
{
    using namespace lu2mi_rng;
    const BatchSys s = sys[blockIdx.x];
    const int nc = s.n + s.nrhs, t = threadIdx.x;
    const unsigned long long x0 = seed_state((unsigned long long)(base + s.id));
    const Lcg step = lcg_pow(256);
    for (int r = blockIdx.y; r < s.n; r += gridDim.y) {
        const Lcg jump = lcg_pow((unsigned long long)r * nc + t + 1);
        unsigned long long x = (jump.a * x0 + jump.c) & M48;
        double *row = A + s.a_off + (long)r * s.lda;
        for (int c = t; c < nc; c += 256) {
            row[c] = to_unit(x) - 0.5;
            x = (step.a * x + step.c) & M48;
        }
        for (long c = nc + t; c < s.lda; c += 256) row[c] = 0.0;
    }
}

// ---------------------------------------------------------------------------
struct lu2mi_batch {
    lu2mi_ctx *ctx;
    int nsys, prec, nmax, nrmax;
    int maxiter;
    double target;
    std::vector<int> n, nrhs;
    std::vector<long> a_off, lda, x_off, p_off, w_off;
    std::vector<int> ldw, order;
    BatchSys *d_sys;
    void *d_work;
    int *d_piv;
    double *d_resid;
    int *d_iters;
    unsigned *d_queue;
    double *d_sum;
    long piv_total;
    // launch shape (autotuner)
    int nt, wp_small, wp_large, split;
    int tuned;
    int wpmax, occupancy, grid;
    size_t smem;
    // captured graph
    hipGraph_t graph;
    hipGraphExec_t exec;
    const double *g_A;
    double *g_X;
    int g_valid;
    hipEvent_t ev0, ev1;
    float last_ms;
};

namespace {
int batch_vec_bytes(const lu2mi_batch *p) { return (3 * p->nmax * p->nrmax + BSZ * (BSZ + 1)) * 8; }

template <typename T, int NT>
int batch_prepare_t(lu2mi_batch *p)
{
    // panel rows: at least WP so the unused tail of a short last panel stays in the allocation
    const size_t panel = (size_t)std::max(p->nmax, 32) * (p->wpmax + 1) * sizeof(T);
    p->smem = std::max(panel, (size_t)batch_vec_bytes(p));
    LU2MI_CHECK(hipFuncSetAttribute((const void *)batch_lu_kernel<T, NT>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->smem));
    int occ = 0;
    LU2MI_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_lu_kernel<T, NT>, NT, p->smem));
    if (occ < 1) {
        lu2mi_set_error_msg("batched solve: kernel does not fit a CU (%zu bytes of LDS)", p->smem);
        return LU2MI_ERR_UNSUPPORTED;
    }
    p->occupancy = occ;
    p->grid = std::max(1, std::min(p->nsys, occ * p->ctx->num_cu));
    return LU2MI_OK;
}

// launch shape for the current configuration; called outside stream capture
int batch_prepare(lu2mi_batch *p)
{
    if (p->prec == LU2MI_PREC_FP32)
        return p->nt == 512 ? batch_prepare_t<float, 512>(p) : batch_prepare_t<float, 256>(p);
    return p->nt == 512 ? batch_prepare_t<double, 512>(p) : batch_prepare_t<double, 256>(p);
}

int batch_upload(lu2mi_batch *p)
{
    std::vector<BatchSys> h(p->nsys);
    p->wpmax = 16;
    for (int q = 0; q < p->nsys; q++) {
        const int s = p->order[q];
        BatchSys &d = h[q];
        d.a_off = p->a_off[s];
        d.lda = p->lda[s];
        d.x_off = p->x_off[s];
        d.w_off = p->w_off[s];
        d.p_off = p->p_off[s];
        d.n = p->n[s];
        d.nrhs = p->nrhs[s];
        d.ldw = p->ldw[s];
        d.wp = (p->n[s] <= p->split) ? p->wp_small : p->wp_large;
        d.id = s;
        d.pad = 0;
        p->wpmax = std::max(p->wpmax, d.wp);
    }
    LU2MI_CHECK(hipStreamSynchronize(p->ctx->stream));   // no launch may still read the descriptors
    if (p->nsys)
        LU2MI_CHECK(hipMemcpy(p->d_sys, h.data(), sizeof(BatchSys) * p->nsys, hipMemcpyHostToDevice));
    p->g_valid = 0;
    return batch_prepare(p);
}

template <typename T, int NT>
int batch_launch_t(lu2mi_batch *p, const double *dA, double *dX)
{
    BatchArgs<T> a;
    a.sys = p->d_sys;
    a.nsys = p->nsys;
    a.A = dA;
    a.X = dX;
    a.Wk = (T *)p->d_work;
    a.piv = p->d_piv;
    a.resid = p->d_resid;
    a.iters = p->d_iters;
    a.queue = p->d_queue;
    a.maxiter = p->maxiter;
    a.target = p->target;
    a.vstride = p->nmax * p->nrmax;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, p->ctx->stream, p->d_queue);
    hipLaunchKernelGGL((batch_lu_kernel<T, NT>), dim3(p->grid), dim3(NT), p->smem, p->ctx->stream, a);
    hipLaunchKernelGGL(batch_summary_kernel, dim3(1), dim3(256), 0, p->ctx->stream, p->d_resid, p->d_iters,
                       p->nsys, p->prec == LU2MI_PREC_FP32 ? p->target : 16.0, p->d_sum);
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

int batch_launch(lu2mi_batch *p, const double *dA, double *dX)
{
    if (p->prec == LU2MI_PREC_FP32)
        return p->nt == 512 ? batch_launch_t<float, 512>(p, dA, dX) : batch_launch_t<float, 256>(p, dA, dX);
    return p->nt == 512 ? batch_launch_t<double, 512>(p, dA, dX) : batch_launch_t<double, 256>(p, dA, dX);
}

void batch_drop_graph(lu2mi_batch *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->g_valid = 0;
}

int batch_timed(lu2mi_batch *p, const double *dA, double *dX, float *ms)
{
    LU2MI_CHECK(hipEventRecord(p->ev0, p->ctx->stream));
    LU2MI_TRY(batch_launch(p, dA, dX));
    LU2MI_CHECK(hipEventRecord(p->ev1, p->ctx->stream));
    LU2MI_CHECK(hipEventSynchronize(p->ev1));
    LU2MI_CHECK(hipEventElapsedTime(ms, p->ev0, p->ev1));
    return LU2MI_OK;
}
}  // namespace

extern "C" int lu2mi_batch_create(lu2mi_ctx *ctx, int nsys, const int *n, const int *nrhs, const long *a_off,
                                  const long *lda, const long *x_off, int prec, int maxiter, double target,
                                  lu2mi_batch **out)
{
    if (!ctx || !out || nsys < 0 || (nsys > 0 && (!n || !nrhs || !a_off || !lda || !x_off)) ||
        (prec != LU2MI_PREC_FP64 && prec != LU2MI_PREC_FP32) || maxiter < 0) {
        lu2mi_set_error_msg("lu2mi_batch_create: bad arguments");
        return LU2MI_ERR_ARG;
    }
    for (int s = 0; s < nsys; s++) {
        if (n[s] < 1 || n[s] > BNMAX || nrhs[s] < 1 || nrhs[s] > BRHSMAX || lda[s] < (long)n[s] + nrhs[s] ||
            a_off[s] < 0 || x_off[s] < 0) {
            lu2mi_set_error_msg("lu2mi_batch_create: system %d: need 1 <= n <= %d, 1 <= nrhs <= %d, "
                                "lda >= n + nrhs, offsets >= 0 (n=%d nrhs=%d lda=%ld)",
                                s, BNMAX, BRHSMAX, n[s], nrhs[s], lda[s]);
            return LU2MI_ERR_ARG;
        }
    }
    LU2MI_CHECK(hipSetDevice(ctx->device));
    lu2mi_batch *p = new lu2mi_batch();
    p->ctx = ctx;
    p->nsys = nsys;
    p->prec = prec;
    p->maxiter = maxiter;
    p->target = target;
    p->n.assign(n, n + nsys);
    p->nrhs.assign(nrhs, nrhs + nsys);
    p->a_off.assign(a_off, a_off + nsys);
    p->lda.assign(lda, lda + nsys);
    p->x_off.assign(x_off, x_off + nsys);
    p->nmax = 1;
    p->nrmax = 1;
    const size_t esz = prec == LU2MI_PREC_FP32 ? 4 : 8;
    long wtot = 0, ptot = 0;
    p->w_off.resize(nsys);
    p->p_off.resize(nsys);
    p->ldw.resize(nsys);
    for (int s = 0; s < nsys; s++) {
        p->nmax = std::max(p->nmax, n[s]);
        p->nrmax = std::max(p->nrmax, nrhs[s]);
        p->ldw[s] = (n[s] + nrhs[s] + 3) & ~3;   // 16-byte rows in FP32, 32-byte in FP64
        p->w_off[s] = wtot;
        wtot += (long)n[s] * p->ldw[s];
        wtot = (wtot + 63) & ~63L;              // each system starts on a 256-byte boundary
        p->p_off[s] = ptot;
        ptot += n[s];
    }
    p->piv_total = ptot;
    p->order.resize(nsys);
    std::iota(p->order.begin(), p->order.end(), 0);
    std::stable_sort(p->order.begin(), p->order.end(), [&](int x, int y) { return n[x] > n[y]; });
    p->nt = 256;
    p->wp_small = 16;
    p->wp_large = 32;
    p->split = 128;
    int rc = LU2MI_OK;
    auto fail = [&](int code) {
        lu2mi_batch_destroy(p);
        return code;
    };
    if (hipMalloc(&p->d_sys, sizeof(BatchSys) * std::max(1, nsys)) != hipSuccess ||
        hipMalloc(&p->d_work, esz * std::max(64L, wtot)) != hipSuccess ||
        hipMalloc(&p->d_piv, sizeof(int) * std::max(1L, ptot)) != hipSuccess ||
        hipMalloc(&p->d_resid, sizeof(double) * std::max(1, nsys)) != hipSuccess ||
        hipMalloc(&p->d_iters, sizeof(int) * std::max(1, nsys)) != hipSuccess ||
        hipMalloc(&p->d_queue, 256) != hipSuccess || hipMalloc(&p->d_sum, 4 * sizeof(double)) != hipSuccess) {
        lu2mi_set_error_msg("lu2mi_batch_create: out of device memory (%ld working elements)", wtot);
        return fail(LU2MI_ERR_NOMEM);
    }
    if (hipEventCreate(&p->ev0) != hipSuccess || hipEventCreate(&p->ev1) != hipSuccess) {
        lu2mi_set_error_msg("lu2mi_batch_create: hipEventCreate failed");
        return fail(LU2MI_ERR_HIP);
    }
    if ((rc = batch_upload(p)) != LU2MI_OK) return fail(rc);
    *out = p;
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_destroy(lu2mi_batch *p)
{
    if (!p) return LU2MI_OK;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    batch_drop_graph(p);
    void *bufs[] = {p->d_sys, p->d_work, p->d_piv, p->d_resid, p->d_iters, p->d_queue, p->d_sum};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    delete p;
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_generate(lu2mi_batch *p, double *dA, long base_seed)
{
    if (!p || (!dA && p->nsys)) {
        lu2mi_set_error_msg("lu2mi_batch_generate: bad arguments");
        return LU2MI_ERR_ARG;
    }
    if (!p->nsys) return LU2MI_OK;
    LU2MI_CHECK(hipSetDevice(p->ctx->device));
    hipLaunchKernelGGL(batch_generate_kernel, dim3(p->nsys, 64), dim3(256), 0, p->ctx->stream, p->d_sys, dA,
                       base_seed);
    LU2MI_CHECK(hipGetLastError());
    LU2MI_CHECK(hipStreamSynchronize(p->ctx->stream));
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_set_config(lu2mi_batch *p, int nt, int wp_small, int wp_large, int split)
{
    if (!p || (nt != 256 && nt != 512) || (wp_small != 16 && wp_small != 32) || (wp_large != 16 && wp_large != 32)) {
        lu2mi_set_error_msg("lu2mi_batch_set_config: nt must be 256 or 512, panel widths 16 or 32");
        return LU2MI_ERR_ARG;
    }
    p->nt = nt;
    p->wp_small = wp_small;
    p->wp_large = wp_large;
    p->split = split;
    return batch_upload(p);
}

extern "C" int lu2mi_batch_get_config(lu2mi_batch *p, int *out4)
{
    if (!p || !out4) return LU2MI_ERR_ARG;
    out4[0] = p->nt;
    out4[1] = p->wp_small;
    out4[2] = p->wp_large;
    out4[3] = p->split;
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_tune(lu2mi_batch *p, const double *dA, double *dX, int reps)
{
    if (!p || (p->nsys && (!dA || !dX))) {
        lu2mi_set_error_msg("lu2mi_batch_tune: bad arguments");
        return LU2MI_ERR_ARG;
    }
    if (!p->nsys) return LU2MI_OK;
    LU2MI_CHECK(hipSetDevice(p->ctx->device));
    reps = std::max(1, reps);
    struct Cand { int nt, ws, wl, split; };
    const Cand cands[] = {{256, 16, 16, BNMAX}, {256, 32, 32, 0}, {256, 16, 32, 128}, {256, 16, 32, 256},
                          {512, 16, 16, BNMAX}, {512, 32, 32, 0}, {512, 16, 32, 128}, {512, 16, 32, 256}};
    float best = 1e30f;
    Cand bc = cands[0];
    for (const Cand &c : cands) {
        LU2MI_TRY(lu2mi_batch_set_config(p, c.nt, c.ws, c.wl, c.split));
        float ms = 0.f, tot = 0.f;
        LU2MI_TRY(batch_timed(p, dA, dX, &ms));   // warm-up
        for (int r = 0; r < reps; r++) {
            LU2MI_TRY(batch_timed(p, dA, dX, &ms));
            tot += ms;
        }
        if (tot / reps < best) {
            best = tot / reps;
            bc = c;
        }
    }
    LU2MI_TRY(lu2mi_batch_set_config(p, bc.nt, bc.ws, bc.wl, bc.split));
    p->tuned = 1;
    p->last_ms = best;
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_solve(lu2mi_batch *p, const double *dA, double *dX, int flags, lu2mi_batch_info *info)
{
    if (!p || (p->nsys && (!dA || !dX))) {
        lu2mi_set_error_msg("lu2mi_batch_solve: bad arguments");
        return LU2MI_ERR_ARG;
    }
    LU2MI_CHECK(hipSetDevice(p->ctx->device));
    hipStream_t st = p->ctx->stream;
    float ms = 0.f;
    if (p->nsys) {
        if (flags & LU2MI_BATCH_EAGER) {
            LU2MI_TRY(batch_timed(p, dA, dX, &ms));
        } else {
            if (!p->g_valid || p->g_A != dA || p->g_X != dX) {
                batch_drop_graph(p);
                LU2MI_CHECK(hipStreamSynchronize(st));
                LU2MI_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                const int rc = batch_launch(p, dA, dX);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(st, &g);
                if (rc != LU2MI_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                LU2MI_CHECK(e);
                p->graph = g;
                LU2MI_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
                p->g_A = dA;
                p->g_X = dX;
                p->g_valid = 1;
            }
            LU2MI_CHECK(hipEventRecord(p->ev0, st));
            LU2MI_CHECK(hipGraphLaunch(p->exec, st));
            LU2MI_CHECK(hipEventRecord(p->ev1, st));
            LU2MI_CHECK(hipEventSynchronize(p->ev1));
            LU2MI_CHECK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
        }
    }
    p->last_ms = ms;
    if (info) {
        double sum[3] = {0.0, 0.0, 0.0};
        if (p->nsys) LU2MI_CHECK(hipMemcpy(sum, p->d_sum, sizeof(sum), hipMemcpyDeviceToHost));
        info->max_resid_hpl2 = sum[0];
        info->max_iterations = (int)sum[1];
        info->n_failed = (int)sum[2];
        info->t_ms = ms;
        info->nt = p->nt;
        info->wp_small = p->wp_small;
        info->wp_large = p->wp_large;
        info->split = p->split;
    }
    return LU2MI_OK;
}

extern "C" int lu2mi_batch_results(lu2mi_batch *p, double *resid, int *iters, int *piv)
{
    if (!p) return LU2MI_ERR_ARG;
    LU2MI_CHECK(hipSetDevice(p->ctx->device));
    LU2MI_CHECK(hipStreamSynchronize(p->ctx->stream));
    if (!p->nsys) return LU2MI_OK;
    if (resid) LU2MI_CHECK(hipMemcpy(resid, p->d_resid, sizeof(double) * p->nsys, hipMemcpyDeviceToHost));
    if (iters) LU2MI_CHECK(hipMemcpy(iters, p->d_iters, sizeof(int) * p->nsys, hipMemcpyDeviceToHost));
    if (piv) LU2MI_CHECK(hipMemcpy(piv, p->d_piv, sizeof(int) * p->piv_total, hipMemcpyDeviceToHost));
    return LU2MI_OK;
}
