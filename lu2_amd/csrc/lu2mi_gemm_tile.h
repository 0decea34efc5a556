// lu2mi_gemm_tile.h -- the 128x128 MFMA tile of the row-major GEMM, shared by
// the stand-alone GEMM kernel and the persistent sub-panel kernel (which runs
// the same tiles between its grid barriers, so both produce identical bits).
//
// FP64 path: v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][k0+l/16] and
// B[k0+l/16][l%16]; the 16x16 result lands as D[l/16 + 4*r][l%16] in the
// lane's r-th double (checked on gfx950).
//
// Tiling: 256 threads (4 waves, 2x2), 128x128 C tile per workgroup, 64x64 per
// wave (4x4 MFMA blocks), K staged through LDS 16 at a time with a register
// double buffer.  The A tile keeps its row-major shape in LDS with one pad
// double per row so the 16 lanes of an MFMA operand read 16 different banks.
#pragma once
#include "lu2mi_internal.h"
#include "lu2mi_dev.h"

namespace lu2mi_gemm {
constexpr int BM = 128, BN = 128, NT = 256;

// ---------------------------------------------------------------------------
// Epilogue modes: EPI_GENERAL  C = alpha*acc + beta*C
//                 EPI_SUB      C = C - acc        (alpha = -1, beta = 1: the LU updates)
//                 EPI_SET      C = acc            (alpha = 1, beta = 0: the TRSM products)
enum { EPI_GENERAL = 0, EPI_SUB = 1, EPI_SET = 2 };

template <typename T, bool EDGE, int EPI>
__device__ __forceinline__ void gemm_tile(const GemmDesc &g, const T *__restrict__ A, const T *__restrict__ B,
                                          T *__restrict__ C, int m0, int n0, int K,
                                          T (*As)[BM * (Mfma<T>::BK + 1)], T (*Bs)[Mfma<T>::BK * BN])
{
    typedef Mfma<T> F;
    typedef typename F::vec_t V;
    constexpr int BKT = F::BK, EPV = F::EPV, AL = BKT + 1;
    constexpr int BCH = BN / EPV;                 // 16-byte chunks per B row
    const int M = g.M, N = g.N;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const long lda = g.lda, ldb = g.ldb;

    V ra[4], rb[4];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * EPV;
            const int br = idx / BCH, bc = (idx % BCH) * EPV;
            if (!EDGE) {
                ra[it] = *(const V *)(A + (long)(m0 + ar) * lda + k0 + ac);
                rb[it] = *(const V *)(B + (long)(k0 + br) * ldb + n0 + bc);
            } else {
                T va[EPV], vb[EPV];
                const int gr = m0 + ar, gk = k0 + br;
#pragma unroll
                for (int e = 0; e < EPV; e++) {
                    const int gc = k0 + ac + e, gn = n0 + bc + e;
                    va[e] = (gr < M && gc < K) ? A[(long)gr * lda + gc] : T(0);
                    vb[e] = (gk < K && gn < N) ? B[(long)gk * ldb + gn] : T(0);
                }
                ra[it] = *(V *)va;
                rb[it] = *(V *)vb;
            }
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * EPV;
            const T *pa = (const T *)&ra[it];
#pragma unroll
            for (int e = 0; e < EPV; e++) As[buf][ar * AL + ac + e] = pa[e];
            const int br = idx / BCH, bc = (idx % BCH) * EPV;
            *(V *)&Bs[buf][br * BN + bc] = rb[it];
        }
    };

    typename F::acc_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++) acc[mi][nj] = (typename F::acc_t){T(0), T(0), T(0), T(0)};

    auto compute = [&](int cur) {
        const T *as = As[cur];
        const T *bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < BKT; kk += 4) {
            T a[4], b[4];
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
                a[mi] = as[(wm * 64 + mi * 16 + (lane & 15)) * AL + kk + (lane >> 4)];
#pragma unroll
            for (int nj = 0; nj < 4; nj++)
                b[nj] = bs[(kk + (lane >> 4)) * BN + wn * 64 + nj * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
#pragma unroll
                for (int nj = 0; nj < 4; nj++) acc[mi][nj] = F::step(a[mi], b[nj], acc[mi][nj]);
        }
    };

    const int nk = (K + BKT - 1) / BKT;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
        // steady state: no conditionals around the loads
        for (int kt = 0; kt < nk - 1; kt++) {
            const int cur = kt & 1;
            load_tiles((kt + 1) * BKT);
            compute(cur);
            store_tiles(cur ^ 1);
            __syncthreads();
        }
        compute((nk - 1) & 1);
    }

    const T alpha = (T)g.alpha, beta = (T)g.beta;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + mi * 16 + F::out_row(lane, r);
                const int col = n0 + wn * 64 + nj * 16 + (lane & 15);
                if (!EDGE || (row < M && col < N)) {
                    T *cp = C + (long)row * g.ldc + col;
                    const T v = acc[mi][nj][r];
                    if (EPI == EPI_SUB) *cp = *cp - v;
                    else if (EPI == EPI_SET) *cp = v;
                    else *cp = (beta == T(0)) ? alpha * v : alpha * v + beta * *cp;
                }
            }
}


// ---------------------------------------------------------------------------
// 64x64 tile, 4 waves of 32x32 (2x2 MFMA blocks each).  Every element of C
// gets the same MFMA over the same K chunks in the same order as in
// gemm_tile (an MFMA output element depends only on its row of A, its column
// of B and its accumulator), so the two tiles give identical bits; it holds a
// quarter of the accumulators, for kernels that keep other state beside it.
// K is the K range of the 128x128 tile that contains this one.
constexpr int BT64 = 64;

template <typename T>
struct Tile64Smem {
    T As[2][BT64 * (Mfma<T>::BK + 1)];
    T Bs[2][Mfma<T>::BK * BT64];
};

// bscale (optional): row k of B is multiplied by bscale[k] as it is loaded,
// one rounding, as a separately scaled copy of B would hold it
template <typename T, bool EDGE, int EPI>
__device__ __forceinline__ void gemm_tile64(const GemmDesc &g, const T *__restrict__ A, const T *__restrict__ B,
                                            T *__restrict__ C, int m0, int n0, int K, Tile64Smem<T> &sm,
                                            const T *__restrict__ bscale = nullptr)
{
    typedef Mfma<T> F;
    typedef typename F::vec_t V;
    constexpr int BKT = F::BK, EPV = F::EPV, AL = BKT + 1;
    constexpr int ACH = BKT / EPV;                 // 16-byte chunks per A row
    constexpr int BCH = BT64 / EPV;                // 16-byte chunks per B row
    constexpr int LPT = BT64 * ACH / NT;           // chunks per thread (A and B alike)
    static_assert(BT64 * ACH == BKT * BCH && LPT * NT == BT64 * ACH, "tile64 load shape");
    const int M = g.M, N = g.N;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const long lda = g.lda, ldb = g.ldb;

    V ra[LPT], rb[LPT];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LPT; it++) {
            const int idx = t + it * NT;
            const int ar = idx / ACH, ac = (idx % ACH) * EPV;
            const int br = idx / BCH, bc = (idx % BCH) * EPV;
            if (!EDGE) {
                ra[it] = *(const V *)(A + (long)(m0 + ar) * lda + k0 + ac);
                rb[it] = *(const V *)(B + (long)(k0 + br) * ldb + n0 + bc);
                if (bscale) {
                    T *pb = (T *)&rb[it];
                    const T sc = bscale[k0 + br];
#pragma unroll
                    for (int e = 0; e < EPV; e++) pb[e] = pb[e] * sc;
                }
            } else {
                T va[EPV], vb[EPV];
                const int gr = m0 + ar, gk = k0 + br;
                const T sc = (bscale && gk < K) ? bscale[gk] : T(1);
#pragma unroll
                for (int e = 0; e < EPV; e++) {
                    const int gc = k0 + ac + e, gn = n0 + bc + e;
                    va[e] = (gr < M && gc < K) ? A[(long)gr * lda + gc] : T(0);
                    vb[e] = (gk < K && gn < N) ? (bscale ? B[(long)gk * ldb + gn] * sc : B[(long)gk * ldb + gn])
                                               : T(0);
                }
                ra[it] = *(V *)va;
                rb[it] = *(V *)vb;
            }
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LPT; it++) {
            const int idx = t + it * NT;
            const int ar = idx / ACH, ac = (idx % ACH) * EPV;
            const T *pa = (const T *)&ra[it];
#pragma unroll
            for (int e = 0; e < EPV; e++) sm.As[buf][ar * AL + ac + e] = pa[e];
            const int br = idx / BCH, bc = (idx % BCH) * EPV;
            *(V *)&sm.Bs[buf][br * BT64 + bc] = rb[it];
        }
    };

    typename F::acc_t acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++) acc[mi][nj] = (typename F::acc_t){T(0), T(0), T(0), T(0)};

    auto compute = [&](int cur) {
        const T *as = sm.As[cur];
        const T *bs = sm.Bs[cur];
#pragma unroll
        for (int kk = 0; kk < BKT; kk += 4) {
            T a[2], b[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++) a[mi] = as[(wm * 32 + mi * 16 + (lane & 15)) * AL + kk + (lane >> 4)];
#pragma unroll
            for (int nj = 0; nj < 2; nj++) b[nj] = bs[(kk + (lane >> 4)) * BT64 + wn * 32 + nj * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int nj = 0; nj < 2; nj++) acc[mi][nj] = F::step(a[mi], b[nj], acc[mi][nj]);
        }
    };

    const int nk = (K + BKT - 1) / BKT;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
        for (int kt = 0; kt < nk - 1; kt++) {
            const int cur = kt & 1;
            load_tiles((kt + 1) * BKT);
            compute(cur);
            store_tiles(cur ^ 1);
            __syncthreads();
        }
        compute((nk - 1) & 1);
    }

    const T alpha = (T)g.alpha, beta = (T)g.beta;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 32 + mi * 16 + F::out_row(lane, r);
                const int col = n0 + wn * 32 + nj * 16 + (lane & 15);
                if (!EDGE || (row < M && col < N)) {
                    T *cp = C + (long)row * g.ldc + col;
                    const T v = acc[mi][nj][r];
                    if (EPI == EPI_SUB) *cp = *cp - v;
                    else if (EPI == EPI_SET) *cp = v;
                    else *cp = (beta == T(0)) ? alpha * v : alpha * v + beta * *cp;
                }
            }
}

// LDS of one tile (double-buffered A and B stages)
template <typename T>
struct TileSmem {
    T As[2][BM * (Mfma<T>::BK + 1)];
    T Bs[2][Mfma<T>::BK * BN];
};
}  // namespace lu2mi_gemm
