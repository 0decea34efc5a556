// lu2mi_gemm.hip -- row-major GEMM on CDNA4 matrix cores.
//
// Replaces the trailing-update DGEMM that lu2 sends to GRAPE-DR or CBLAS
// (lu2.c:312-343 mmmulandsub -> lu2lib.c:2079-2135 mydgemm ->
// gdrdgemm.c:155-199 mygdrdgemm), and the U12 product of the explicit-inverse
// TRSM (lu2.c:511-512).  C = alpha*A*B + beta*C, all row-major with
// arbitrary leading dimensions, optionally batched.
//
// FP64 path: v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][k0+l/16] and
// B[k0+l/16][l%16]; the 16x16 result lands as D[l/16 + 4*r][l%16] in the
// lane's r-th double (checked on gfx950).
//
// Tiling: 256 threads (4 waves, 2x2), 128x128 C tile per workgroup, 64x64 per
// wave (4x4 MFMA blocks), K staged through LDS 16 at a time with a register
// double buffer.  The A tile keeps its row-major shape in LDS with one pad
// double per row so the 16 lanes of an MFMA operand read 16 different banks.
// Workgroups are remapped so that each XCD (blocks go round-robin over the 8
// XCDs) owns a contiguous run of tiles in GROUP_M-row bands, which keeps the
// shared A rows and B columns in that XCD's L2.
#include "lu2mi_internal.h"
#include "lu2mi_dev.h"

namespace {
constexpr int BM = 128, BN = 128, NT = 256;
}  // namespace
using namespace lu2mi_tiles;

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

template <typename T, bool VEC, int EPI, int TAG>
__global__ __launch_bounds__(NT) void gemm_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    constexpr int BKT = Mfma<T>::BK;
    __shared__ T As[2][BM * (BKT + 1)];
    __shared__ T Bs[2][BKT * BN];

    const int total = tiles_m * tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    int tm, tn;
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const long bz = blockIdx.y;
    const T *A = (const T *)g.A + bz * g.sA;
    const T *B = (const T *)g.B + bz * g.sB;
    T *C = (T *)g.C + bz * g.sC;
    const int K = g.a_lower ? min(g.K, m0 + BM) : g.K;
    const bool interior = VEC && (m0 + BM <= g.M) && (n0 + BN <= g.N) && (K % BKT == 0);
    if (interior) gemm_tile<T, false, EPI>(g, A, B, C, m0, n0, K, As, Bs);
    else gemm_tile<T, true, EPI>(g, A, B, C, m0, n0, K, As, Bs);
}

// ---------------------------------------------------------------------------
static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <typename T>
static int launch_gemm(const GemmDesc &g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return LU2MI_OK;
    if (g.K < 0 || g.lda < (g.K > 0 ? g.K : 1) || g.ldb < g.N || g.ldc < g.N) {
        lu2mi_set_error_msg("gemm: bad leading dimension (M=%d N=%d K=%d lda=%ld ldb=%ld ldc=%ld)",
                            g.M, g.N, g.K, g.lda, g.ldb, g.ldc);
        return LU2MI_ERR_ARG;
    }
    const int tiles_m = (g.M + BM - 1) / BM, tiles_n = (g.N + BN - 1) / BN;
    const long total = (long)tiles_m * tiles_n;
    if (total > 0x7fffffffL || g.batch > 65535) {
        lu2mi_set_error_msg("gemm: grid too large");
        return LU2MI_ERR_ARG;
    }
    // 16-byte vector loads need 16-byte aligned rows of A and B
    constexpr long EPV = Mfma<T>::EPV;
    const bool vec = aligned16(g.A) && aligned16(g.B) && (g.lda % EPV == 0) && (g.ldb % EPV == 0) &&
                     (g.sA % EPV == 0) && (g.sB % EPV == 0);
    const int epi = (g.alpha == -1.0 && g.beta == 1.0) ? EPI_SUB : (g.alpha == 1.0 && g.beta == 0.0) ? EPI_SET
                                                                                                   : EPI_GENERAL;
    dim3 grid((unsigned)total, (unsigned)g.batch), block(NT);
    if (vec && epi == EPI_SUB && g.tag == 1)
        hipLaunchKernelGGL((gemm_kernel<T, true, EPI_SUB, 1>), grid, block, 0, s, g, tiles_m, tiles_n);
    else if (vec && epi == EPI_SUB)
        hipLaunchKernelGGL((gemm_kernel<T, true, EPI_SUB, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
    else if (vec && epi == EPI_SET)
        hipLaunchKernelGGL((gemm_kernel<T, true, EPI_SET, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
    else if (vec)
        hipLaunchKernelGGL((gemm_kernel<T, true, EPI_GENERAL, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((gemm_kernel<T, false, EPI_GENERAL, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

int lu2mi_gemm_f64(const GemmDesc &g, hipStream_t s) { return launch_gemm<double>(g, s); }
int lu2mi_gemm_f32(const GemmDesc &g, hipStream_t s) { return launch_gemm<float>(g, s); }
