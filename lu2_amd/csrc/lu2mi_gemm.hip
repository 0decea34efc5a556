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

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

namespace {
constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int ALD = BK + 1;
constexpr int GROUP_M = 8;
constexpr int NXCD = 8;

__device__ __forceinline__ int xcd_remap(int bid, int total)
{
    const int q = total / NXCD, r = total % NXCD;
    const int xcd = bid % NXCD, local = bid / NXCD;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int &tm, int &tn)
{
    const int width = GROUP_M * tiles_n;
    const int group = t / width;
    const int first_m = group * GROUP_M;
    const int gm = min(tiles_m - first_m, GROUP_M);
    const int in = t - group * width;
    tm = first_m + in % gm;
    tn = in / gm;
}
}  // namespace

// ---------------------------------------------------------------------------
// Epilogue modes: EPI_GENERAL  C = alpha*acc + beta*C
//                 EPI_SUB      C = C - acc        (alpha = -1, beta = 1: the LU updates)
//                 EPI_SET      C = acc            (alpha = 1, beta = 0: the TRSM products)
enum { EPI_GENERAL = 0, EPI_SUB = 1, EPI_SET = 2 };

template <bool EDGE, int EPI>
__device__ __forceinline__ void gemm_f64_tile(const GemmDesc &g, const double *__restrict__ A,
                                              const double *__restrict__ B, double *__restrict__ C,
                                              int m0, int n0, int K, double (*As)[BM * ALD],
                                              double (*Bs)[BK * BN])
{
    const int M = g.M, N = g.N;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const long lda = g.lda, ldb = g.ldb;

    double2 ra[4], rb[4];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * 2;
            const int br = idx >> 6, bc = (idx & 63) * 2;
            if (!EDGE) {
                ra[it] = *(const double2 *)(A + (long)(m0 + ar) * lda + k0 + ac);
                rb[it] = *(const double2 *)(B + (long)(k0 + br) * ldb + n0 + bc);
            } else {
                const int gr = m0 + ar, gc = k0 + ac;
                double2 v = {0.0, 0.0};
                if (gr < M) {
                    const double *p = A + (long)gr * lda + gc;
                    if (gc < K) v.x = p[0];
                    if (gc + 1 < K) v.y = p[1];
                }
                ra[it] = v;
                const int gk = k0 + br, gn = n0 + bc;
                double2 u = {0.0, 0.0};
                if (gk < K) {
                    const double *q = B + (long)gk * ldb + gn;
                    if (gn < N) u.x = q[0];
                    if (gn + 1 < N) u.y = q[1];
                }
                rb[it] = u;
            }
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * 2;
            As[buf][ar * ALD + ac] = ra[it].x;
            As[buf][ar * ALD + ac + 1] = ra[it].y;
            const int br = idx >> 6, bc = (idx & 63) * 2;
            *(double2 *)&Bs[buf][br * BN + bc] = rb[it];
        }
    };

    d4_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++) acc[mi][nj] = (d4_t){0.0, 0.0, 0.0, 0.0};

    auto compute = [&](int cur) {
        const double *as = As[cur];
        const double *bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a[4], b[4];
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
                a[mi] = as[(wm * 64 + mi * 16 + (lane & 15)) * ALD + kk + (lane >> 4)];
#pragma unroll
            for (int nj = 0; nj < 4; nj++)
                b[nj] = bs[(kk + (lane >> 4)) * BN + wn * 64 + nj * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
#pragma unroll
                for (int nj = 0; nj < 4; nj++)
                    acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi], b[nj], acc[mi][nj], 0, 0, 0);
        }
    };

    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
        // steady state: no conditionals around the loads
        for (int kt = 0; kt < nk - 1; kt++) {
            const int cur = kt & 1;
            load_tiles((kt + 1) * BK);
            compute(cur);
            store_tiles(cur ^ 1);
            __syncthreads();
        }
        compute((nk - 1) & 1);
    }

    const double alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + mi * 16 + (lane >> 4) + 4 * r;
                const int col = n0 + wn * 64 + nj * 16 + (lane & 15);
                if (!EDGE || (row < M && col < N)) {
                    double *cp = C + (long)row * g.ldc + col;
                    const double v = acc[mi][nj][r];
                    if (EPI == EPI_SUB) *cp = *cp - v;
                    else if (EPI == EPI_SET) *cp = v;
                    else *cp = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * *cp);
                }
            }
}

template <bool VEC, int EPI, int TAG>
__global__ __launch_bounds__(NT) void gemm_f64_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    __shared__ double As[2][BM * ALD];
    __shared__ double Bs[2][BK * BN];

    const int total = tiles_m * tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    int tm, tn;
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const long bz = blockIdx.y;
    const double *A = (const double *)g.A + bz * g.sA;
    const double *B = (const double *)g.B + bz * g.sB;
    double *C = (double *)g.C + bz * g.sC;
    const int K = g.a_lower ? min(g.K, m0 + BM) : g.K;
    const bool interior = VEC && (m0 + BM <= g.M) && (n0 + BN <= g.N) && (K % BK == 0);
    if (interior) gemm_f64_tile<false, EPI>(g, A, B, C, m0, n0, K, As, Bs);
    else gemm_f64_tile<true, EPI>(g, A, B, C, m0, n0, K, As, Bs);
}

// ---------------------------------------------------------------------------
// FP32 path: v_mfma_f32_16x16x4f32 with FP32 accumulation.  Operand layout as
// the FP64 instruction; the result is D[4*(l/16) + r][l%16].
template <bool VEC>
__global__ __launch_bounds__(NT) void gemm_f32_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    __shared__ float As[2][BM * ALD];
    __shared__ float Bs[2][BK * BN];

    const int total = tiles_m * tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    int tm, tn;
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const long bz = blockIdx.y;
    const float *__restrict__ A = (const float *)g.A + bz * g.sA;
    const float *__restrict__ B = (const float *)g.B + bz * g.sB;
    float *__restrict__ C = (float *)g.C + bz * g.sC;
    const int M = g.M, N = g.N;
    const int K = g.a_lower ? min(g.K, m0 + BM) : g.K;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;

    float2 ra[4], rb[4];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * 2;
            const int gr = m0 + ar, gc = k0 + ac;
            float2 v = {0.f, 0.f};
            if (gr < M) {
                const float *p = A + (long)gr * g.lda + gc;
                if (VEC && gc + 1 < K) {
                    v = *(const float2 *)p;
                } else {
                    if (gc < K) v.x = p[0];
                    if (gc + 1 < K) v.y = p[1];
                }
            }
            ra[it] = v;
            const int br = idx >> 6, bc = (idx & 63) * 2;
            const int gk = k0 + br, gn = n0 + bc;
            float2 u = {0.f, 0.f};
            if (gk < K) {
                const float *q = B + (long)gk * g.ldb + gn;
                if (VEC && gn + 1 < N) {
                    u = *(const float2 *)q;
                } else {
                    if (gn < N) u.x = q[0];
                    if (gn + 1 < N) u.y = q[1];
                }
            }
            rb[it] = u;
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int idx = t + it * NT;
            const int ar = idx >> 3, ac = (idx & 7) * 2;
            As[buf][ar * ALD + ac] = ra[it].x;
            As[buf][ar * ALD + ac + 1] = ra[it].y;
            const int br = idx >> 6, bc = (idx & 63) * 2;
            *(float2 *)&Bs[buf][br * BN + bc] = rb[it];
        }
    };

    f4_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++) acc[mi][nj] = (f4_t){0.f, 0.f, 0.f, 0.f};

    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt++) {
            const int cur = kt & 1;
            if (kt + 1 < nk) load_tiles((kt + 1) * BK);
            const float *as = As[cur];
            const float *bs = Bs[cur];
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                float a[4], b[4];
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
                    a[mi] = as[(wm * 64 + mi * 16 + (lane & 15)) * ALD + kk + (lane >> 4)];
#pragma unroll
                for (int nj = 0; nj < 4; nj++)
                    b[nj] = bs[(kk + (lane >> 4)) * BN + wn * 64 + nj * 16 + (lane & 15)];
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
#pragma unroll
                    for (int nj = 0; nj < 4; nj++)
                        acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi], b[nj],
                                                                          acc[mi][nj], 0, 0, 0);
            }
            if (kt + 1 < nk) store_tiles(cur ^ 1);
            __syncthreads();
        }
    }

    const float alpha = (float)g.alpha, beta = (float)g.beta;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + mi * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wn * 64 + nj * 16 + (lane & 15);
                if (row < M && col < N) {
                    float *cp = C + (long)row * g.ldc + col;
                    const float v = acc[mi][nj][r];
                    *cp = (beta == 0.f) ? alpha * v : fmaf(alpha, v, beta * *cp);
                }
            }
}

// ---------------------------------------------------------------------------
static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <typename T>
static int launch_gemm(const GemmDesc &g, hipStream_t s, bool is_f64)
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
    const size_t esz = sizeof(T);
    (void)esz;
    const bool vec = aligned16(g.A) && aligned16(g.B) && (g.lda % 2 == 0) && (g.ldb % 2 == 0) &&
                     (g.sA % 2 == 0) && (g.sB % 2 == 0);
    dim3 grid((unsigned)total, (unsigned)g.batch), block(NT);
    if (is_f64) {
        const int epi = (g.alpha == -1.0 && g.beta == 1.0) ? EPI_SUB
                        : (g.alpha == 1.0 && g.beta == 0.0) ? EPI_SET : EPI_GENERAL;
        if (vec && epi == EPI_SUB && g.tag == 1)
            hipLaunchKernelGGL((gemm_f64_kernel<true, EPI_SUB, 1>), grid, block, 0, s, g, tiles_m, tiles_n);
        else if (vec && epi == EPI_SUB)
            hipLaunchKernelGGL((gemm_f64_kernel<true, EPI_SUB, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
        else if (vec && epi == EPI_SET)
            hipLaunchKernelGGL((gemm_f64_kernel<true, EPI_SET, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
        else if (vec)
            hipLaunchKernelGGL((gemm_f64_kernel<true, EPI_GENERAL, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
        else
            hipLaunchKernelGGL((gemm_f64_kernel<false, EPI_GENERAL, 0>), grid, block, 0, s, g, tiles_m, tiles_n);
    } else {
        const bool vecf = ((uintptr_t)g.A % 8 == 0) && ((uintptr_t)g.B % 8 == 0) &&
                          (g.lda % 2 == 0) && (g.ldb % 2 == 0) && (g.sA % 2 == 0) && (g.sB % 2 == 0);
        if (vecf) hipLaunchKernelGGL(gemm_f32_kernel<true>, grid, block, 0, s, g, tiles_m, tiles_n);
        else hipLaunchKernelGGL(gemm_f32_kernel<false>, grid, block, 0, s, g, tiles_m, tiles_n);
    }
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

int lu2mi_gemm_f64(const GemmDesc &g, hipStream_t s) { return launch_gemm<double>(g, s, true); }
int lu2mi_gemm_f32(const GemmDesc &g, hipStream_t s) { return launch_gemm<float>(g, s, false); }
