// lu2mi_gemm.hip -- row-major GEMM on CDNA4 matrix cores.
//
// Replaces the trailing-update DGEMM that lu2 sends to GRAPE-DR or CBLAS
// (lu2.c:312-343 mmmulandsub -> lu2lib.c:2079-2135 mydgemm ->
// gdrdgemm.c:155-199 mygdrdgemm), and the U12 product of the explicit-inverse
// TRSM (lu2.c:511-512).  C = alpha*A*B + beta*C, all row-major with
// arbitrary leading dimensions, optionally batched.
//
// The tile itself (MFMA layout, LDS staging) is in lu2mi_gemm_tile.h.
// Workgroups are remapped so that each XCD (blocks go round-robin over the 8
// XCDs) owns a contiguous run of tiles in GROUP_M-row bands, which keeps the
// shared A rows and B columns in that XCD's L2.
#include "lu2mi_gemm_tile.h"

using namespace lu2mi_tiles;
using namespace lu2mi_gemm;

// Narrow products (M or N <= 64: the recursion nodes near the leaves, their
// TRSM and inverse blocks) and products too small to fill the device in
// 128x128 tiles run in 64x64 tiles, which waste less of the matrix cores on
// padding and spread over more workgroups; every element keeps the K range
// of the 128x128 tile that contains it, so the results are the 128x128
// kernel's bits.
template <typename T, int EPI>
__global__ __launch_bounds__(NT) void gemm64_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    __shared__ Tile64Smem<T> sm;
    const int m0 = (blockIdx.x / tiles_n) * BT64, n0 = (blockIdx.x % tiles_n) * BT64;
    const long bz = blockIdx.y;
    const T *A = (const T *)g.A + bz * g.sA;
    const T *B = (const T *)g.B + bz * g.sB;
    T *C = (T *)g.C + bz * g.sC;
    const int K = g.a_lower ? min(g.K, (m0 & ~(BM - 1)) + BM) : g.K;
    const bool interior = (m0 + BT64 <= g.M) && (n0 + BT64 <= g.N) && (K % Mfma<T>::BK == 0);
    if (interior) gemm_tile64<T, false, EPI>(g, A, B, C, m0, n0, K, sm);
    else gemm_tile64<T, true, EPI>(g, A, B, C, m0, n0, K, sm);
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

// The trailing update (tag 1, FP64) in the direct-to-LDS tile of
// lu2mi_gemm2; edge tiles run the same arithmetic with bounds-checked loads.
template <int EPI>
__global__ __launch_bounds__(lu2mi_gemm2::NT, 2) void gemm2_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    namespace G2 = lu2mi_gemm2;
    __shared__ G2::Smem sm;
    const int total = tiles_m * tiles_n;
    const int tile_id = xcd_remap(blockIdx.x, total);
    int tm, tn;
    tile_coords(tile_id, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * G2::BM, n0 = tn * G2::BN;
    const double *A = (const double *)g.A;
    const double *B = (const double *)g.B;
    double *C = (double *)g.C;
    const bool interior = (m0 + G2::BM <= g.M) && (n0 + G2::BN <= g.N) && (g.K % G2::BK == 0);
    if (interior) G2::tile<false, EPI>(g, A, B, C, m0, n0, g.K, sm);
    else G2::tile<true, EPI>(g, A, B, C, m0, n0, g.K, sm);
}

// The U12 product C = Linv * rows (lu2.c:511-512): A lower triangular, so row
// tile r stops at the diagonal and carries r+1 units of K work.  Workgroups
// are placed on the CUs in a fixed rotation, so tiles of unequal length leave
// CUs idle (a 2048-row product took as long as with the whole K range): each
// workgroup here takes row tiles p and T-1-p of one column (T+1 units
// whatever p), the grid covering ceil(T/2) row pairs.  The tiles are the
// single-tile kernels' (same bits).
template <typename T>
__global__ __launch_bounds__(NT) void gemm_lower_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    constexpr int BKT = Mfma<T>::BK;
    __shared__ T As[2][BM * (BKT + 1)];
    __shared__ T Bs[2][BKT * BN];
    const int slots_m = (tiles_m + 1) / 2;
    int tp, tn;
    tile_coords(xcd_remap(blockIdx.x, slots_m * tiles_n), slots_m, tiles_n, tp, tn);
    const T *A = (const T *)g.A;
    const T *B = (const T *)g.B;
    T *C = (T *)g.C;
    const int n0 = tn * BN;
    auto one = [&](int tm) {
        const int m0 = tm * BM;
        const int K = min(g.K, m0 + BM);
        if ((m0 + BM <= g.M) && (n0 + BN <= g.N) && (K % BKT == 0))
            gemm_tile<T, false, EPI_SET>(g, A, B, C, m0, n0, K, As, Bs);
        else
            gemm_tile<T, true, EPI_SET>(g, A, B, C, m0, n0, K, As, Bs);
    };
    one(tp);
    if (tiles_m - 1 - tp != tp) {
        __syncthreads();   // every wave is done with the first tile's LDS
        one(tiles_m - 1 - tp);
    }
}

__global__ __launch_bounds__(lu2mi_gemm2::NT, 2) void gemm2_lower_kernel(GemmDesc g, int tiles_m, int tiles_n)
{
    namespace G2 = lu2mi_gemm2;
    __shared__ G2::Smem sm;
    const int slots_m = (tiles_m + 1) / 2;
    int tp, tn;
    tile_coords(xcd_remap(blockIdx.x, slots_m * tiles_n), slots_m, tiles_n, tp, tn);
    const double *A = (const double *)g.A;
    const double *B = (const double *)g.B;
    double *C = (double *)g.C;
    const int n0 = tn * G2::BN;
    auto one = [&](int tm) {
        const int m0 = tm * G2::BM;
        const int K = min(g.K, m0 + G2::BM);
        if ((m0 + G2::BM <= g.M) && (n0 + G2::BN <= g.N) && (K % G2::BK == 0))
            G2::tile<false, EPI_SET>(g, A, B, C, m0, n0, K, sm);
        else
            G2::tile<true, EPI_SET>(g, A, B, C, m0, n0, K, sm);
    };
    one(tp);
    if (tiles_m - 1 - tp != tp) {
        __syncthreads();   // every wave is done with the first tile's LDS
        one(tiles_m - 1 - tp);
    }
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
    // 64x64 tiles for narrow products, and for any other product whose
    // 128x128 grid would not give every CU eight such workgroups (four rounds
    // at two per CU; per n=65536 factorization 13 ms faster with a two-round
    // threshold than with one, and another ~10 ms with four)
    static int num_cu = 0;
    if (!num_cu) {
        int dev = 0;
        LU2MI_CHECK(hipGetDevice(&dev));
        LU2MI_CHECK(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const bool underfilled = total * g.batch < 8L * num_cu;
    if (vec && g.tag == 0 && (g.M <= BT64 || g.N <= BT64 || underfilled)) {
        const int tm64 = (g.M + BT64 - 1) / BT64, tn64 = (g.N + BT64 - 1) / BT64;
        dim3 grid64((unsigned)(tm64 * tn64), (unsigned)g.batch), block64(NT);
        if (epi == EPI_SUB)
            hipLaunchKernelGGL((gemm64_kernel<T, EPI_SUB>), grid64, block64, 0, s, g, tm64, tn64);
        else if (epi == EPI_SET)
            hipLaunchKernelGGL((gemm64_kernel<T, EPI_SET>), grid64, block64, 0, s, g, tm64, tn64);
        else
            hipLaunchKernelGGL((gemm64_kernel<T, EPI_GENERAL>), grid64, block64, 0, s, g, tm64, tn64);
        LU2MI_CHECK(hipGetLastError());
        return LU2MI_OK;
    }
    if (g.a_lower && epi == EPI_SET && vec && g.batch == 1) {
        // the U12 product: pairs of row tiles per workgroup
        dim3 gridp((unsigned)((long)((tiles_m + 1) / 2) * tiles_n)), blockp(NT);
        if (sizeof(T) == 8 && g.tag == 1 && aligned16(g.C) && g.ldc % 2 == 0)
            hipLaunchKernelGGL(gemm2_lower_kernel, gridp, blockp, 0, s, g, tiles_m, tiles_n);
        else
            hipLaunchKernelGGL(gemm_lower_kernel<T>, gridp, blockp, 0, s, g, tiles_m, tiles_n);
        LU2MI_CHECK(hipGetLastError());
        return LU2MI_OK;
    }
    dim3 grid((unsigned)total, (unsigned)g.batch), block(NT);
    if (sizeof(T) == 8 && g.tag == 1 && epi == EPI_SUB && vec && g.batch == 1 &&
        !g.a_lower && aligned16(g.C) && g.ldc % 2 == 0) {
        hipLaunchKernelGGL(gemm2_kernel<EPI_SUB>, grid, block, 0, s, g, tiles_m, tiles_n);
        LU2MI_CHECK(hipGetLastError());
        return LU2MI_OK;
    }
    if (vec && epi == EPI_SUB && g.tag != 0)
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
