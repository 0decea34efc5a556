// lu2mi_lowp_wide.hip -- the wide (256 x 256) tile of the 16-bit trailing
// update, in a file of its own: its waves keep 256 FP32 accumulators each,
// which live in the accumulation registers (the rest of the library is built
// with MFMA accumulators in ordinary VGPRs).
#include "lu2mi_lowp.h"

using namespace lu2mi_lowp;

namespace {
// Wide tile: C[M x N] -= 2^-(sa+sb) A16 B16t^T in 256 x 256 output tiles of
// eight waves (two per SIMD), each wave 128 x 64 (8 x 4 blocks of 16x16x32
// matrix-core ops, 128 FP32 accumulators per lane) inside one 128-row scaling
// block of L and one 128-column block of U, so each wave has one unscale
// factor.  A wave reads 3/256 byte of LDS operand per flop (a 64 x 64 wave
// tile reads 1/32, which at the FP16 rate is the whole LDS bandwidth), and the
// workgroup fetches half as many operand bytes per flop as a 128 x 128 tile.
// K is staged BKW halves at a time (64 for one operand plane, 32 for the
// split format's two) through two LDS stages, filled by direct 16-byte
// global-to-LDS loads issued behind the stage's fragment reads.  Tile row r
// keeps its 16-byte chunk c at slot CPR*r + (c ^ swz(r)) so that the 8 (and
// 16) lanes of a read phase hit different bank groups.
// Every accumulator gets the same matrix-core ops over the same K steps in
// the same order as in gemm16_kernel (lane l supplies k = 8(l/16)+e of each
// 32-step; FP16x2: hi*lo, lo*hi, then hi*hi per step), so the two tiles give
// identical bits.
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void gbl_void;
constexpr int TW = 256;   // wide output tile edge

template <int CPR>
__device__ __forceinline__ int wswz(int r)
{
    return (CPR == 8) ? ((r ^ (r >> 3)) & 7) : (((r >> 1) ^ (r >> 3)) & 3);
}

template <int FMT>
__global__ __launch_bounds__(512, 1) void gemm16w_kernel(const typename H16<FMT>::t *__restrict__ A16,
                                                         const typename H16<FMT>::t *__restrict__ B16t, int M,
                                                         int N, int K, long pa, long pb, const int *__restrict__ sa,
                                                         const int *__restrict__ sb, int nsa, int nsb,
                                                         float *__restrict__ C, long ldc)
{
    typedef typename H16<FMT>::t H;
    typedef typename H16<FMT>::v8 V8;
    constexpr int NP = H16<FMT>::NP;
    constexpr int BKW = (NP == 1) ? 64 : 32;   // K per stage
    constexpr int CPR = BKW / 8;               // 16-byte chunks per tile row
    constexpr int SL = TW * CPR;               // 16-byte slots per operand plane per stage
    constexpr int NTW = 512;                   // 8 waves: two per SIMD
    constexpr int IPW = SL / NTW;              // direct loads per lane per operand plane
    using namespace lu2mi_tiles;
    extern __shared__ __align__(16) char smem16[];
    // stage b, operand o (0 = A, 1 = B), plane p
    auto slots = [&](int b, int o, int p) { return (V8 *)smem16 + ((b * 2 + o) * NP + p) * SL; };

    const int tiles_m = (M + TW - 1) / TW, tiles_n = (N + TW - 1) / TW;
    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, tiles_m * tiles_n), tiles_m, tiles_n, tm, tn);
    const int m0 = tm * TW, n0 = tn * TW;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 2, wn = w & 3, l16 = lane & 15, lk = lane >> 4;   // wave tile: 128 rows x 64 columns

    // this lane's direct-load slots: p = 64 (IPW w + i) + lane -> row p/CPR,
    // chunk (p % CPR) ^ swz(row)
    // (element offset row * K + column, relative to the tile's first row)
    int loff[IPW];
#pragma unroll
    for (int i = 0; i < IPW; i++) {
        const int p = 64 * (IPW * w + i) + lane, r = p / CPR;
        loff[i] = r * K + 8 * ((p % CPR) ^ wswz<CPR>(r));
    }
    auto issue = [&](int kt, int b) {
        const int k0 = kt * BKW;
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int i = 0; i < IPW; i++) {
                const H *ga = A16 + p * pa + (long)m0 * K + k0 + loff[i];
                const H *gb = B16t + p * pb + (long)n0 * K + k0 + loff[i];
                __builtin_amdgcn_global_load_lds((gbl_void *)ga, (lds_void *)(slots(b, 0, p) + 64 * (IPW * w + i)),
                                                 16, 0, 0);
                __builtin_amdgcn_global_load_lds((gbl_void *)gb, (lds_void *)(slots(b, 1, p) + 64 * (IPW * w + i)),
                                                 16, 0, 0);
            }
    };

    f4_t acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};

    auto step = [&](int b, int ks) {   // one 32-deep matrix-core step of stage b
        const int c = 4 * ks + lk;      // this lane's chunk
        V8 af[NP][8], bf[NP][4];
#pragma unroll
        for (int p = 0; p < NP; p++) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int ra = wm * 128 + i * 16 + l16;
                af[p][i] = slots(b, 0, p)[CPR * ra + (c ^ wswz<CPR>(ra))];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int rb = wn * 64 + j * 16 + l16;
                bf[p][j] = slots(b, 1, p)[CPR * rb + (c ^ wswz<CPR>(rb))];
            }
        }
        if (NP == 2) {   // small terms first, as in gemm16_kernel
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[0][i], bf[NP - 1][j], acc[i][j]);
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[NP - 1][i], bf[0][j], acc[i][j]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[0][i], bf[0][j], acc[i][j]);
    };

    const int nk = K / BKW;
    issue(0, 0);
    for (int kt = 0; kt < nk; kt++) {
        const int b = kt & 1;
        __syncthreads();   // stage kt landed (each wave's loads complete first); stage kt-1 read
        if (kt + 1 < nk) issue(kt + 1, b ^ 1);
#pragma unroll
        for (int ks = 0; ks < BKW / 32; ks++) step(b, ks);
    }

    const int rb_i = min(2 * tm + wm, nsa - 1), cb_i = min(2 * tn + (wn >> 1), nsb - 1);
    const float scale = ldexpf(1.0f, -(sa[rb_i] + sb[cb_i]));
    const int r0 = m0 + wm * 128, c0 = n0 + wn * 64;
    const bool full = (r0 + 128 <= M) && (c0 + 64 <= N);
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = r0 + i * 16 + 4 * lk + r, col = c0 + j * 16 + l16;
                if (full || (row < M && col < N)) {
                    float *p = C + (long)row * ldc + col;
                    *p = fmaf(-scale, acc[i][j][r], *p);
                }
            }
}

}  // namespace

template <int FMT>
static int launch_wide(lu2mi_ctx *ctx, const void *A16, const void *B16t, int M, int N, int K, long pa, long pb,
                       const int *sa, const int *sb, int nsa, int nsb, float *C, long ldc)
{
    typedef typename H16<FMT>::t H;
    constexpr int NP = H16<FMT>::NP, BKW = (NP == 1) ? 64 : 32;
    const size_t shm = (size_t)2 * 2 * NP * TW * BKW * sizeof(H);   // two stages of A and B (128 KiB)
    static bool attr = false;
    if (!attr) {
        LU2MI_CHECK(hipFuncSetAttribute((const void *)gemm16w_kernel<FMT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)shm));
        attr = true;
    }
    const int tw_m = (M + TW - 1) / TW, tw_n = (N + TW - 1) / TW;
    hipLaunchKernelGGL(gemm16w_kernel<FMT>, dim3(tw_m * tw_n), dim3(512), shm, ctx->stream, (const H *)A16,
                       (const H *)B16t, M, N, K, pa, pb, sa, sb, nsa, nsb, C, ldc);
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

int lu2mi_gemm16w_launch(lu2mi_ctx *ctx, int fmt, const void *A16, const void *B16t, int M, int N, int K, long pa,
                         long pb, const int *sa, const int *sb, int nsa, int nsb, float *C, long ldc)
{
    if (fmt == LU2MI_LOWP_FP16)
        return launch_wide<LU2MI_LOWP_FP16>(ctx, A16, B16t, M, N, K, pa, pb, sa, sb, nsa, nsb, C, ldc);
    if (fmt == LU2MI_LOWP_BF16)
        return launch_wide<LU2MI_LOWP_BF16>(ctx, A16, B16t, M, N, K, pa, pb, sa, sb, nsa, nsb, C, ldc);
    return launch_wide<LU2MI_LOWP_FP16X2>(ctx, A16, B16t, M, N, K, pa, pb, sa, sb, nsa, nsb, C, ldc);
}
