// lu2mi_lowp.hip -- the trailing update A22 -= L21 U12 with 16-bit operands
// (FP16 or BF16) on the matrix cores, FP32 accumulation, per-block
// power-of-two scaling.  The working matrix stays FP32; only the update's
// operands are rounded to 16 bits.  This is the low-precision form of
// lu2.c:580 (mmmulandsub -> mydgemm) for the mixed-precision solve.
//
// Per update (M x N output, K = nb):
//   scale_rows16: for every 128-row block of L21 a power of two 2^s puts the
//     block's largest |entry| in [2^13, 2^14), far from FP16's overflow
//     (65504) and from its subnormals; the block is scaled and rounded to
//     16 bits (A16, M_pad x K, rows past M zero)
//   scale_cols16: the same for every 128-column block of U12, written
//     transposed (B16t, N_pad x K) so both operands are K-contiguous
//   gemm16: 128x128 output tiles, FP32 accumulators; C -= 2^-(sa+sb) acc,
//     exact since the scale is a power of two
// FP16x2 splits every scaled operand v into hi = fp16(v) and lo = fp16(v-hi)
// (v-hi is exact in FP32), about 22 significant bits, and accumulates
// hi*hi + hi*lo + lo*hi: close to FP32 operands at FP16 matrix-core rates.
#include "lu2mi_lowp.h"

using namespace lu2mi_lowp;

namespace {
// power-of-two exponent putting max in [2^13, 2^14); 0 for zero or non-finite
__device__ __forceinline__ int block_shift(float m)
{
    if (!(m > 0.0f) || !(m <= 3.0e38f)) return 0;
    int e;
    frexpf(m, &e);   // m = f 2^e, f in [0.5, 1)
    return 14 - e;
}

__device__ __forceinline__ float block_max(float m, float *red)
{
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    return m;
}
}  // namespace

// one workgroup per 128-row block of L (M x K, row-major, leading dimension ldl)
template <int FMT>
__global__ __launch_bounds__(256) void scale_rows16_kernel(const float *__restrict__ L, long ldl, int M, int K,
                                                           typename H16<FMT>::t *__restrict__ A16, long plane,
                                                           int *__restrict__ sa)
{
    typedef typename H16<FMT>::t H;
    __shared__ float red[4];
    const int rb = blockIdx.x, t = threadIdx.x;
    const int r0 = rb * TB, rows = min(TB, M - r0);
    float m = 0.0f;
    for (long idx = t; idx < (long)rows * K; idx += 256) {
        const int r = (int)(idx / K), k = (int)(idx - (long)r * K);
        m = fmaxf(m, fabsf(L[(long)(r0 + r) * ldl + k]));
    }
    const int s = block_shift(block_max(m, red));
    if (t == 0) sa[rb] = s;
    for (long idx = t; idx < (long)TB * K; idx += 256) {
        const int r = (int)(idx / K), k = (int)(idx - (long)r * K);
        const float v = (r < rows) ? ldexpf(L[(long)(r0 + r) * ldl + k], s) : 0.0f;
        const H hi = (H)v;
        A16[(long)(r0 + r) * K + k] = hi;
        if (H16<FMT>::NP == 2) A16[plane + (long)(r0 + r) * K + k] = (H)(v - (float)hi);
    }
}

// one workgroup per 128-column block of U (K x N, row-major, leading
// dimension ldu); output transposed: B16t[c][k].  K is a multiple of 64.
template <int FMT>
__global__ __launch_bounds__(256) void scale_cols16_kernel(const float *__restrict__ U, long ldu, int K, int N,
                                                           typename H16<FMT>::t *__restrict__ B16t, long plane,
                                                           int *__restrict__ sb)
{
    typedef typename H16<FMT>::t H;
    typedef typename H16<FMT>::v8 V8;
    __shared__ float red[4];
    constexpr int NP = H16<FMT>::NP;
    __shared__ H tile[NP][64][TB + 8];
    const int cb = blockIdx.x, t = threadIdx.x;
    const int c0 = cb * TB, cols = min(TB, N - c0);
    float m = 0.0f;
    for (long idx = t; idx < (long)K * TB; idx += 256) {
        const int k = (int)(idx / TB), c = (int)(idx - (long)k * TB);
        if (c < cols) m = fmaxf(m, fabsf(U[(long)k * ldu + c0 + c]));
    }
    const int s = block_shift(block_max(m, red));
    if (t == 0) sb[cb] = s;
    for (int k0 = 0; k0 < K; k0 += 64) {
        for (int idx = t; idx < 64 * TB; idx += 256) {
            const int kk = idx / TB, c = idx - kk * TB;
            const float v = (c < cols) ? ldexpf(U[(long)(k0 + kk) * ldu + c0 + c], s) : 0.0f;
            const H hi = (H)v;
            tile[0][kk][c] = hi;
            if (NP == 2) tile[NP - 1][kk][c] = (H)(v - (float)hi);
        }
        __syncthreads();
        for (int q = t; q < NP * TB * 8; q += 256) {
            const int p = q / (TB * 8), qq = q - p * TB * 8;
            const int c = qq >> 3, kq = (qq & 7) * 8;
            V8 v;
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = tile[p][kq + e][c];
            *(V8 *)(B16t + p * plane + (long)(c0 + c) * K + k0 + kq) = v;
        }
        __syncthreads();
    }
}

// C[M x N] -= 2^-(sa[tm]+sb[tn]) * A16 B16t^T; A16 and B16t padded to whole
// 128-row tiles.  256 threads, 2x2 waves of 64x64, 16x16x32 matrix-core
// blocks.  K is staged 64 at a time through LDS; the next stage's global
// loads are in flight in registers while the current stage computes.  One
// operand plane keeps two LDS buffers (one barrier per stage); the split
// format keeps one (two barriers per stage) so that two workgroups still fit
// a CU.  Lane l supplies A[l%16][8(l/16)+e] and B[8(l/16)+e][l%16]: the same K
// assignment in both operands, so the product is the GEMM's whatever order
// the instruction sums in; the result lands at row 4(l/16)+r, column l%16.
template <int FMT>
__global__ __launch_bounds__(256, 2) void gemm16_kernel(const typename H16<FMT>::t *__restrict__ A16,
                                                        const typename H16<FMT>::t *__restrict__ B16t, int M, int N,
                                                        int K, long pa, long pb, const int *__restrict__ sa,
                                                        const int *__restrict__ sb, float *__restrict__ C, long ldc)
{
    typedef typename H16<FMT>::t H;
    typedef typename H16<FMT>::v8 V8;
    constexpr int NP = H16<FMT>::NP;
    constexpr int NBUF = (NP == 1) ? 2 : 1;
    constexpr int CPR = BKH / 8;              // 16-byte chunks per staged row
    constexpr int LPT = TB * CPR / 256;       // chunk loads per thread per operand plane
    using namespace lu2mi_tiles;
    __shared__ __align__(16) H As[NBUF][NP][TB][BKH + PADH];
    __shared__ __align__(16) H Bs[NBUF][NP][TB][BKH + PADH];
    const int tiles_m = (M + TB - 1) / TB, tiles_n = (N + TB - 1) / TB;
    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, tiles_m * tiles_n), tiles_m, tiles_n, tm, tn);
    const int m0 = tm * TB, n0 = tn * TB;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1, l16 = lane & 15, lk = lane >> 4;

    V8 ra[NP][LPT], rb[NP][LPT];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int it = 0; it < LPT; it++) {
                const int q = t + it * 256, row = q / CPR, part = (q % CPR) * 8;
                ra[p][it] = *(const V8 *)(A16 + p * pa + (long)(m0 + row) * K + k0 + part);
                rb[p][it] = *(const V8 *)(B16t + p * pb + (long)(n0 + row) * K + k0 + part);
            }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int it = 0; it < LPT; it++) {
                const int q = t + it * 256, row = q / CPR, part = (q % CPR) * 8;
                *(V8 *)&As[buf][p][row][part] = ra[p][it];
                *(V8 *)&Bs[buf][p][row][part] = rb[p][it];
            }
    };
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};

    const int nk = K / BKH;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = (NBUF == 2) ? (kt & 1) : 0;
        if (kt + 1 < nk) load((kt + 1) * BKH);
#pragma unroll
        for (int ks = 0; ks < BKH; ks += 32) {
            V8 af[NP][4], bf[NP][4];
#pragma unroll
            for (int p = 0; p < NP; p++) {
#pragma unroll
                for (int i = 0; i < 4; i++) af[p][i] = *(const V8 *)&As[cur][p][wm * 64 + i * 16 + l16][ks + lk * 8];
#pragma unroll
                for (int j = 0; j < 4; j++) bf[p][j] = *(const V8 *)&Bs[cur][p][wn * 64 + j * 16 + l16][ks + lk * 8];
            }
            // one product over all 16 accumulators before the next: consecutive
            // matrix-core ops never wait on each other's accumulator
            if (NP == 2) {   // small terms first
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[0][i], bf[NP - 1][j], acc[i][j]);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[NP - 1][i], bf[0][j], acc[i][j]);
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = H16<FMT>::mfma(af[0][i], bf[0][j], acc[i][j]);
        }
        if (kt + 1 < nk) {
            if (NBUF == 1) __syncthreads();   // every wave is done reading the single buffer
            store((NBUF == 2) ? (cur ^ 1) : 0);
            __syncthreads();
        }
    }
    const float scale = ldexpf(1.0f, -(sa[tm] + sb[tn]));
    const bool full = (m0 + TB <= M) && (n0 + TB <= N);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + i * 16 + 4 * lk + r, col = n0 + wn * 64 + j * 16 + l16;
                if (full || (row < M && col < N)) {
                    float *p = C + (long)row * ldc + col;
                    *p = fmaf(-scale, acc[i][j][r], *p);
                }
            }
}

namespace {
template <int FMT>
int launch16(lu2mi_ctx *ctx, int M, int N, int K, const float *L, long ldl, const float *U, long ldu, float *C,
              long ldc, int tiles_m, int tiles_n, int *sa, int *sb)
{
    typedef typename H16<FMT>::t H;
    const long pa = (long)(tiles_m + 1) * TB * K, pb = (long)(tiles_n + 1) * TB * K;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(scale_rows16_kernel<FMT>, dim3(tiles_m), dim3(256), 0, st, L, ldl, M, K, (H *)ctx->m_a16, pa,
                       sa);
    hipLaunchKernelGGL(scale_cols16_kernel<FMT>, dim3(tiles_n), dim3(256), 0, st, U, ldu, K, N, (H *)ctx->m_b16, pb,
                       sb);
    const long tw = (long)((M + 255) / 256) * ((N + 255) / 256);
    if (ctx->wide16 && tw >= ctx->num_cu) {
        // the wide tile, when it still gives every CU a workgroup (same bits)
        return lu2mi_gemm16w_launch(ctx, FMT, ctx->m_a16, ctx->m_b16, M, N, K, pa, pb, sa, sb, tiles_m, tiles_n, C,
                                    ldc);
    }
    hipLaunchKernelGGL(gemm16_kernel<FMT>, dim3(tiles_m * tiles_n), dim3(256), 0, st, (const H *)ctx->m_a16,
                       (const H *)ctx->m_b16, M, N, K, pa, pb, sa, sb, C, ldc);
    return LU2MI_OK;
}
}  // namespace

// C -= L U with 16-bit operands; L: M x K (ldl), U: K x N (ldu), C: M x N (ldc),
// all FP32 row-major on the context's device.  K a multiple of 64.
int lu2mi_update16(lu2mi_ctx *ctx, int fmt, int M, int N, int K, const float *L, long ldl, const float *U, long ldu,
                   float *C, long ldc)
{
    if (M <= 0 || N <= 0) return LU2MI_OK;
    if ((fmt != LU2MI_LOWP_FP16 && fmt != LU2MI_LOWP_BF16 && fmt != LU2MI_LOWP_FP16X2) || K <= 0 || K % 64) {
        lu2mi_set_error_msg("16-bit update: format %d, K=%d (need FP16/BF16/FP16x2 and K a multiple of 64)", fmt, K);
        return LU2MI_ERR_ARG;
    }
    const int tiles_m = (M + TB - 1) / TB, tiles_n = (N + TB - 1) / TB;
    const size_t np = (fmt == LU2MI_LOWP_FP16X2) ? 2 : 1;
    // operand copies padded to whole 128-row blocks, plus one more block so
    // that a wide tile's second half past the last block reads inside them
    LU2MI_TRY(lu2mi_dev_buf(&ctx->m_a16, &ctx->m_a16_bytes, np * (tiles_m + 1) * TB * K * 2));
    LU2MI_TRY(lu2mi_dev_buf(&ctx->m_b16, &ctx->m_b16_bytes, np * (tiles_n + 1) * TB * K * 2));
    LU2MI_TRY(lu2mi_dev_buf(&ctx->m_s16, &ctx->m_s16_bytes, sizeof(int) * (size_t)(tiles_m + tiles_n)));
    int *sa = (int *)ctx->m_s16, *sb = sa + tiles_m;
    if (fmt == LU2MI_LOWP_FP16)
        LU2MI_TRY(launch16<LU2MI_LOWP_FP16>(ctx, M, N, K, L, ldl, U, ldu, C, ldc, tiles_m, tiles_n, sa, sb));
    else if (fmt == LU2MI_LOWP_BF16)
        LU2MI_TRY(launch16<LU2MI_LOWP_BF16>(ctx, M, N, K, L, ldl, U, ldu, C, ldc, tiles_m, tiles_n, sa, sb));
    else
        LU2MI_TRY(launch16<LU2MI_LOWP_FP16X2>(ctx, M, N, K, L, ldl, U, ldu, C, ldc, tiles_m, tiles_n, sa, sb));
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

extern "C" int lu2mi_gemm16_dev(lu2mi_ctx *ctx, int fmt, int m, int n, int k, const float *dL, long ldl,
                                const float *dU, long ldu, float *dC, long ldc)
{
    if (!ctx || m < 0 || n < 0 || k < 0 || (m && n && k && (!dL || !dU || !dC)) || ldl < k || ldu < n || ldc < n) {
        lu2mi_set_error_msg("lu2mi_gemm16_dev: bad arguments");
        return LU2MI_ERR_ARG;
    }
    LU2MI_CHECK(hipSetDevice(ctx->device));
    LU2MI_TRY(lu2mi_update16(ctx, fmt, m, n, k, dL, ldl, dU, ldu, dC, ldc));
    LU2MI_CHECK(hipStreamSynchronize(ctx->stream));
    return LU2MI_OK;
}
