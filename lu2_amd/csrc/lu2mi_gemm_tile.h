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

// ---------------------------------------------------------------------------
// Trailing-update tile, FP64 only: 128x128 per workgroup, 4 waves of 64x64
// (4x4 blocks of v_mfma_f64_16x16x4f64), K staged 16 at a time through two
// LDS stages of 32 KB.  The operands go from global memory to LDS by direct
// 16-byte loads (global_load_lds_dwordx4): no registers hold them in flight
// and no LDS store instructions are issued.  The loads of stage k+1 are
// issued right after the barrier that opens stage k.
//
// K order.  In MFMA step s (0..3) of a stage, lane group q = lane/16 supplies
// k = 4q + s, so a lane's four k of an A row are 32 contiguous bytes (two
// 16-byte LDS reads per 16-row block instead of four 8-byte reads).
// Column pairs.  The wave's 64 columns are two groups of 32; MFMA block 2p+e
// (p = group, e = 0/1) takes the even (e=0) or odd (e=1) columns of group p,
// so one 16-byte read of a B row gives a lane the operands of both blocks and
// the epilogue moves C as 16-byte pairs.
// LDS layout.  A: 16-byte chunk c (k = 2c, 2c+1) of tile row r at slot
// 8r + (c ^ f(r)), f(r) = (r & 7) ^ ((r >> 3) & 1): any 8 (and any 16)
// consecutive rows read at the same chunk land in different bank groups.
// B: row k, column pair j at slot 64k + j (consecutive lanes read
// consecutive pairs).  Every element of C gets the same MFMA sequence in the
// interior tile (direct loads) and the edge tile (bounds-checked loads through
// registers, zero fill, LDS stores into the same slots).
namespace lu2mi_gemm2 {
constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int SLOTS = BM * BK / 2;                 // 16-byte slots per operand per stage (1024)

struct Smem {
    double2 A0[SLOTS], A1[SLOTS], B0[SLOTS], B1[SLOTS];   // two stages, 64 KB
};

__device__ __forceinline__ int a_slot(int r, int c) { return 8 * r + (c ^ ((r & 7) ^ ((r >> 3) & 1))); }

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void gbl_void;

template <bool EDGE, int EPI>
__device__ __forceinline__ void tile(const GemmDesc &g, const double *__restrict__ A, const double *__restrict__ B,
                                     double *__restrict__ C, int m0, int n0, int K, Smem &sm)
{
    const int M = g.M, N = g.N;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1, q = lane >> 4, j16 = lane & 15;
    const long lda = g.lda, ldb = g.ldb;

    // this thread's four A slots and four B slots of a stage: slot p =
    // 64 (4w + i) + lane (i = 0..3); A slot p holds row p/8, chunk (p%8)^f(row);
    // B slot p holds row p/64, pair p%64
    int a_row[4], a_col[4], b_row[4], b_col[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int p = 64 * (4 * w + i) + lane;
        const int r = p >> 3;
        a_row[i] = r;
        a_col[i] = 2 * ((p & 7) ^ ((r & 7) ^ ((r >> 3) & 1)));
        b_row[i] = p >> 6;
        b_col[i] = 2 * (p & 63);
    }

    d4_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int nj = 0; nj < 4; nj++) acc[mi][nj] = (d4_t){0.0, 0.0, 0.0, 0.0};

    // interior: direct global -> LDS loads of stage kt into the given buffers
    auto issue = [&](int kt, double2 *As, double2 *Bs) {
        const int k0 = kt * BK;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double *src = A + (long)(m0 + a_row[i]) * lda + k0 + a_col[i];
            __builtin_amdgcn_global_load_lds((gbl_void *)src, (lds_void *)(As + 64 * (4 * w + i)), 16, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double *src = B + (long)(k0 + b_row[i]) * ldb + n0 + b_col[i];
            __builtin_amdgcn_global_load_lds((gbl_void *)src, (lds_void *)(Bs + 64 * (4 * w + i)), 16, 0, 0);
        }
    };
    // edge: bounds-checked loads into registers (zero fill), stored later
    double2 ra[4], rb[4];
    auto fetch = [&](int kt) {
        const int k0 = kt * BK;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int gr = m0 + a_row[i], gk = k0 + a_col[i];
            const double *pa = A + (long)gr * lda + gk;
            ra[i].x = (gr < M && gk < K) ? pa[0] : 0.0;
            ra[i].y = (gr < M && gk + 1 < K) ? pa[1] : 0.0;
            const int bk = k0 + b_row[i], bn = n0 + b_col[i];
            const double *pb = B + (long)bk * ldb + bn;
            rb[i].x = (bk < K && bn < N) ? pb[0] : 0.0;
            rb[i].y = (bk < K && bn + 1 < N) ? pb[1] : 0.0;
        }
    };
    auto stash = [&](double2 *As, double2 *Bs) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            As[64 * (4 * w + i) + lane] = ra[i];
            Bs[64 * (4 * w + i) + lane] = rb[i];
        }
    };

    // one stage: the fragments for MFMA steps 0-1 and 2-3 are read first
    // (sixteen 16-byte LDS reads), then the 64 MFMAs of the stage.  The
    // scheduling barriers keep every MFMA of the stage ahead of the wait for
    // the next stage's loads (the compiler would otherwise hoist that wait).
    auto compute = [&](const double2 *As, const double2 *Bs, auto &&after_reads) {
        double2 af[2][4], bf[2][2][2];   // [half][mi] / [half][p][s&1]
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int mi = 0; mi < 4; mi++) af[h][mi] = As[a_slot(wm * 64 + mi * 16 + j16, 2 * q + h)];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int s1 = 0; s1 < 2; s1++) bf[h][p][s1] = Bs[64 * (4 * q + 2 * h + s1) + wn * 32 + 16 * p + j16];
            __builtin_amdgcn_sched_barrier(0);   // steps 0-1 need only the first eight reads
        }
        after_reads();   // the next stage's loads, issued behind this stage's LDS reads
#pragma unroll
        for (int h = 0; h < 2; h++) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int s1 = 0; s1 < 2; s1++)
#pragma unroll
                for (int mi = 0; mi < 4; mi++) {
                    const double a = s1 ? af[h][mi].y : af[h][mi].x;
#pragma unroll
                    for (int p = 0; p < 2; p++) {
                        acc[mi][2 * p] =
                            __builtin_amdgcn_mfma_f64_16x16x4f64(a, bf[h][p][s1].x, acc[mi][2 * p], 0, 0, 0);
                        acc[mi][2 * p + 1] =
                            __builtin_amdgcn_mfma_f64_16x16x4f64(a, bf[h][p][s1].y, acc[mi][2 * p + 1], 0, 0, 0);
                    }
                }
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    const int nk = (K + BK - 1) / BK;
    if (!EDGE) {
        // stages alternate between the two buffers; the loop is unrolled by
        // two so that each buffer is a distinct LDS object
        // (__syncthreads waits for this wave's outstanding loads first)
        if (nk > 0) issue(0, sm.A0, sm.B0);
        int kt = 0;
        for (; kt + 1 < nk; kt += 2) {
            __syncthreads();
            compute(sm.A0, sm.B0, [&] { issue(kt + 1, sm.A1, sm.B1); });
            __syncthreads();
            compute(sm.A1, sm.B1, [&] {
                if (kt + 2 < nk) issue(kt + 2, sm.A0, sm.B0);
            });
        }
        if (kt < nk) {
            __syncthreads();
            compute(sm.A0, sm.B0, [] {});
        }
    } else {
        // edge tiles (few): the next stage is loaded and stored to LDS before
        // the current one is computed, so no registers carry it across
        if (nk > 0) {
            fetch(0);
            stash(sm.A0, sm.B0);
        }
        int kt = 0;
        for (; kt + 1 < nk; kt += 2) {
            __syncthreads();
            fetch(kt + 1);
            stash(sm.A1, sm.B1);
            compute(sm.A0, sm.B0, [] {});
            __syncthreads();
            if (kt + 2 < nk) {
                fetch(kt + 2);
                stash(sm.A0, sm.B0);
            }
            compute(sm.A1, sm.B1, [] {});
        }
        if (kt < nk) {
            __syncthreads();
            compute(sm.A0, sm.B0, [] {});
        }
    }

    // epilogue: acc[mi][2p+e][r] is C[m0 + 64wm + 16mi + q + 4r][n0 + 64wn + 32p + 2j16 + e]
    const double alpha = g.alpha, beta = g.beta;
    constexpr int epi = EPI == lu2mi_gemm::EPI_SUB ? 1 : EPI == lu2mi_gemm::EPI_SET ? 2 : 0;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = m0 + wm * 64 + mi * 16 + q + 4 * r;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int col = n0 + wn * 64 + 32 * p + 2 * j16;
                double v0 = acc[mi][2 * p][r], v1 = acc[mi][2 * p + 1][r];
                double *cp = C + (long)row * g.ldc + col;
                if (!EDGE) {
                    double2 c2 = *(double2 *)cp;
                    if (epi == 1) { c2.x = c2.x - v0; c2.y = c2.y - v1; }
                    else if (epi == 2) { c2.x = v0; c2.y = v1; }
                    else if (beta == 0.0) { c2.x = alpha * v0; c2.y = alpha * v1; }
                    else { c2.x = alpha * v0 + beta * c2.x; c2.y = alpha * v1 + beta * c2.y; }
                    *(double2 *)cp = c2;
                } else if (row < M) {
                    if (col < N) {
                        double c0 = cp[0];
                        cp[0] = epi == 1 ? c0 - v0 : epi == 2 ? v0 : beta == 0.0 ? alpha * v0 : alpha * v0 + beta * c0;
                    }
                    if (col + 1 < N) {
                        double c1 = cp[1];
                        cp[1] = epi == 1 ? c1 - v1 : epi == 2 ? v1 : beta == 0.0 ? alpha * v1 : alpha * v1 + beta * c1;
                    }
                }
            }
        }
}
}  // namespace lu2mi_gemm2
