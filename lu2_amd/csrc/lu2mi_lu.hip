// lu2mi_lu.hip -- GPU-resident blocked recursive LU with partial pivoting and
// the solve that follows it, in lu2's factor form.
//
// Algorithm (lu2.c:1122-1144 lumcolumn, lu2.c:1088-1120 recursive panel,
// lu2.c:524-581 process_right_part, lu2.c:485-521 solve_triangle,
// lu2.c:110-116 backward_sub) re-planned for one MI355X:
//
//  per nb-block i:
//    panel(i, nb)                       recursive halving down to a W-column leaf
//      leaf   : one persistent kernel factors W columns of rows [c, n) held in
//               registers; one grid barrier per column exchanges the
//               per-workgroup pivot candidates (wavefront argmax, lowest row
//               on ties) together with the candidate rows, so the winning row
//               arrives with the decision.  It also forms the leaf's W x W
//               block of (D^-1 L11)^-1.
//      rowperm: the leaf's swaps applied to the other panel columns, with
//               the 1/d scaling of the L part left of the diagonal
//               (lu2.c:1112-1118, lu2tlib.c:1119-1126).
//      node   : right half rows scaled by 1/d, TRSM by the explicit inverse
//               (lu2.c:467-512) as an MFMA GEMM, trailing MFMA GEMM
//               (lu2.c:580), and the off-diagonal block of the inverse.
//    rowperm of the block's nb swaps over columns [i+nb, n] (b included),
//      1/d scaling, block rows redirected to the staging buffer (awork)
//    U12 = Linv * staging                 (lu2.c:511-512)
//    A22 -= L21 * U12                     (lu2.c:580, 97.7% of the flops)
//  back substitution on the unit-upper U (lu2.c:110-116), blocked.
//
//  Lookahead schedule (default): the block-level steps run first on the next
//  panel's nb columns only, panel(i+nb) then starts on a second stream, and
//  the block's remaining columns follow on the first; same arithmetic per
//  element as the serial order.  Optionally, up to 256 columns of a panel
//  run as one persistent launch (subpanel_kernel) with grid barriers between
//  the steps above.
//
// Layout in HBM: lu2's row-major augmented matrix, lda >= n+1 (RDIM = n+16
// in lu2.c:28), b in column n.  Row-major rows are contiguous, so row swaps
// and scalings are coalesced along rows.
#include "lu2mi_internal.h"
#include "lu2mi_dev.h"
#include "lu2mi_gemm_tile.h"
#include <limits.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <vector>

// drand48 on the device: jump-ahead generator in lu2mi_dev.h.
using namespace lu2mi_rng;

__global__ __launch_bounds__(256) void randomsetmat_kernel(double *A, long lda, int n,
                                                           unsigned long long x0)
{
    const long r = blockIdx.x;
    const int t = threadIdx.x;
    // element (r, col) is draw number r*(n+1)+col; its value is the state
    // after draw+1 steps
    const unsigned long long first = (unsigned long long)r * (unsigned long long)(n + 1) + t + 1;
    const Lcg jump = lcg_pow(first);
    unsigned long long x = (jump.a * x0 + jump.c) & M48;
    const Lcg step = lcg_pow(256);
    double *row = A + r * lda;
    for (int col = t; col <= n; col += 256) {
        row[col] = (double)x * (1.0 / 281474976710656.0) - 0.5;
        x = (step.a * x + step.c) & M48;
    }
    for (long col = n + 1 + t; col < lda; col += 256) row[col] = 0.0;
}

// a[i][i] += sum_j |a[i][j]| over the n columns of A (diagonal dominance for
// the 16-bit solves); one workgroup per row, fixed reduction order
__global__ __launch_bounds__(256) void diag_dominant_kernel(double *A, long lda, int n)
{
    __shared__ double red[4];
    const long r = blockIdx.x;
    const int t = threadIdx.x;
    const double *row = A + r * lda;
    double s = 0.0;
    for (int c = t; c < n; c += 256) s += fabs(row[c]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((t & 63) == 0) red[t >> 6] = s;
    __syncthreads();
    if (t == 0) A[r * lda + r] += (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Leaf factorization: persistent kernel over rows [c, n), W columns; all
// workgroups are co-resident and meet at a grid barrier once per column.
namespace {
constexpr int SLOT_WORDS = 40;  // 3 header words + up to 32 row values; the last word is the sequence number
constexpr int NTL = 256;        // threads per workgroup in the leaf kernel
constexpr int LEAF_GMAX = 128;  // most workgroups a leaf launch may use

__device__ __forceinline__ void st_relaxed(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed(unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long dbits(double v) { return __double_as_longlong(v); }
__device__ __forceinline__ double bitsd(unsigned long long v) { return __longlong_as_double(v); }
// device-scope (coherent across XCDs) scalar loads and stores of a vector element
__device__ __forceinline__ double ld_dev(const double *p)
{
    return bitsd(__hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ float ld_dev(const float *p)
{
    return __uint_as_float(__hip_atomic_load((unsigned *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// device-scope loads through a buffer resource: same coherence as ld_dev but
// without a wait after every load, so independent loads overlap
__device__ __forceinline__ double ld_buf(__amdgpu_buffer_rsrc_t r, int byte_off, double)
{
    const unsigned int lo = __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 16);
    const unsigned int hi = __builtin_amdgcn_raw_buffer_load_b32(r, byte_off + 4, 0, 16);
    return bitsd(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float ld_buf(__amdgpu_buffer_rsrc_t r, int byte_off, float)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 16));
}
__device__ __forceinline__ void st_dev(double *p, double v)
{
    __hip_atomic_store((unsigned long long *)p, dbits(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(float *p, float v)
{
    __hip_atomic_store((unsigned *)p, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publication of a sequence word (leaf slots, back-substitution frontier)
// after the data it announces.  Every exchanged word is written and read with
// device-coherent (agent-scope, sc1) accesses, so no cache maintenance is
// needed, only completion order: the compiler barriers keep the data stores
// ahead of the sequence store in program order and s_waitcnt vmcnt(0) makes
// them complete before it is issued.  strict: an agent-scope release store
// instead (on gfx950 it adds buffer_wbl2, an L2 writeback, per publication;
// the matching acquire on the reader adds buffer_inv, an L2 invalidate).
__device__ __forceinline__ void publish_seq(unsigned long long *p, unsigned long long v, int strict)
{
    if (strict) {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
// reader side, after the spin on a sequence word has seen the new value
__device__ __forceinline__ void acquire_seq(int strict)
{
    if (strict) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// rank-1 update of one element: x - l * (u * s), with u*s rounded first (the
// stored U value) and one fused multiply-subtract, written out so that every
// copy of a row computes bit-identical values
__device__ __forceinline__ double pending_update(double x, double l, double u, double s)
{
    return fma(-l, u * s, x);
}
__device__ __forceinline__ float pending_update(float x, float l, float u, float s)
{
    return fmaf(-l, u * s, x);
}

__device__ __forceinline__ bool better(double v, int r, double bv, int br)
{
    return (v > bv) || (v == bv && r < br);
}
}  // namespace

// maximum that keeps a NaN (fmax drops it): a NaN solution must not pass as
// a small residual
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

template <typename T>
struct LeafArgs {
    T *A;
    long lda;
    int n, c, i, compat;
    int G;                        // number of workgroups (grid size)
    int *piv;                     // global pivots, piv[j] for j in [c, c+W)
    int *orig;                    // block-level origin row ids
    T *invd;                      // 1/d per pivot row
    int *pair_dst, *pair_src;     // 2W pairs for the panel row permutation
    T *linv;                      // nb x nb inverse of D^-1 L11 (this leaf's block)
    int ldl;
    unsigned long long *slots;    // 2 * (G+1) * SLOT_WORDS, G <= LEAF_GMAX
    unsigned long long seq_base;  // step sequence numbers seq_base+1 .. seq_base+W
    int *err;
    int strict;                   // release/acquire publication (publish_seq)
};

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Per column step j (one grid-wide exchange):
//   gather   every workgroup waits for the step's sequence number in each
//            slot and copies all slots (candidate value, row, origins and
//            the W values of the candidate row, plus the current row j) into
//            LDS with device-scope buffer loads;
//   reduce   each wavefront reduces the G candidates (largest |x|, lowest
//            row on ties) -- identical result everywhere, no extra barrier;
//   swap     the owners of rows j and p exchange rows, row j scaled by 1/d;
//   column j+1 updated first, its candidate reduced in the workgroup and
//            published (lookahead), then the other columns updated while
//            the exchange is in flight.
template <typename T, int W>
struct LeafSmem {
    T stage[NTL * (W + 1)];
    unsigned long long xch[2][(LEAF_GMAX + 1) * (W + 3)];
    double red_v[NTL / 64];
    int red_r[NTL / 64];
    int pivs[W];
};

// the leaf as workgroup bid of a.G (a stand-alone launch, or a phase of the
// persistent sub-panel kernel)
template <typename T, int W, int RPT>
__device__ __forceinline__ void leaf_body(const LeafArgs<T> &a, const int bid, LeafSmem<T, W> &sm)
{
    constexpr int SW = W + 3;
    T *const stage = sm.stage;
    auto &xch = sm.xch;
    double *const red_v = sm.red_v;
    int *const red_r = sm.red_r;
    int *const pivs = sm.pivs;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int G = a.G;
    const long TT = (long)G * NTL;
    const int n = a.n, c = a.c;
    T *__restrict__ A = a.A;
    unsigned long long *const slots = a.slots;
    const __amdgpu_buffer_rsrc_t srsrc =
        __builtin_amdgcn_make_buffer_rsrc(slots, 0, (int)(2 * (G + 1) * SLOT_WORDS * 8), 0x00020000);

    T row[RPT][W];
    int lorig[RPT], borig[RPT], rq[RPT];

    // coalesced load of the rows through LDS, one 256-row layer at a time
#pragma clang loop unroll(full)
    for (int q = 0; q < RPT; q++) {
        const long base = (long)c + (long)bid * NTL + (long)q * TT;
        rq[q] = (int)min((long)INT_MAX, base + t);
        for (int e = t; e < NTL * W; e += NTL) {
            const int rr = e / W, kk = e % W;
            const long gr = base + rr;
            stage[rr * (W + 1) + kk] = (gr < n) ? A[gr * a.lda + c + kk] : T(0);
        }
        __syncthreads();
#pragma clang loop unroll(full)
        for (int k = 0; k < W; k++) row[q][k] = stage[t * (W + 1) + k];
        __syncthreads();
        lorig[q] = rq[q];
        borig[q] = (rq[q] < n) ? ((c == a.i) ? rq[q] : a.orig[rq[q]]) : -1;
    }

    // local candidate of column jc, workgroup reduction, publication of the
    // slot (and of the current row by its owner in workgroup 0)
    // pprev/sprev: the previous step's pivot slot and 1/d when that step's
    // update of columns > jc is still pending (lookahead); the published
    // values then carry it, computed exactly as the bulk update computes it
    auto publish = [&](const int jc, const unsigned long long *pprev, const T sprev)
                       __attribute__((always_inline)) {
        const int j = c + jc;
        const int par = jc & 1;
        const unsigned long long seq = a.seq_base + (unsigned long long)jc + 1;
        const bool odd_rule = a.compat && (jc & 1);
        double bv = -1.0;
        int br = INT_MAX;
#pragma clang loop unroll(full)
        for (int q = 0; q < RPT; q++) {
            const int r = rq[q];
            if (r >= j && r < n) {
                const T xa = row[q][jc];
                const T xb = row[q][jc > 0 ? jc - 1 : 0];
                const double v = fabs((double)(odd_rule ? xb : xa));
                if (v > bv) { bv = v; br = r; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int orr = __shfl_xor(br, off);
            if (better(ov, orr, bv, br)) { bv = ov; br = orr; }
        }
        if (lane == 0) { red_v[wave] = bv; red_r[wave] = br; }
        __syncthreads();
        double wv = red_v[0];
        int wr = red_r[0];
#pragma unroll
        for (int w2 = 1; w2 < NTL / 64; w2++)
            if (better(red_v[w2], red_r[w2], wv, wr)) { wv = red_v[w2]; wr = red_r[w2]; }
        unsigned long long *sl = slots + ((long)par * (G + 1) + bid) * SLOT_WORDS;
        auto value = [&](int q, int k) __attribute__((always_inline)) -> T {
            if (pprev && k > jc) return pending_update(row[q][k], row[q][jc > 0 ? jc - 1 : 0], (T)bitsd(pprev[3 + k]), sprev);
            return row[q][k];
        };
#pragma clang loop unroll(full)
        for (int q = 0; q < RPT; q++)
            if (rq[q] == wr && wr < n) {
                st_relaxed(sl + 0, dbits(wv));
                st_relaxed(sl + 1, (unsigned long long)(unsigned)wr);
                st_relaxed(sl + 2, (unsigned long long)(unsigned)lorig[q] |
                                       ((unsigned long long)(unsigned)borig[q] << 32));
#pragma clang loop unroll(full)
                for (int k = 0; k < W; k++) st_relaxed(sl + 3 + k, dbits((double)value(q, k)));
                publish_seq(sl + SLOT_WORDS - 1, seq, a.strict);
            }
        if (wr >= n && t == 0) {
            st_relaxed(sl + 0, dbits(-1.0));
            st_relaxed(sl + 1, (unsigned long long)(unsigned)INT_MAX);
            publish_seq(sl + SLOT_WORDS - 1, seq, a.strict);
        }
        if (bid == 0 && t == jc) {
            unsigned long long *cs = slots + ((long)par * (G + 1) + G) * SLOT_WORDS;
            st_relaxed(cs + 2, (unsigned long long)(unsigned)lorig[0] |
                                   ((unsigned long long)(unsigned)borig[0] << 32));
#pragma clang loop unroll(full)
            for (int k = 0; k < W; k++) st_relaxed(cs + 3 + k, dbits((double)value(0, k)));
            publish_seq(cs + SLOT_WORDS - 1, seq, a.strict);
        }
    };

    publish(0, nullptr, T(0));
#pragma clang loop unroll(full)
    for (int jl = 0; jl < W; jl++) {
        const int j = c + jl;
        const int par = jl & 1;
        const unsigned long long seq = a.seq_base + (unsigned long long)jl + 1;
        unsigned long long *const xb = xch[par];
        // ---- gather: thread g waits for slot g, then copies it to LDS
        if (t <= G) {
            unsigned long long *sl = slots + ((long)par * (G + 1) + t) * SLOT_WORDS;
            unsigned long spins = 0;
            while (ld_relaxed(sl + SLOT_WORDS - 1) != seq) {
                __builtin_amdgcn_s_sleep(1);
                if (((++spins) & 1023) == 0) {
                    if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                    if (spins > (1UL << 25)) {
                        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            acquire_seq(a.strict);
            const int boff = (int)(((long)par * (G + 1) + t) * SLOT_WORDS * 8);
#pragma clang loop unroll(full)
            for (int k = 0; k < SW; k++) {
                const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(srsrc, boff + 8 * k, 0, 16);
                xb[t * SW + k] = ((unsigned long long)v.y << 32) | v.x;
            }
        }
        __syncthreads();
        // ---- every wavefront reduces the G candidates (same result everywhere)
        double bv = -1.0;
        int br = INT_MAX, bg = -1;
        for (int g2 = lane; g2 < G; g2 += 64) {
            const double v = bitsd(xb[g2 * SW + 0]);
            const int r = (int)(unsigned)xb[g2 * SW + 1];
            if (better(v, r, bv, br)) { bv = v; br = r; bg = g2; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int orr = __shfl_xor(br, off);
            const int og = __shfl_xor(bg, off);
            if (better(ov, orr, bv, br)) { bv = ov; br = orr; bg = og; }
        }
        // no comparable candidate (every |x| is NaN): pivot on the current
        // row, as BLAS idamax returns the first index for NaN input
        const bool none = (br < j || br >= n || bg < 0);
        const int p = none ? j : br;
        const unsigned long long *pslot = xb + (none ? G : bg) * SW;
        const unsigned long long *cslot = xb + G * SW;
        const int p_lorig = (int)(unsigned)(pslot[2] & 0xffffffffULL), p_borig = (int)(unsigned)(pslot[2] >> 32);
        const int c_lorig = (int)(unsigned)(cslot[2] & 0xffffffffULL), c_borig = (int)(unsigned)(cslot[2] >> 32);
#define PIVROW(k) ((T)bitsd(pslot[3 + (k)]))
#define CURROW(k) ((T)bitsd(cslot[3 + (k)]))
        const T d = PIVROW(jl);
        const T s = T(1) / d;
        // ---- swap rows j and p; row j becomes the pivot row scaled by 1/d
#pragma clang loop unroll(full)
        for (int q = 0; q < RPT; q++) {
            const int r = rq[q];
            if (r == j) {
#pragma clang loop unroll(full)
                for (int k = 0; k < W; k++) row[q][k] = (k == jl) ? d : PIVROW(k) * s;
                lorig[q] = p_lorig;
                borig[q] = p_borig;
            } else if (r == p) {
#pragma clang loop unroll(full)
                for (int k = 0; k < W; k++) row[q][k] = CURROW(k);
                lorig[q] = c_lorig;
                borig[q] = c_borig;
            }
        }
        // ---- column j+1 first, then publish its candidate (lookahead)
        if (jl + 1 < W) {
#pragma clang loop unroll(full)
            for (int q = 0; q < RPT; q++) {
                const int r = rq[q];
                if (r > j && r < n) {
                    T &x1 = row[q][jl + 1 < W ? jl + 1 : 0];
                    x1 = pending_update(x1, row[q][jl], PIVROW(jl + 1 < W ? jl + 1 : 0), s);
                }
            }
            publish(jl + 1 < W ? jl + 1 : 0, pslot, s);
        }
        // ---- the remaining columns while the exchange is in flight
#pragma clang loop unroll(full)
        for (int q = 0; q < RPT; q++) {
            const int r = rq[q];
            if (r > j && r < n) {
                const T l = row[q][jl];
#pragma clang loop unroll(full)
                for (int k = jl + 2; k < W; k++) row[q][k] = pending_update(row[q][k], l, PIVROW(k), s);
            }
        }
        if (t == 0) {
            pivs[jl] = p;
            if (bid == 0) {
                a.piv[j] = p;
                a.invd[j] = s;
            }
        }
#undef PIVROW
#undef CURROW
    }
    __syncthreads();

    // ---- write back rows (coalesced through LDS), origins and row pairs
#pragma clang loop unroll(full)
    for (int q = 0; q < RPT; q++) {
        const long base = (long)c + (long)bid * NTL + (long)q * TT;
#pragma clang loop unroll(full)
        for (int k = 0; k < W; k++) stage[t * (W + 1) + k] = row[q][k];
        __syncthreads();
        for (int e = t; e < NTL * W; e += NTL) {
            const int rr = e / W, kk = e % W;
            const long gr = base + rr;
            if (gr < n) A[gr * a.lda + c + kk] = stage[rr * (W + 1) + kk];
        }
        __syncthreads();
        const int r = rq[q];
        if (r < n) {
            a.orig[r] = borig[q];
#pragma clang loop unroll(full)
            for (int k = 0; k < W; k++)
                if (pivs[k] == r) {
                    a.pair_dst[W + k] = r;
                    a.pair_src[W + k] = lorig[q];
                }
        }
    }
    if (bid == 0) {
        if (t < W) {
            a.pair_dst[t] = c + t;
            a.pair_src[t] = lorig[0];
            // leaf block of the inverse of the unit-lower D^-1 L11
#pragma clang loop unroll(full)
            for (int k = 0; k < W; k++) stage[t * (W + 1) + k] = row[0][k];
        }
        __syncthreads();
        if (t < W) {
            T x[W];
#pragma clang loop unroll(full)
            for (int r = 0; r < W; r++) {
                if (r < t) x[r] = T(0);
                else if (r == t) x[r] = T(1);
                else {
                    T acc = T(0);
#pragma clang loop unroll(full)
                    for (int k = 0; k < r; k++)
                        if (k >= t) acc += stage[r * (W + 1) + k] * x[k];
                    x[r] = -acc;
                }
            }
            const long off = (long)(c - a.i);
#pragma clang loop unroll(full)
            for (int r = 0; r < W; r++) a.linv[(off + r) * a.ldl + off + t] = x[r];
        }
    }
}

template <typename T, int W, int RPT>
__global__ __launch_bounds__(NTL) void leaf_kernel(LeafArgs<T> a)
{
    __shared__ LeafSmem<T, W> sm;
    leaf_body<T, W, RPT>(a, blockIdx.x, sm);
}

// ---------------------------------------------------------------------------
// Row permutation of a column range: all sources of a column strip are read
// into LDS before any destination is written, so in-place permutations are
// safe.  Destinations in [d0, d1) are scaled by invd[dst] on columns < sce;
// destinations in [w0, w1) are written to the staging buffer instead.
template <typename T>
struct PermArgs {
    T *A;
    long lda;
    const int *dst, *src;
    int np;
    int c0, c1, k0, k1;   // columns [c0, c1) except [k0, k1)
    int sce;              // scale columns < sce
    int d0, d1;
    const T *invd;
    T *W;
    long ldw;
    int w0, w1;
    int sw;               // strip width
};

template <typename T>
__device__ __forceinline__ void rowperm_strip(const PermArgs<T> &a, const int strip, T *buf)
{
    const int cs = a.c0 + strip * a.sw;
    const int total = a.np * a.sw;
    for (int e = threadIdx.x; e < total; e += 256) {
        const int pr = e / a.sw, cc = e % a.sw;
        const int col = cs + cc;
        if (col < a.c1 && (col < a.k0 || col >= a.k1)) buf[e] = a.A[(long)a.src[pr] * a.lda + col];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < total; e += 256) {
        const int pr = e / a.sw, cc = e % a.sw;
        const int col = cs + cc;
        if (col < a.c1 && (col < a.k0 || col >= a.k1)) {
            const int d = a.dst[pr];
            T v = buf[e];
            if (d >= a.d0 && d < a.d1 && col < a.sce) v *= a.invd[d];
            if (d >= a.w0 && d < a.w1) a.W[(long)(d - a.w0) * a.ldw + (col - a.c0)] = v;
            else a.A[(long)d * a.lda + col] = v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void rowperm_kernel(PermArgs<T> a)
{
    extern __shared__ unsigned char smem_raw[];
    rowperm_strip<T>(a, blockIdx.x, (T *)smem_raw);
}

// pairs of the block's row permutation: (i+k, orig[i+k]) and (piv[i+k], orig[piv[i+k]])
// The block-level permutation (lu2.c:524-556) without LDS staging, in two
// launches over (pair, 1024-column chunk).  Destinations inside the block go
// to the staging buffer W (scaled by 1/d), so pass 0 reads A and writes only
// W.  A row below the block only ever receives content that started in the
// block (content that enters the block from below stays there), so pass 1
// reads block rows of A, which no pass writes, and writes rows below, which
// pass 1 never reads.  Every read of a row below happens in pass 0.
template <typename T>
__global__ __launch_bounds__(256) void staged_perm_kernel(PermArgs<T> a, int pass)
{
    const int pr = blockIdx.y;
    const int d = a.dst[pr];
    const bool to_w = (d >= a.w0 && d < a.w1);
    if (to_w != (pass == 0)) return;
    const T *srow = a.A + (long)a.src[pr] * a.lda;
    const bool scale = (d >= a.d0 && d < a.d1);
    const T sc = scale ? a.invd[d] : T(1);
    T *drow = to_w ? a.W + (long)(d - a.w0) * a.ldw - a.c0 : a.A + (long)d * a.lda;
    const int col0 = a.c0 + blockIdx.x * 1024;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int col = col0 + q * 256 + threadIdx.x;
        if (col < a.c1) {
            T v = srow[col];
            if (scale && col < a.sce) v *= sc;
            drow[col] = v;
        }
    }
}

__global__ void block_pairs_kernel(const int *piv, const int *orig, int i, int nb, int *dst, int *src)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nb) {
        dst[k] = i + k;
        src[k] = orig[i + k];
        const int p = piv[i + k];
        dst[nb + k] = p;
        src[nb + k] = orig[p];
    }
}

// W[r][col] = A[r0+r][c0+col] * invd[r0+r]  (rows scaled by 1/d before the TRSM)
template <typename T>
__global__ void scale_copy_kernel(const T *A, long lda, int r0, int c0, int h, int nc, const T *invd,
                                  T *W)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long)h * nc) {
        const int r = (int)(e / nc), col = (int)(e % nc);
        W[e] = A[(long)(r0 + r) * lda + c0 + col] * invd[r0 + r];
    }
}

// ---------------------------------------------------------------------------
// Persistent sub-panel: one launch factors w columns [c, c+w) of the block's
// panel (w a power of two <= SUB_WMAX, 16-column leaves): every leaf, its row
// swaps over the panel's other columns, and every recursion node inside the
// sub-panel (1/d scaling and staging, explicit-inverse TRSM, update of the
// rows below, off-diagonal block of the inverse), in the order of the
// recursion of lu2.c:1187-1235, separated by grid barriers instead of kernel
// boundaries.  The steps run the same device code as the separate launches
// (leaf body, row-swap strip, GEMM tile), so the factors are bit-identical;
// the difference is one dispatch instead of about seven per leaf, which is
// what a panel factored beside the trailing update waits on.
constexpr int SUB_W = 16;       // leaf width inside the sub-panel kernel
constexpr int SUB_WMAX = 256;   // widest sub-panel
constexpr int SUB_SWMAX = 128;  // widest row-swap strip (LDS: 2*SUB_W pairs x strip)
constexpr size_t SLOTS_BYTES = sizeof(unsigned long long) * 2 * (size_t)(1024 + 1) * SLOT_WORDS;

template <typename T>
struct SubArgs {
    LeafArgs<T> leaf;            // first leaf: c = sub-panel start, seq_base of its first column
    int nb, w, sw;               // block size, sub-panel width, row-swap strip width
    T *wbuf, *tmp;               // h x h staging of the scaled rows; h x h product scratch
    unsigned long long *gbar;    // grid barrier arrivals (monotonic within a factorization)
    unsigned long long gbar_base;
};

// Barrier over the G co-resident workgroups: arrivals counted on one word,
// agent-scope fences on both sides so that every step's global writes (on
// any XCD) are visible to the next step.  Bounded: a timeout raises *err and
// lets every wave run to the end of the kernel.
__device__ __forceinline__ void grid_sync(unsigned long long *ctr, unsigned long long target, int *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(ctr, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (((++spins) & 1023) == 0) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (spins > (1UL << 25)) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// C = alpha A B + beta C over the G workgroups in 64x64 tiles, each with the
// K range of the stand-alone GEMM's 128x128 tile that contains it
template <typename T, int EPI>
__device__ __forceinline__ void gemm_steps(int M, int N, int K, double alpha, const T *A, long lda, const T *B,
                                           long ldb, double beta, T *C, long ldc, int a_lower, int bid, int G,
                                           lu2mi_gemm::Tile64Smem<T> &ts)
{
    using namespace lu2mi_gemm;
    GemmDesc g;
    g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc; g.a_lower = a_lower;
    const int tiles_m = (M + BT64 - 1) / BT64, tiles_n = (N + BT64 - 1) / BT64;
    const int total = (M > 0 && N > 0) ? tiles_m * tiles_n : 0;
    for (int tile = bid; tile < total; tile += G) {
        const int m0 = (tile / tiles_n) * BT64, n0 = (tile % tiles_n) * BT64;
        const int Kt = a_lower ? min(K, (m0 & ~(BM - 1)) + BM) : K;
        const bool interior = (m0 + BT64 <= M) && (n0 + BT64 <= N) && (Kt % Mfma<T>::BK == 0);
        if (interior) gemm_tile64<T, false, EPI>(g, A, B, C, m0, n0, Kt, ts);
        else gemm_tile64<T, true, EPI>(g, A, B, C, m0, n0, Kt, ts);
        __syncthreads();   // the next tile restages the same LDS
    }
}

template <typename T, int RPT>
__global__ __launch_bounds__(NTL, 2) void subpanel_kernel(SubArgs<T> s)
{
    using namespace lu2mi_gemm;
    constexpr int W = SUB_W;
    __shared__ union {
        LeafSmem<T, W> leaf;
        Tile64Smem<T> tile;
        T perm[2 * W * SUB_SWMAX];
    } sm;
    const int bid = blockIdx.x, G = s.leaf.G;
    const int i = s.leaf.i, n = s.leaf.n, c0 = s.leaf.c, nb = s.nb, L = s.w / W;
    const long lda = s.leaf.lda;
    const int ldl = s.leaf.ldl;
    T *const A = s.leaf.A;
    T *const linv = s.leaf.linv;
    unsigned long long target = s.gbar_base;
    auto sync = [&]() __attribute__((always_inline)) {
        target += (unsigned long long)G;
        grid_sync(s.gbar, target, s.leaf.err);
    };
    // node (cn, 2h) after its left half: rows [cn, cn+h) of the right half
    // scaled by 1/d and staged, X = inv(D^-1 L11) * staged rows, then the
    // rows below (host path: Factor::panel)
    auto node_mid = [&](const int cn, const int h) __attribute__((always_inline)) {
        const long off = cn - i;
        for (long e = (long)bid * NTL + threadIdx.x; e < (long)h * h; e += (long)G * NTL) {
            const int r = (int)(e / h), col = (int)(e % h);
            s.wbuf[e] = A[(long)(cn + r) * lda + cn + h + col] * s.leaf.invd[cn + r];
        }
        sync();
        gemm_steps<T, EPI_SET>(h, h, h, 1.0, linv + off * ldl + off, ldl, s.wbuf, h, 0.0,
                               A + (long)cn * lda + cn + h, lda, 1, bid, G, sm.tile);
        sync();
        gemm_steps<T, EPI_SUB>(n - cn - h, h, h, -1.0, A + (long)(cn + h) * lda + cn, lda,
                               A + (long)cn * lda + cn + h, lda, 1.0, A + (long)(cn + h) * lda + cn + h, lda, 0,
                               bid, G, sm.tile);
        sync();
    };
    // node (cn, 2h) after its right half: X21 = -inv22 * L21 * inv11
    auto node_end = [&](const int cn, const int h) __attribute__((always_inline)) {
        const long off = cn - i;
        gemm_steps<T, EPI_SET>(h, h, h, 1.0, A + (long)(cn + h) * lda + cn, lda, linv + off * ldl + off, ldl, 0.0,
                               s.tmp, h, 0, bid, G, sm.tile);
        sync();
        gemm_steps<T, EPI_GENERAL>(h, h, h, -1.0, linv + (off + h) * ldl + off + h, ldl, s.tmp, h, 0.0,
                                   linv + (off + h) * ldl + off, ldl, 1, bid, G, sm.tile);
        sync();
    };
    // post-order walk: before leaf k the nodes whose right half ended at
    // leaf k-1 (deepest first) get their inverse block, then the node whose
    // left half ended there gets its right half updated
    for (int k = 0; k <= L; k++) {
        if (k > 0) {
            const int tz = __builtin_ctz(k);
            for (int lv = 1; lv <= tz; lv++) node_end(c0 + (k - (1 << lv)) * W, W << (lv - 1));
            if (k == L) break;
            node_mid(c0 + (k - (1 << tz)) * W, W << tz);
        }
        LeafArgs<T> la = s.leaf;
        la.c = c0 + k * W;
        la.seq_base = s.leaf.seq_base + (unsigned long long)k * W;
        leaf_body<T, W, RPT>(la, bid, sm.leaf);
        sync();
        // the leaf's swaps on the panel's other columns; its diagonal rows'
        // L part (columns left of the leaf) scaled by 1/d
        PermArgs<T> p;
        p.A = A; p.lda = lda; p.dst = la.pair_dst; p.src = la.pair_src; p.np = 2 * W;
        p.c0 = i; p.c1 = i + nb; p.k0 = la.c; p.k1 = la.c + W; p.sce = la.c; p.d0 = la.c; p.d1 = la.c + W;
        p.invd = la.invd; p.W = nullptr; p.ldw = 0; p.w0 = 0; p.w1 = 0; p.sw = s.sw;
        const int strips = (nb + s.sw - 1) / s.sw;
        for (int st = bid; st < strips; st += G) {
            rowperm_strip<T>(p, st, sm.perm);
            __syncthreads();
        }
        sync();
    }
}

// grid barriers one sub-panel launch of L leaves passes: two per leaf, three
// per node before its right half, two per node after it
__host__ __device__ constexpr int sub_barriers(int L) { return 2 * L + 5 * (L - 1); }

// ---------------------------------------------------------------------------
// Recursion nodes of half-width h <= 64, one workgroup (one 64x64 tile) per
// step instead of two launches each; same arithmetic as the separate steps.
// node_trsm64: X = inv(D^-1 L11) * (rows [c, c+h) of the right half scaled
// by 1/d), the scaling done as the rows are loaded (the staging copy's
// values); the tile reads all of its rows before it writes them back.
template <typename T>
__global__ __launch_bounds__(256) void node_trsm64_kernel(T *A, long lda, int c, int h, const T *linv, int ldl,
                                                          long off, const T *invd)
{
    using namespace lu2mi_gemm;
    __shared__ Tile64Smem<T> sm;
    GemmDesc g;
    g.M = h; g.N = h; g.K = h; g.alpha = 1.0; g.beta = 0.0;
    g.lda = ldl; g.ldb = lda; g.ldc = lda; g.a_lower = 1;
    T *X = A + (long)c * lda + c + h;
    gemm_tile64<T, true, EPI_SET>(g, linv + off * ldl + off, X, X, 0, 0, h, sm, invd + c);
}

// node_inv64: the node's off-diagonal inverse block, X21 = -inv22 * (L21 * inv11),
// both products in one workgroup (tmp passes between them through memory)
template <typename T>
__global__ __launch_bounds__(256) void node_inv64_kernel(const T *A, long lda, int c, int h, T *linv, int ldl,
                                                         long off, T *tmp)
{
    using namespace lu2mi_gemm;
    __shared__ Tile64Smem<T> sm;
    GemmDesc g;
    g.M = h; g.N = h; g.K = h; g.alpha = 1.0; g.beta = 0.0;
    g.lda = lda; g.ldb = ldl; g.ldc = h; g.a_lower = 0;
    gemm_tile64<T, true, EPI_SET>(g, A + (long)(c + h) * lda + c, linv + off * ldl + off, tmp, 0, 0, h, sm);
    __threadfence();
    __syncthreads();
    g.alpha = -1.0; g.lda = ldl; g.ldb = h; g.ldc = ldl; g.a_lower = 1;
    gemm_tile64<T, true, EPI_GENERAL>(g, linv + (off + h) * ldl + off + h, tmp, linv + (off + h) * ldl + off, 0, 0,
                                      h, sm);
}

// column-major panel (m columns of ld rows, rows [r0, r0+R)) <-> row-major (R x ldp)
__global__ void cm_to_rm_kernel(const double *cm, long ld, int r0, int R, int m, double *rm, long ldp, int back)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)R * m) return;
    const int r = (int)(e / m), col = (int)(e % m);
    if (back) ((double *)cm)[(long)col * ld + r0 + r] = rm[(long)r * ldp + col];
    else rm[(long)r * ldp + col] = cm[(long)col * ld + r0 + r];
}

template <typename T>
__global__ void fill_kernel(T *p, long count, T v)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
        p[e] = v;
}

template <typename T>
__global__ void copy_col_kernel(const T *A, long lda, int n, int col, T *x)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) x[r] = A[(long)r * lda + col];
}

// ---------------------------------------------------------------------------
// Back substitution on the unit-upper U (lu2.c:110-116) in one persistent
// launch.  Rows are cut into 64-row blocks, solved bottom-up; block number
// idx (0 = bottom) belongs to workgroup idx % G (all G co-resident).  A
// workgroup accumulates U[block rows][solved block] * x[solved block] for
// every block below it as soon as the shared frontier (count of solved
// blocks) shows it solved -- the A tile of the next block it will need is
// loaded into registers before it waits -- then solves its 64x64 diagonal
// block right-looking in one wavefront (each finished x broadcast with
// readlane), publishes x with device-scope stores and advances the frontier.
// With 64-row blocks the adjacent block (the last one solved before this one)
// is left to wavefront 0: the other blocks' sums are reduced across lanes
// while that block is still being solved, its 64x64 tile waits transposed in
// LDS, and at the hand-over one wavefront applies it (lane = row) and goes on
// to the diagonal solve with no workgroup barrier in between.  x holds y on
// entry.  TF is the factor's element type, TX the vector's (FP64 for
// refinement).
constexpr int BSB = 64;

__device__ __forceinline__ double lane_bcast(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <typename TF, typename TX, int BS>
__global__ __launch_bounds__(BS * 4) void backsub_kernel(const TF *__restrict__ A, long lda, int n, TX *x,
                                                         unsigned long long *front, int *err, int strict)
{
    // BS-row blocks, BS/16 waves of 16 rows each; the columns of a solved
    // block are taken 64 at a time (lane = column), in BS/64 halves
    constexpr int NT = BS * 4, NH = BS / 64;
    constexpr bool ADJ = (BS == 64);   // the adjacent block applied by wavefront 0
    __shared__ TF Ud[BS][BS + 1];
    __shared__ TF UaT[ADJ ? BS : 1][BS + 1];   // UaT[k][row]: the tile against the adjacent block
    __shared__ TX part[BS];
    __shared__ int s_front;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nblk = (n + BS - 1) / BS;
    const int G = gridDim.x;
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(x, 0, (int)(sizeof(TX) * n), 0x00020000);
    for (int idx = blockIdx.x; idx < nblk; idx += G) {
        const int b = nblk - 1 - idx;
        const int r0 = b * BS, r1 = min(n, r0 + BS), nrb = r1 - r0;
        for (int e = t; e < BS * BS; e += NT) {
            const int rr = e / BS, kk = e % BS;
            Ud[rr][kk] = (rr < nrb && kk < nrb) ? A[(long)(r0 + rr) * lda + r0 + kk] : TF(0);
        }
        if constexpr (ADJ) {
            // columns [r0 + BS, r0 + 2 BS) of the block's rows (none for the bottom block)
            const int c0 = r0 + BS, nc = min(BS, n - c0);
            for (int e = t; e < BS * BS; e += NT) {
                const int rr = e / BS, kk = e % BS;
                UaT[kk][rr] = (rr < nrb && kk < nc) ? A[(long)(r0 + rr) * lda + c0 + kk] : TF(0);
            }
        }
        // wave w accumulates rows r0 + 16w .. r0 + 16w + 15, lane = column in the solved block
        TX acc[16];
        TF pa[NH][16];
#pragma unroll
        for (int i = 0; i < 16; i++) acc[i] = TX(0);
        // the 64 columns h of solved block d for this thread: rows r0 + 16w + i,
        // column c0(d) + 64 h + lane
        auto fetch = [&](int d, int h, TF *v) {
            const int col = (nblk - 1 - d) * BS + 64 * h + lane;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = r0 + w * 16 + i;
                v[i] = (r < r1 && col < n) ? A[(long)r * lda + col] : TF(0);
            }
        };
        auto accumulate = [&](int d, int h, const TF *v) {
            const int col = (nblk - 1 - d) * BS + 64 * h + lane;
            const TX xv = (col < n) ? ld_buf(xr, (int)(sizeof(TX) * col), TX(0)) : TX(0);
#pragma unroll
            for (int i = 0; i < 16; i++) acc[i] += (TX)v[i] * xv;
        };
        // this block's y, loaded before the waits: only this workgroup writes
        // these rows of x, and only once it solves the block
        TX yv[NH];
        if (t < 64)
#pragma unroll
            for (int h = 0; h < NH; h++) yv[h] = (t + 64 * h < nrb) ? x[r0 + t + 64 * h] : TX(0);
        // solved blocks summed by all waves (lane = column): all below this
        // one, or all but the adjacent one
        const int need = ADJ ? idx - 1 : idx;
        int done = 0;
        if (need > 0)
#pragma unroll
            for (int h = 0; h < NH; h++) fetch(0, h, pa[h]);
        while (done < need) {
            if (t == 0) {
                unsigned long spins = 0;
                int f;
                while ((f = (int)ld_relaxed(front)) <= done) {
                    __builtin_amdgcn_s_sleep(1);
                    if (((++spins) & 1023) == 0) {
                        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { f = idx; break; }
                        if (spins > (1UL << 25)) {
                            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            f = idx;
                            break;
                        }
                    }
                }
                s_front = f;
            }
            __syncthreads();
            const int F = min(s_front, need);
            acquire_seq(strict);
#pragma unroll
            for (int h = 0; h < NH; h++) accumulate(done, h, pa[h]);   // the prefetched tile
            int d = done + 1;
            if (NH == 1) {
                for (; d + 1 < F; d += 2) {      // blocks already solved: two tiles in flight
                    TF v0[16], v1[16];
                    fetch(d, 0, v0);
                    fetch(d + 1, 0, v1);
                    accumulate(d, 0, v0);
                    accumulate(d + 1, 0, v1);
                }
            }
            for (; d < F; d++) {
                TF v0[16], v1[16];
                fetch(d, 0, v0);
                if (NH > 1) fetch(d, NH - 1, v1);
                accumulate(d, 0, v0);
                if (NH > 1) accumulate(d, NH - 1, v1);
            }
            done = F;
            if (done < need)
#pragma unroll
                for (int h = 0; h < NH; h++) fetch(done, h, pa[h]);   // ahead of the next wait
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            TX v = acc[i];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) part[w * 16 + i] = v;
        }
        __syncthreads();
        if (t < 64) {
            // right-looking: x[rr] is final when step rr begins; rows above
            // subtract U[row][rr] x[rr]; lane t holds rows t + 64 h
            TX xt[NH];
#pragma unroll
            for (int h = 0; h < NH; h++) xt[h] = (t + 64 * h < nrb) ? yv[h] - part[t + 64 * h] : TX(0);
            if constexpr (ADJ) {
                if (idx > 0) {
                    // the adjacent block: this wavefront alone waits for it
                    unsigned long spins = 0;
                    while ((int)ld_relaxed(front) < idx) {
                        __builtin_amdgcn_s_sleep(1);
                        if (((++spins) & 1023) == 0) {
                            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                            if (spins > (1UL << 25)) {
                                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                break;
                            }
                        }
                    }
                    acquire_seq(strict);
                    const int ca = r0 + BS + lane;
                    const TX xa = (ca < n) ? ld_buf(xr, (int)(sizeof(TX) * ca), TX(0)) : TX(0);
                    TX s0 = TX(0), s1 = TX(0);
#pragma unroll
                    for (int k = 0; k < BS; k += 2) {
                        s0 += (TX)UaT[k][lane] * lane_bcast(xa, k);
                        s1 += (TX)UaT[k + 1][lane] * lane_bcast(xa, k + 1);
                    }
                    xt[0] -= s0 + s1;
                }
            }
            for (int rr = nrb - 1; rr > 0; rr--) {
                TX xk;
                if (NH == 1 || rr < 64) xk = lane_bcast(xt[0], rr & 63);
                else xk = lane_bcast(xt[NH - 1], rr & 63);
#pragma unroll
                for (int h = 0; h < NH; h++)
                    if (t + 64 * h < rr) xt[h] -= (TX)Ud[t + 64 * h][rr] * xk;
            }
#pragma unroll
            for (int h = 0; h < NH; h++)
                if (t + 64 * h < nrb) st_dev(x + r0 + t + 64 * h, xt[h]);
            // the block's x all come from this one wave, whose s_waitcnt
            // vmcnt(0) in publish_seq covers every lane's store
            if (!strict && t == 0) publish_seq(front, (unsigned long long)(idx + 1), 0);
        }
        __syncthreads();
        // strict: the workgroup barrier orders the other lanes' stores before
        // thread 0's release store
        if (strict && t == 0) publish_seq(front, (unsigned long long)(idx + 1), 1);
    }
}

// ---------------------------------------------------------------------------
// FP64 residual: per row |(A x)_i - b_i| and sum_j |a_ij|; per column sum |a_ij|.
// A wavefront takes RESID_ROWS rows, so each x[c] it loads serves all of them;
// each row is summed in the same order as by a wavefront of its own.
constexpr int RESID_ROWS = 4;
__global__ __launch_bounds__(256) void resid_rows_kernel(const double *A, long lda, int n, const double *x,
                                                         double *rerr, double *rabs)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long r0 = ((long)blockIdx.x * 4 + wave) * RESID_ROWS;
    if (r0 >= n) return;
    const int nr = (int)min((long)RESID_ROWS, (long)n - r0);
    const double *a = A + r0 * lda;
    double s[RESID_ROWS], sa[RESID_ROWS];
#pragma unroll
    for (int q = 0; q < RESID_ROWS; q++) s[q] = sa[q] = 0.0;
    for (int c = lane; c < n; c += 64) {
        const double xc = x[c];
#pragma unroll
        for (int q = 0; q < RESID_ROWS; q++) {
            if (q < nr) {
                const double e = a[q * lda + c];
                s[q] = fma(e, xc, s[q]);
                sa[q] += fabs(e);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RESID_ROWS; q++) {
        for (int off = 32; off > 0; off >>= 1) {
            s[q] += __shfl_xor(s[q], off);
            sa[q] += __shfl_xor(sa[q], off);
        }
    }
    if (lane == 0) {
        for (int q = 0; q < nr; q++) {
            rerr[r0 + q] = fabs(s[q] - a[q * lda + n]);
            rabs[r0 + q] = sa[q];
        }
    }
}

// A workgroup takes 64 columns; its 4 waves take the rows r = w (mod 4), each
// lane four rows per step (four loads in flight), and the 16 partial sums per
// column are added in a fixed order, so the result does not depend on timing.
__global__ __launch_bounds__(256) void resid_cols_kernel(const double *A, long lda, int n, double *cabs)
{
    __shared__ double part[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = blockIdx.x * 64 + lane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (col < n) {
        const double *a = A + col;
        long r = wave;
        for (; r + 12 < n; r += 16) {
            s0 += fabs(a[r * lda]);
            s1 += fabs(a[(r + 4) * lda]);
            s2 += fabs(a[(r + 8) * lda]);
            s3 += fabs(a[(r + 12) * lda]);
        }
        if (r < n) s0 += fabs(a[r * lda]);
        if (r + 4 < n) s1 += fabs(a[(r + 4) * lda]);
        if (r + 8 < n) s2 += fabs(a[(r + 8) * lda]);
    }
    part[wave][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wave == 0 && col < n) cabs[col] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// out[0]=max rerr, [1]=max rabs, [2]=max cabs, [3]=max|x|, [4]=sum|x|, [5]=max|b|
__global__ __launch_bounds__(256) void resid_final_kernel(const double *A, long lda, int n, const double *x,
                                                          const double *rerr, const double *rabs,
                                                          const double *cabs, double *out)
{
    __shared__ double sh[6][256];
    const int t = threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int r = t; r < n; r += 256) {
        v[0] = nan_max(v[0], rerr[r]);
        v[1] = nan_max(v[1], rabs[r]);
        v[2] = nan_max(v[2], cabs[r]);
        v[3] = nan_max(v[3], fabs(x[r]));
        v[4] += fabs(x[r]);
        v[5] = nan_max(v[5], fabs(A[(long)r * lda + n]));
    }
    for (int k = 0; k < 6; k++) sh[k][t] = v[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int k = 0; k < 6; k++) sh[k][t] = (k == 4) ? sh[k][t] + sh[k][t + st] : nan_max(sh[k][t], sh[k][t + st]);
        __syncthreads();
    }
    if (t < 6) out[t] = sh[t][0];
    if (t == 0) {
        // the scaled forms: HPL-2 with eps = 2^-53, and lu2_mpi.c:2170-2229's
        // resid1..3 with lu2's eps = 2e-16
        const double eps = 1.1102230246251565e-16, eps_lu2 = 2e-16, dn = (double)n;
        const double r = sh[0][0], ainf = sh[1][0], a1 = sh[2][0], xinf = sh[3][0], x1 = sh[4][0], bmax = sh[5][0];
        out[6] = r / (eps * (ainf * xinf + bmax) * dn);
        out[7] = r / (eps_lu2 * a1 * dn);
        out[8] = r / (eps_lu2 * a1 * x1);
        out[9] = r / (eps_lu2 * ainf * xinf * dn);
    }
}

// Scratch of the residual entry point that allocates nothing
// (lu2mi_residual_dev_noalloc): per-row |(Ax)_i - b_i| and sum_j |a_ij|, per
// column sum_i |a_ij| -- 3n doubles for n up to LU2MI_RESID_NMAX.  Each
// device holds its own copy until the process ends; host code reaches it only
// through the address lu2mi_scratch_address() obtains (ctx->d_scratch), and
// calls that overlap on one device share it (they race).
static __device__ __align__(256) double scratch[3 * (size_t)LU2MI_RESID_NMAX];

// ===========================================================================
// host side
// ===========================================================================
static char g_errmsg[512];
void lu2mi_set_error_msg(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof g_errmsg, fmt, ap);
    va_end(ap);
}
extern "C" const char *lu2mi_last_error(void) { return g_errmsg; }

int lu2mi_dev_buf(void **p, size_t *cur, size_t need)
{
    if (*cur >= need && *p) return LU2MI_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cur = 0;
    if (need == 0) return LU2MI_OK;
    if (hipMalloc(p, need) != hipSuccess) {
        *p = nullptr;
        lu2mi_set_error_msg("hipMalloc(%zu) failed", need);
        return LU2MI_ERR_NOMEM;
    }
    *cur = need;
    return LU2MI_OK;
}

int lu2mi_ctx_reserve(lu2mi_ctx *ctx, int n, int nb, int esz)
{
    (void)esz;
    if (n <= ctx->n_cap && nb <= ctx->nb_cap && ctx->d_w) return LU2MI_OK;
    const int ncap = n > ctx->n_cap ? n : ctx->n_cap;
    const int nbcap = nb > ctx->nb_cap ? nb : ctx->nb_cap;
    (void)hipFree(ctx->d_piv); (void)hipFree(ctx->d_orig); (void)hipFree(ctx->d_invd);
    (void)hipFree(ctx->d_invd_f); (void)hipFree(ctx->d_pairs); (void)hipFree(ctx->d_linv);
    (void)hipFree(ctx->d_tmp); (void)hipFree(ctx->d_y);
    ctx->d_piv = nullptr; ctx->d_orig = nullptr; ctx->d_invd = nullptr; ctx->d_invd_f = nullptr;
    ctx->d_pairs = nullptr; ctx->d_linv = nullptr; ctx->d_tmp = nullptr; ctx->d_y = nullptr;
    LU2MI_CHECK(hipMalloc(&ctx->d_piv, sizeof(int) * (size_t)ncap));
    LU2MI_CHECK(hipMalloc(&ctx->d_orig, sizeof(int) * (size_t)ncap));
    LU2MI_CHECK(hipMalloc(&ctx->d_invd, sizeof(double) * (size_t)ncap));
    LU2MI_CHECK(hipMalloc(&ctx->d_invd_f, sizeof(float) * (size_t)ncap));
    // pairs and Linv: two of each (blocks i and i+1 alternate, so the next
    // panel can be factored while block i's rows are still being moved);
    // d_tmp: the panel's h x h product scratch, then its h x h staging rows
    LU2MI_CHECK(hipMalloc(&ctx->d_pairs, sizeof(int) * 2 * 4 * (size_t)(nbcap > 64 ? nbcap : 64)));
    LU2MI_CHECK(hipMalloc(&ctx->d_linv, sizeof(double) * 2 * (size_t)nbcap * nbcap));
    LU2MI_CHECK(hipMalloc(&ctx->d_tmp, sizeof(double) * 2 * (size_t)nbcap * nbcap));
    LU2MI_CHECK(hipMalloc(&ctx->d_y, sizeof(double) * 4 * (size_t)ncap + 64));
    size_t wneed = sizeof(double) * (size_t)nbcap * (size_t)(ncap + 1);
    LU2MI_TRY(lu2mi_dev_buf(&ctx->d_w, &ctx->w_bytes, wneed));
    // leaf exchange slots, then the sub-panel grid barrier word (both cleared
    // at the start of every factorization)
    size_t sneed = SLOTS_BYTES + 256;
    LU2MI_TRY(lu2mi_dev_buf(&ctx->d_slots, &ctx->slots_bytes, sneed));
    ctx->n_cap = ncap;
    ctx->nb_cap = nbcap;
    return LU2MI_OK;
}

// ---------------------------------------------------------------------------
template <typename T> struct Prec;
template <> struct Prec<double> {
    static int gemm(const GemmDesc &g, hipStream_t s) { return lu2mi_gemm_f64(g, s); }
    static double *invd(lu2mi_ctx *c) { return c->d_invd; }
};
template <> struct Prec<float> {
    static int gemm(const GemmDesc &g, hipStream_t s) { return lu2mi_gemm_f32(g, s); }
    static float *invd(lu2mi_ctx *c) { return c->d_invd_f; }
};

template <typename T>
static int gemm_call(lu2mi_ctx *ctx, int M, int N, int K, double alpha, const T *A, long lda,
                     const T *B, long ldb, double beta, T *C, long ldc, int a_lower, int tag = 0)
{
    GemmDesc g;
    g.M = M; g.N = N; g.K = K; g.batch = 1;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.sA = 0;
    g.B = B; g.ldb = ldb; g.sB = 0;
    g.C = C; g.ldc = ldc; g.sC = 0;
    g.a_lower = a_lower;
    g.tag = tag;
    return Prec<T>::gemm(g, ctx->stream);
}

// x <- U^-1 x with the unit-upper U stored in A (one persistent launch)
template <typename TF, typename TX, int BS>
static int backsub_bs(lu2mi_ctx *ctx, const TF *A, long lda, int n, TX *x)
{
    const int nblk = (n + BS - 1) / BS;
    int per_cu = 0;
    LU2MI_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)backsub_kernel<TF, TX, BS>, 4 * BS, 0));
    // every workgroup the device keeps co-resident: more workgroups mean
    // shorter backlogs of solved tiles each
    int G = per_cu * ctx->num_cu < nblk ? per_cu * ctx->num_cu : nblk;
    if (G < 1) G = 1;
    LU2MI_CHECK(hipMemsetAsync(ctx->d_bar, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL((backsub_kernel<TF, TX, BS>), dim3(G), dim3(4 * BS), 0, ctx->stream, A, lda, n, x, ctx->d_bar,
                       ctx->d_err, ctx->strict_sync);
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

// 64-row blocks, or 128 (LU2MI_OPT_BACKSUB128): half as many frontier steps,
// each a cross-workgroup hand-over, for a longer diagonal solve per step
template <typename TF, typename TX>
static int backsub(lu2mi_ctx *ctx, const TF *A, long lda, int n, TX *x)
{
    if (ctx->backsub128) return backsub_bs<TF, TX, 128>(ctx, A, lda, n, x);
    return backsub_bs<TF, TX, BSB>(ctx, A, lda, n, x);
}

template <typename T>
struct Factor {
    lu2mi_ctx *ctx;
    T *A;
    long lda;
    int n, nb, compat, W;
    T *linv, *tmp, *wbuf;
    T *pwbuf;              // the panel's staging rows (wbuf holds the block's)
    int *pairs;
    T *linv_base = nullptr;   // two nb x nb inverses, alternating by block
    int *pairs_base = nullptr;
    // when set, every block keeps its inverse diagonal block and its row
    // pairs here (for later solves with the stored factors)
    T *linv_all = nullptr;
    int *pairs_all = nullptr;

    int leaf(int i, int c)
    {
        const int R = n - c;
        // rows per thread: at most 64 workgroups for the 16-wide leaf (fewer
        // slots to gather per column), 128 for the 32-wide one
        const int gcap = (W == 16) ? 64 : 128;
        int rpt = (R + NTL * gcap - 1) / (NTL * gcap);
        if (rpt < 1) rpt = 1;
        if (W == 32 && rpt > 2) rpt = 2;
        if (W == 16 && rpt > 4) rpt = 4;
        if (rpt == 3) rpt = 4;
        const int G = (R + NTL * rpt - 1) / (NTL * rpt);
        if (G > LEAF_GMAX) {
            lu2mi_set_error_msg("leaf: %d rows need %d workgroups (> %d)", R, G, LEAF_GMAX);
            return LU2MI_ERR_UNSUPPORTED;
        }
        LeafArgs<T> la;
        la.A = A; la.lda = lda; la.n = n; la.c = c; la.i = i; la.compat = compat; la.G = G;
        la.piv = ctx->d_piv; la.orig = ctx->d_orig; la.invd = Prec<T>::invd(ctx);
        la.pair_dst = pairs; la.pair_src = pairs + 2 * W;
        la.linv = linv; la.ldl = nb;
        la.slots = (unsigned long long *)ctx->d_slots;
        la.seq_base = ctx->bar_count; la.err = ctx->d_err; la.strict = ctx->strict_sync;
        void *args[] = {&la};
        const void *fn = nullptr;
        if (W == 32) fn = (rpt == 1) ? (const void *)leaf_kernel<T, 32, 1> : (const void *)leaf_kernel<T, 32, 2>;
        else fn = (rpt == 1) ? (const void *)leaf_kernel<T, 16, 1>
                 : (rpt == 2) ? (const void *)leaf_kernel<T, 16, 2> : (const void *)leaf_kernel<T, 16, 4>;
        // All G workgroups must be resident together (they meet at the grid
        // barrier).  An ordinary launch on an otherwise idle device places
        // them all when occupancy allows it; the runtime's cooperative launch
        // is not used because it holds a per-device resource that a second
        // process on the same GPU cannot get.  The barrier's bounded spin
        // turns any shortfall into LU2MI_ERR_DEVICE rather than a hang.
        int per_cu = 0;
        LU2MI_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NTL, 0));
        if ((long)per_cu * ctx->num_cu < G) {
            lu2mi_set_error_msg("leaf: %d workgroups cannot be co-resident (%d per CU x %d CUs)", G, per_cu,
                                ctx->num_cu);
            return LU2MI_ERR_UNSUPPORTED;
        }
        LU2MI_CHECK(hipLaunchKernel(fn, dim3(G), dim3(NTL), args, 0, ctx->stream));
        ctx->bar_count += (unsigned long long)W;
        // apply the leaf's swaps to the other panel columns; L part (columns
        // left of the leaf) of the leaf's diagonal rows scaled by 1/d
        return rowperm(pairs, pairs + 2 * W, 2 * W, i, i + nb, c, c + W, c, c, c + W, nullptr, 0, 0, 0);
    }

    int rowperm(const int *dst, const int *src, int np, int c0, int c1, int k0, int k1, int sce, int d0,
                int d1, T *Wb, long ldw, int w0, int w1)
    {
        if (c1 <= c0) return LU2MI_OK;
        static bool lds_attr = false;
        if (!lds_attr) {
            LU2MI_CHECK(hipFuncSetAttribute((const void *)rowperm_kernel<T>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
            lds_attr = true;
        }
        PermArgs<T> p;
        p.A = A; p.lda = lda; p.dst = dst; p.src = src; p.np = np;
        p.c0 = c0; p.c1 = c1; p.k0 = k0; p.k1 = k1; p.sce = sce; p.d0 = d0; p.d1 = d1;
        p.invd = Prec<T>::invd(ctx); p.W = Wb; p.ldw = ldw; p.w0 = w0; p.w1 = w1;
        if (Wb && w0 == d0 && w1 == d1 && k0 == k1) {
            // block level: destinations in [w0, w1) are the block's rows
            p.sw = 0;
            const dim3 grid((c1 - c0 + 1023) / 1024, np);
            hipLaunchKernelGGL(staged_perm_kernel<T>, grid, dim3(256), 0, ctx->stream, p, 0);
            hipLaunchKernelGGL(staged_perm_kernel<T>, grid, dim3(256), 0, ctx->stream, p, 1);
            LU2MI_CHECK(hipGetLastError());
            return LU2MI_OK;
        }
        // strip width: every source row of the strip fits in LDS; strips of at
        // least 64 bytes per row, and narrow enough for ~2 strips per CU
        const int seg = 64 / (int)sizeof(T);
        int sw = (c1 - c0 + 2 * ctx->num_cu - 1) / (2 * ctx->num_cu);
        sw = (sw + seg - 1) / seg * seg;
        if (sw > 128) sw = 128;
        const int swmax = (int)(131072 / (np * (int)sizeof(T)));
        if (sw > swmax) sw = swmax;
        if (sw < 1) {
            lu2mi_set_error_msg("rowperm: %d row pairs do not fit in LDS", np);
            return LU2MI_ERR_UNSUPPORTED;
        }
        p.sw = sw;
        const int strips = (c1 - c0 + sw - 1) / sw;
        const size_t shm = (size_t)np * sw * sizeof(T);
        hipLaunchKernelGGL(rowperm_kernel<T>, dim3(strips), dim3(256), shm, ctx->stream, p);
        LU2MI_CHECK(hipGetLastError());
        return LU2MI_OK;
    }

    // rows per thread and workgroups of a 16-wide leaf over R rows: at most
    // 64 workgroups (fewer slots to gather per column), up to 4 rows each
    static void leaf16_shape(int R, int &rpt, int &G)
    {
        rpt = (R + NTL * 64 - 1) / (NTL * 64);
        if (rpt < 1) rpt = 1;
        if (rpt > 4) rpt = 4;
        if (rpt == 3) rpt = 4;
        G = (R + NTL * rpt - 1) / (NTL * rpt);
    }

    // the sub-panel kernel applies to 16-wide leaves on 16-byte aligned rows
    bool sub_ok(int w) const
    {
        constexpr long EPV = 16 / sizeof(T);
        auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
        return ctx->sub_w > 0 && W == SUB_W && w <= ctx->sub_w && w <= SUB_WMAX && w >= W && al(A) &&
               lda % EPV == 0 && nb % EPV == 0 && al(linv) && al(pwbuf) && al(tmp);
    }

    int subpanel(int i, int c, int w)
    {
        int rpt, G;
        leaf16_shape(n - c, rpt, G);
        if (G > LEAF_GMAX) {
            lu2mi_set_error_msg("sub-panel: %d rows need %d workgroups (> %d)", n - c, G, LEAF_GMAX);
            return LU2MI_ERR_UNSUPPORTED;
        }
        SubArgs<T> sa;
        LeafArgs<T> &la = sa.leaf;
        la.A = A; la.lda = lda; la.n = n; la.c = c; la.i = i; la.compat = compat; la.G = G;
        la.piv = ctx->d_piv; la.orig = ctx->d_orig; la.invd = Prec<T>::invd(ctx);
        la.pair_dst = pairs; la.pair_src = pairs + 2 * SUB_W;
        la.linv = linv; la.ldl = nb;
        la.slots = (unsigned long long *)ctx->d_slots;
        la.seq_base = ctx->bar_count; la.err = ctx->d_err; la.strict = ctx->strict_sync;
        sa.nb = nb; sa.w = w;
        // strips of at least 64 bytes per row, about one per workgroup
        const int seg = 64 / (int)sizeof(T);
        int sw = (nb + G - 1) / G;
        sw = (sw + seg - 1) / seg * seg;
        sa.sw = sw > SUB_SWMAX ? SUB_SWMAX : sw;
        sa.wbuf = pwbuf; sa.tmp = tmp;
        sa.gbar = (unsigned long long *)((char *)ctx->d_slots + SLOTS_BYTES);
        sa.gbar_base = ctx->gbar_count;
        const void *fn = (rpt == 1) ? (const void *)subpanel_kernel<T, 1>
                         : (rpt == 2) ? (const void *)subpanel_kernel<T, 2> : (const void *)subpanel_kernel<T, 4>;
        // all G workgroups must be co-resident (grid barriers), as for the leaf
        int per_cu = 0;
        LU2MI_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NTL, 0));
        if ((long)per_cu * ctx->num_cu < G) {
            lu2mi_set_error_msg("sub-panel: %d workgroups cannot be co-resident (%d per CU x %d CUs)", G, per_cu,
                                ctx->num_cu);
            return LU2MI_ERR_UNSUPPORTED;
        }
        void *args[] = {&sa};
        LU2MI_CHECK(hipLaunchKernel(fn, dim3(G), dim3(NTL), args, 0, ctx->stream));
        ctx->bar_count += (unsigned long long)w;   // W sequence numbers per leaf
        ctx->gbar_count += (unsigned long long)G * sub_barriers(w / SUB_W);
        return LU2MI_OK;
    }

    int panel(int i, int c, int w)
    {
        if (sub_ok(w)) return subpanel(i, c, w);
        if (w <= W) return leaf(i, c);
        const int h = w / 2;
        LU2MI_TRY(panel(i, c, h));
        const long off = (long)(c - i);
        if (h <= lu2mi_gemm::BT64) {   // one workgroup per step (see node_trsm64_kernel)
            hipLaunchKernelGGL(node_trsm64_kernel<T>, dim3(1), dim3(256), 0, ctx->stream, A, lda, c, h, linv, nb, off,
                               Prec<T>::invd(ctx));
            LU2MI_CHECK(hipGetLastError());
            LU2MI_TRY(gemm_call<T>(ctx, n - c - h, h, h, -1.0, A + (long)(c + h) * lda + c, lda,
                                   A + (long)c * lda + c + h, lda, 1.0, A + (long)(c + h) * lda + c + h, lda, 0));
            LU2MI_TRY(panel(i, c + h, h));
            hipLaunchKernelGGL(node_inv64_kernel<T>, dim3(1), dim3(256), 0, ctx->stream, A, lda, c, h, linv, nb, off,
                               tmp);
            LU2MI_CHECK(hipGetLastError());
            return LU2MI_OK;
        }
        // right half of this node: rows [c, c+h) scaled by 1/d, then
        // X = inv(D^-1 L11) * scaled rows, then the rows below
        {
            const long e = (long)h * h;
            hipLaunchKernelGGL(scale_copy_kernel<T>, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, ctx->stream,
                               A, lda, c, c + h, h, h, Prec<T>::invd(ctx), pwbuf);
            LU2MI_CHECK(hipGetLastError());
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, 1.0, linv + off * nb + off, nb, pwbuf, h, 0.0,
                                   A + (long)c * lda + c + h, lda, 1));
            LU2MI_TRY(gemm_call<T>(ctx, n - c - h, h, h, -1.0, A + (long)(c + h) * lda + c, lda,
                                   A + (long)c * lda + c + h, lda, 1.0, A + (long)(c + h) * lda + c + h, lda, 0));
        }
        LU2MI_TRY(panel(i, c + h, h));
        // off-diagonal block of the inverse: X21 = -inv22 * L21 * inv11
        {
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, 1.0, A + (long)(c + h) * lda + c, lda, linv + off * nb + off, nb,
                                   0.0, tmp, h, 0));
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, -1.0, linv + (off + h) * nb + off + h, nb, tmp, h, 0.0,
                                   linv + (off + h) * nb + off, nb, 1));
        }
        return LU2MI_OK;
    }

    // scratch pointers of block i (per-block copies when the factors are kept)
    void set_block(int i)
    {
        if (linv_all) {
            linv = linv_all + (long)(i / nb) * nb * nb;
            pairs = pairs_all + (long)(i / nb) * 4 * nb;
        } else if (linv_base) {
            linv = linv_base + (long)((i / nb) & 1) * nb * nb;
            pairs = pairs_base + (long)((i / nb) & 1) * 4 * nb;
        }
    }

    // HIP events around one subsystem's launches on the current stream
    // (recorded only with timing on; slot -1 = not recorded)
    int tbegin(int cat, int &slot)
    {
        slot = -1;
        KernelTimer &kt = ctx->tcat[cat];
        if (!ctx->timing_on || kt.used >= kt.cap) return LU2MI_OK;
        slot = kt.used++;
        LU2MI_CHECK(hipEventRecord(kt.ev[2 * slot], ctx->stream));
        return LU2MI_OK;
    }
    int tend(int cat, int slot, double flops, double bytes)
    {
        if (slot < 0) return LU2MI_OK;
        KernelTimer &kt = ctx->tcat[cat];
        LU2MI_CHECK(hipEventRecord(kt.ev[2 * slot + 1], ctx->stream));
        kt.flops[slot] = flops;
        kt.bytes[slot] = bytes;
        return LU2MI_OK;
    }

    int start_panel(int i)
    {
        set_block(i);
        int ts;
        LU2MI_TRY(tbegin(LU2MI_T_PANEL, ts));
        hipLaunchKernelGGL(fill_kernel<T>, dim3(256), dim3(256), 0, ctx->stream, linv, (long)nb * nb, T(0));
        LU2MI_CHECK(hipGetLastError());
        LU2MI_TRY(panel(i, i, nb));
        return tend(LU2MI_T_PANEL, ts, (double)(n - i) * nb * nb, 0.0);
    }

    // block-level row permutation of columns [c0, c1) with its timing
    int block_perm(int i, int c0, int c1, T *wb, long ldw)
    {
        int ts;
        LU2MI_TRY(tbegin(LU2MI_T_SWAP, ts));
        LU2MI_TRY(rowperm(pairs, pairs + 2 * nb, 2 * nb, c0, c1, 0, 0, n + 1, i, i + nb, wb, ldw, i, i + nb));
        return tend(LU2MI_T_SWAP, ts, 0.0, 32.0 * nb * (double)(c1 - c0));
    }

    // U12 of columns [c0, c0+N) = inv(D^-1 L11) * staged rows (lu2.c:511-512)
    int u12(int i, int c0, int N, const T *wb, long ldw)
    {
        int ts;
        LU2MI_TRY(tbegin(LU2MI_T_U12, ts));
        LU2MI_TRY(gemm_call<T>(ctx, nb, N, nb, 1.0, linv, nb, wb, ldw, 0.0, A + (long)i * lda + c0, lda, 1,
                               (sizeof(T) == 8 && ctx->trail_kernel) ? 1 : 0));
        return tend(LU2MI_T_U12, ts, (double)nb * nb * N, 0.0);
    }

    // A22[:, c0 .. c0+N) -= L21 * U12[:, c0 .. c0+N) for block i (lu2.c:580), timed
    int trailing(int i, int c0, int N)
    {
        const int M = n - i - nb;
        int ts;
        LU2MI_TRY(tbegin(LU2MI_T_TRAIL, ts));
        if (sizeof(T) == 4 && ctx->lowp)   // operands rounded to FP16/BF16, FP32 accumulation
            LU2MI_TRY(lu2mi_update16(ctx, ctx->lowp, M, N, nb, (const float *)(A + (long)(i + nb) * lda + i), lda,
                                     (const float *)(A + (long)i * lda + c0), lda,
                                     (float *)(A + (long)(i + nb) * lda + c0), lda));
        else
            LU2MI_TRY(gemm_call<T>(ctx, M, N, nb, -1.0, A + (long)(i + nb) * lda + i, lda, A + (long)i * lda + c0,
                                   lda, 1.0, A + (long)(i + nb) * lda + c0, lda, 0, ctx->trail_kernel ? 1 : 2));
        return tend(LU2MI_T_TRAIL, ts, 2.0 * M * (double)N * nb, 0.0);
    }

    // Block loop of lu2.c:1122-1144.  With lookahead, block i's row
    // permutation, U12 product and trailing update are each split: the next
    // panel's nb columns first, then the rest; the next panel is factored on
    // the second stream as soon as its columns are up to date, while the rest
    // of block i runs.  The two touch disjoint columns; the next panel has its
    // own Linv and pair buffers (they alternate by block) and its own staging
    // rows, so block i's stay intact until its rest is done.
    int run(int *dpiv_out, T *dx)
    {
        const bool la = ctx->lookahead != 0;
        hipStream_t s1 = ctx->stream;
        bool panel_pending = false;   // panel(i) issued on the second stream
        LU2MI_TRY(start_panel(0));
        for (int i = 0; i < n; i += nb) {
            if (panel_pending) {
                LU2MI_CHECK(hipStreamWaitEvent(s1, ctx->ev_panel, 0));
                panel_pending = false;
            }
            set_block(i);
            int *const pairs_i = pairs;
            const int ncols = n + 1 - (i + nb);
            const int M = n - i - nb;
            const int na = (la && M > 0) ? nb : ncols;   // columns before the split
            // block-level row permutation over [i+nb, n], 1/d scaling of the
            // block rows, block rows staged in w (lu2.c:502-509 awork copy)
            hipLaunchKernelGGL(block_pairs_kernel, dim3((nb + 255) / 256), dim3(256), 0, s1, ctx->d_piv,
                               ctx->d_orig, i, nb, pairs_i, pairs_i + 2 * nb);
            LU2MI_CHECK(hipGetLastError());
            const long ldw = (ncols + 3) & ~3L;   // 16-byte rows for FP64 and FP32
            LU2MI_TRY(block_perm(i, i + nb, i + nb + na, wbuf, ldw));
            // U12 = inv(D^-1 L11) * staged rows   (lu2.c:511-512)
            LU2MI_TRY(u12(i, i + nb, na, wbuf, ldw));
            if (M <= 0) continue;
            if (la) {
                LU2MI_TRY(trailing(i, i + nb, nb));                    // the next panel's columns
                LU2MI_CHECK(hipEventRecord(ctx->ev_next, s1));
                LU2MI_CHECK(hipStreamWaitEvent(ctx->stream2, ctx->ev_next, 0));
                ctx->stream = ctx->stream2;                           // panel(i+nb) on the second stream
                const int prc = start_panel(i + nb);
                ctx->stream = s1;
                LU2MI_TRY(prc);
                LU2MI_CHECK(hipEventRecord(ctx->ev_panel, ctx->stream2));
                panel_pending = true;
                // the rest of block i, b included: staging rows from column nb
                // of w on, so that w keeps the layout of the unsplit product
                set_block(i);
                LU2MI_TRY(block_perm(i, i + 2 * nb, n + 1, wbuf + nb, ldw));
                LU2MI_TRY(u12(i, i + 2 * nb, ncols - nb, wbuf + nb, ldw));
                LU2MI_TRY(trailing(i, i + 2 * nb, ncols - nb));
            } else {
                LU2MI_TRY(trailing(i, i + nb, ncols));
                LU2MI_TRY(start_panel(i + nb));
            }
        }
        if (panel_pending) LU2MI_CHECK(hipStreamWaitEvent(s1, ctx->ev_panel, 0));
        // x = column n, then unit-upper back substitution
        int ts;
        LU2MI_TRY(tbegin(LU2MI_T_BACKSUB, ts));
        hipLaunchKernelGGL(copy_col_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A, lda, n, n, dx);
        LU2MI_CHECK(hipGetLastError());
        LU2MI_TRY(backsub(ctx, A, lda, n, dx));
        LU2MI_TRY(tend(LU2MI_T_BACKSUB, ts, (double)n * n, 4.0 * (double)n * n));
        if (dpiv_out && dpiv_out != ctx->d_piv)
            LU2MI_CHECK(hipMemcpyAsync(dpiv_out, ctx->d_piv, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
        return LU2MI_OK;
    }
};

template <typename T>
static int factor_impl(lu2mi_ctx *ctx, int n, T *dA, long lda, int nb, int pivmode, int *dpiv_out, T *dx,
                       T *keep_linv = nullptr, int *keep_pairs = nullptr)
{
    if (!ctx || !dA || n <= 0 || lda < n + 1 || nb < 32 || nb > 4096 || (nb & (nb - 1)) || n % nb) {
        lu2mi_set_error_msg("lumcolumn: need nb a power of two in [32, 4096], n %% nb == 0, lda >= n+1 "
                            "(n=%d nb=%d lda=%ld)", n, nb, lda);
        return LU2MI_ERR_ARG;
    }
    LU2MI_TRY(lu2mi_ctx_reserve(ctx, n, nb, sizeof(T)));
    T *x = dx ? dx : (T *)ctx->d_y;
    Factor<T> f;
    f.ctx = ctx; f.A = dA; f.lda = lda; f.n = n; f.nb = nb;
    f.compat = (pivmode == LU2MI_PIV_COMPAT);
    f.W = (ctx->leaf_w == 16) ? 16 : 32;
    if (f.W > nb / 2) f.W = 16;
    // a W=32 leaf holds at most 2 rows per thread in registers over LEAF_GMAX workgroups
    if (f.W == 32 && n > 2 * NTL * LEAF_GMAX) f.W = 16;
    f.linv = (T *)ctx->d_linv; f.tmp = (T *)ctx->d_tmp; f.wbuf = (T *)ctx->d_w; f.pairs = ctx->d_pairs;
    f.pwbuf = (T *)((double *)ctx->d_tmp + (size_t)ctx->nb_cap * ctx->nb_cap);
    f.linv_base = (T *)ctx->d_linv; f.pairs_base = ctx->d_pairs;
    f.linv_all = keep_linv;
    f.pairs_all = keep_pairs;
    for (int c = 0; c < LU2MI_T_COUNT; c++) ctx->tcat[c].used = 0;
    LU2MI_CHECK(hipMemsetAsync(ctx->d_err, 0, sizeof(int), ctx->stream));
    // slot sequence numbers restart with every factorization
    LU2MI_CHECK(hipMemsetAsync(ctx->d_slots, 0, ctx->slots_bytes, ctx->stream));
    ctx->bar_count = 0;
    ctx->gbar_count = 0;
    auto t0 = std::chrono::steady_clock::now();
    int rc = f.run(dpiv_out, x);
    if (rc != LU2MI_OK) return rc;
    LU2MI_CHECK(hipStreamSynchronize(ctx->stream));
    auto t1 = std::chrono::steady_clock::now();
    ctx->t_last_ms[3] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    int herr = 0;
    LU2MI_CHECK(hipMemcpy(&herr, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        lu2mi_set_error_msg("leaf kernel grid barrier timed out");
        return LU2MI_ERR_DEVICE;
    }
    return LU2MI_OK;
}

int lu2mi_factor_f64(lu2mi_ctx *ctx, int n, double *dA, long lda, int nb, int pivmode, int *dpiv, double *dx)
{
    return factor_impl<double>(ctx, n, dA, lda, nb, pivmode, dpiv, dx);
}
int lu2mi_factor_f32(lu2mi_ctx *ctx, int n, float *dA, long lda, int nb, int pivmode, int *dpiv, float *dx,
                     float *keep_linv, int *keep_pairs)
{
    return factor_impl<float>(ctx, n, dA, lda, nb, pivmode, dpiv, dx, keep_linv, keep_pairs);
}
int lu2mi_factor_f64_keep(lu2mi_ctx *ctx, int n, double *dA, long lda, int nb, int pivmode, int *dpiv,
                          double *dx, double *keep_linv, int *keep_pairs)
{
    return factor_impl<double>(ctx, n, dA, lda, nb, pivmode, dpiv, dx, keep_linv, keep_pairs);
}
// lu2tlib.c:1130 cm_column_decomposition_recursive on a device copy: the
// panel's rows [i, n) of m columns (column-major, leading dimension ld)
// factored in place with the recursive panel; pv[k] = global pivot row.
int lu2mi_cm_panel_f64(lu2mi_ctx *ctx, double *dcm, long ld, int n, int m, int i, int pivmode, int *hpv)
{
    if (m != 16 && m != 32) {
        lu2mi_set_error_msg("cm panel: %d columns (16 or 32 supported)", m);
        return LU2MI_ERR_UNSUPPORTED;
    }
    const int R = n - i;
    if (R < m) {
        lu2mi_set_error_msg("cm panel: %d rows for %d columns", R, m);
        return LU2MI_ERR_ARG;
    }
    const long ldp = 2 * m;
    LU2MI_TRY(lu2mi_ctx_reserve(ctx, R, 2 * m > 64 ? 2 * m : 64, sizeof(double)));
    double *P = nullptr;
    LU2MI_CHECK(hipMalloc(&P, sizeof(double) * (size_t)R * ldp));
    const long tot = (long)R * m;
    hipLaunchKernelGGL(cm_to_rm_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, dcm, ld, i,
                       R, m, P, ldp, 0);
    Factor<double> f;
    f.ctx = ctx; f.A = P; f.lda = ldp; f.n = R; f.nb = m; f.compat = (pivmode == LU2MI_PIV_COMPAT); f.W = m;
    f.linv = (double *)ctx->d_linv; f.tmp = (double *)ctx->d_tmp; f.wbuf = (double *)ctx->d_w;
    f.pwbuf = (double *)ctx->d_tmp + (size_t)ctx->nb_cap * ctx->nb_cap;
    f.pairs = ctx->d_pairs;
    int rc = LU2MI_OK;
    if (hipMemsetAsync(ctx->d_err, 0, sizeof(int), ctx->stream) != hipSuccess ||
        hipMemsetAsync(ctx->d_slots, 0, ctx->slots_bytes, ctx->stream) != hipSuccess)
        rc = LU2MI_ERR_HIP;
    ctx->bar_count = 0;
    ctx->gbar_count = 0;
    if (rc == LU2MI_OK) rc = f.panel(0, 0, m);
    if (rc == LU2MI_OK) {
        hipLaunchKernelGGL(cm_to_rm_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, dcm, ld,
                           i, R, m, P, ldp, 1);
        int piv[32], herr = 0;
        if (hipMemcpyAsync(piv, ctx->d_piv, sizeof(int) * m, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(&herr, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = LU2MI_ERR_HIP;
        else if (herr)
            rc = LU2MI_ERR_DEVICE;
        else
            for (int k = 0; k < m; k++) hpv[k] = i + piv[k];
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(P);
    if (rc == LU2MI_ERR_HIP) lu2mi_set_error_msg("cm panel: HIP call failed");
    if (rc == LU2MI_ERR_DEVICE) lu2mi_set_error_msg("cm panel: device wait timed out");
    return rc;
}

int lu2mi_backsub_f32_f64(lu2mi_ctx *ctx, const float *A, long lda, int n, double *x)
{
    return backsub<float, double>(ctx, A, lda, n, x);
}
int lu2mi_backsub_f64_f64(lu2mi_ctx *ctx, const double *A, long lda, int n, double *x)
{
    return backsub<double, double>(ctx, A, lda, n, x);
}

// ---------------------------------------------------------------------------
int lu2mi_randomsetmat_impl(lu2mi_ctx *ctx, int n, int seed, double *dA, long lda)
{
    if (!ctx || !dA || n <= 0 || lda < n + 1) {
        lu2mi_set_error_msg("randomsetmat: bad arguments");
        return LU2MI_ERR_ARG;
    }
    const unsigned long long x0 = (((unsigned long long)(unsigned)seed) << 16) | 0x330EULL;
    hipLaunchKernelGGL(randomsetmat_kernel, dim3(n), dim3(256), 0, ctx->stream, dA, lda, n, x0);
    LU2MI_CHECK(hipGetLastError());
    LU2MI_CHECK(hipStreamSynchronize(ctx->stream));
    return LU2MI_OK;
}

int lu2mi_diag_dominant_impl(lu2mi_ctx *ctx, int n, double *dA, long lda)
{
    if (!dA || n <= 0 || lda < n + 1) {
        lu2mi_set_error_msg("make_diag_dominant: bad arguments");
        return LU2MI_ERR_ARG;
    }
    hipLaunchKernelGGL(diag_dominant_kernel, dim3(n), dim3(256), 0, ctx->stream, dA, lda, n);
    LU2MI_CHECK(hipGetLastError());
    LU2MI_CHECK(hipStreamSynchronize(ctx->stream));
    return LU2MI_OK;
}

// the three residual kernels on ctx's stream: buf holds 3n doubles of
// partial results, dout (device) receives the ten figures
static int residual_launch(lu2mi_ctx *ctx, int n, const double *dA0, long lda, const double *dx, double *buf,
                           double *dout)
{
    double *rerr = buf, *rabs = buf + n, *cabs = buf + 2 * (size_t)n;
    hipLaunchKernelGGL(resid_rows_kernel, dim3((n + 4 * RESID_ROWS - 1) / (4 * RESID_ROWS)), dim3(256), 0, ctx->stream,
                       dA0, lda, n, dx, rerr, rabs);
    hipLaunchKernelGGL(resid_cols_kernel, dim3((n + 63) / 64), dim3(256), 0, ctx->stream, dA0, lda, n, cabs);
    hipLaunchKernelGGL(resid_final_kernel, dim3(1), dim3(256), 0, ctx->stream, dA0, lda, n, dx, rerr, rabs, cabs,
                       dout);
    LU2MI_CHECK(hipGetLastError());
    return LU2MI_OK;
}

int lu2mi_residual_impl(lu2mi_ctx *ctx, int n, const double *dA0, long lda, const double *dx, double *out)
{
    if (!ctx || !dA0 || !dx || !out || n <= 0 || lda < n + 1) {
        lu2mi_set_error_msg("residual: bad arguments");
        return LU2MI_ERR_ARG;
    }
    double *buf = nullptr;
    LU2MI_CHECK(hipMalloc(&buf, sizeof(double) * (3 * (size_t)n + 16)));
    int rc = residual_launch(ctx, n, dA0, lda, dx, buf, buf + 3 * (size_t)n);
    hipError_t e = hipSuccess;
    if (rc == LU2MI_OK) {
        e = hipMemcpyAsync(out, buf + 3 * (size_t)n, 10 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    (void)hipFree(buf);
    if (rc != LU2MI_OK) return rc;
    if (e != hipSuccess) {
        lu2mi_set_error_msg("residual: %s", hipGetErrorString(e));
        return LU2MI_ERR_HIP;
    }
    return LU2MI_OK;
}

// address of this device's copy of scratch (called once per device, with the
// device current, when a context is created)
int lu2mi_scratch_address(double **p)
{
    *p = nullptr;
    LU2MI_CHECK(hipGetSymbolAddress((void **)p, HIP_SYMBOL(scratch)));
    return LU2MI_OK;
}

int lu2mi_residual_noalloc_impl(lu2mi_ctx *ctx, int n, const double *dA0, long lda, const double *dx, double *dout)
{
    if (!ctx || !dA0 || !dx || !dout || n <= 0 || lda < n + 1) {
        lu2mi_set_error_msg("residual (no allocation): bad arguments");
        return LU2MI_ERR_ARG;
    }
    if (n > LU2MI_RESID_NMAX) {
        lu2mi_set_error_msg("residual (no allocation): n = %d exceeds the scratch limit %d", n, LU2MI_RESID_NMAX);
        return LU2MI_ERR_ARG;
    }
    if (!ctx->d_scratch) {
        lu2mi_set_error_msg("residual (no allocation): scratch address unknown");
        return LU2MI_ERR_HIP;
    }
    return residual_launch(ctx, n, dA0, lda, dx, ctx->d_scratch, dout);
}
