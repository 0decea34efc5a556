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
// Layout in HBM: lu2's row-major augmented matrix, lda >= n+1 (RDIM = n+16
// in lu2.c:28), b in column n.  Row-major rows are contiguous, so row swaps
// and scalings are coalesced along rows.
#include "lu2mi_internal.h"
#include <limits.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <vector>

// ---------------------------------------------------------------------------
// drand48 on the device (lu2.c:70-85).  x_{k+1} = (a x_k + c) mod 2^48,
// x_0 = (seed << 16) | 0x330E, value = x / 2^48.  Jump-ahead composes the
// affine map; every element is bit-identical to glibc's sequential draws.
namespace {
constexpr unsigned long long M48 = (1ULL << 48) - 1;
struct Lcg {
    unsigned long long a, c;
};
__device__ __forceinline__ Lcg lcg_then(Lcg f, Lcg g)  // g after f
{
    return Lcg{(g.a * f.a) & M48, (g.a * f.c + g.c) & M48};
}
__device__ Lcg lcg_pow(unsigned long long steps)
{
    Lcg r{1, 0}, b{0x5DEECE66DULL, 0xBULL};
    while (steps) {
        if (steps & 1) r = lcg_then(r, b);
        b = lcg_then(b, b);
        steps >>= 1;
    }
    return r;
}
}  // namespace

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

// ---------------------------------------------------------------------------
// Leaf factorization: persistent kernel over rows [c, n), W columns; all
// workgroups are co-resident and meet at a grid barrier once per column.
namespace {
constexpr int SLOT_WORDS = 40;  // 3 header words + up to 32 row values, padded
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

// Monotonic-counter grid barrier.  Bounded spin: on time-out the device error
// flag is raised and every later barrier falls through, so the kernel always
// drains; the host turns the flag into LU2MI_ERR_DEVICE.
__device__ __forceinline__ void grid_barrier(unsigned long long *bar, unsigned long long target,
                                             int *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1ULL, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (((++spins) & 1023) == 0) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (spins > (1UL << 25)) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

__device__ __forceinline__ bool better(double v, int r, double bv, int br)
{
    return (v > bv) || (v == bv && r < br);
}
}  // namespace

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
    unsigned long long *bar;
    unsigned long long bar_base;
    int *err;
};

template <typename T, int W, int RPT>
__global__ __launch_bounds__(NTL) void leaf_kernel(LeafArgs<T> a)
{
    __shared__ T stage[NTL * (W + 1)];
    // every workgroup's slot (header + row) plus the current-row slot
    __shared__ unsigned long long xch[(LEAF_GMAX + 1) * (W + 3)];
    __shared__ double red_v[NTL / 64];
    __shared__ int red_r[NTL / 64], red_g[NTL / 64];
    __shared__ int s_win[8];
    __shared__ int pivs[W];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int G = a.G;
    const long TT = (long)G * NTL;
    const int gt = blockIdx.x * NTL + t;
    const int n = a.n, c = a.c;
    T *__restrict__ A = a.A;

    T row[RPT][W];
    int lorig[RPT], borig[RPT], rq[RPT];

    // coalesced load of the rows through LDS, one 256-row layer at a time
#pragma clang loop unroll(full)
    for (int q = 0; q < RPT; q++) {
        const long base = (long)c + (long)blockIdx.x * NTL + (long)q * TT;
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

    unsigned long long *const slots = a.slots;
#pragma clang loop unroll(full)
    for (int jl = 0; jl < W; jl++) {
        const int j = c + jl;
        const int par = jl & 1;
        const bool odd_rule = a.compat && (jl & 1);
        // ---- local candidate: first max |x| over owned rows r >= j
        double bv = -1.0;
        int br = INT_MAX;
#pragma clang loop unroll(full)
        for (int q = 0; q < RPT; q++) {
            const int r = rq[q];
            if (r >= j && r < n) {
                const T xa = row[q][jl];
                const T xb = row[q][jl > 0 ? jl - 1 : 0];
                const T x = odd_rule ? xb : xa;
                const double v = fabs((double)x);
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
        if (t == 0) {
            double v = red_v[0];
            int r = red_r[0];
            for (int w2 = 1; w2 < NTL / 64; w2++)
                if (better(red_v[w2], red_r[w2], v, r)) { v = red_v[w2]; r = red_r[w2]; }
            s_win[0] = r;
            red_v[0] = v;
        }
        __syncthreads();
        {
            // publish: the owner of the workgroup's candidate row writes the
            // slot (value, row, origins, W values); workgroup 0's thread jl
            // writes the current row j.  Writers fence before the barrier.
            const int wr = s_win[0];
            const double wv = red_v[0];
            unsigned long long *sl = slots + ((long)par * (G + 1) + blockIdx.x) * SLOT_WORDS;
            bool wrote = false;
#pragma clang loop unroll(full)
            for (int q = 0; q < RPT; q++)
                if (rq[q] == wr && wr < n) {
                    st_relaxed(sl + 0, dbits(wv));
                    st_relaxed(sl + 1, (unsigned long long)(unsigned)wr);
                    st_relaxed(sl + 2, (unsigned long long)(unsigned)lorig[q] |
                                           ((unsigned long long)(unsigned)borig[q] << 32));
#pragma clang loop unroll(full)
                    for (int k = 0; k < W; k++) st_relaxed(sl + 3 + k, dbits((double)row[q][k]));
                    wrote = true;
                }
            if (wr >= n && t == 0) {
                st_relaxed(sl + 0, dbits(-1.0));
                st_relaxed(sl + 1, (unsigned long long)(unsigned)INT_MAX);
                wrote = true;
            }
            if (blockIdx.x == 0 && t == jl) {
                unsigned long long *cs = slots + ((long)par * (G + 1) + G) * SLOT_WORDS;
                st_relaxed(cs + 2, (unsigned long long)(unsigned)lorig[0] |
                                       ((unsigned long long)(unsigned)borig[0] << 32));
#pragma clang loop unroll(full)
                for (int k = 0; k < W; k++) st_relaxed(cs + 3 + k, dbits((double)row[0][k]));
                wrote = true;
            }
            if (wrote) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        }
        grid_barrier(a.bar, a.bar_base + (unsigned long long)(jl + 1) * G, a.err);

        // ---- one round trip: every slot (headers and rows) into LDS
        constexpr int SW = W + 3;
        for (int e = t; e < (G + 1) * SW; e += NTL) {
            const int g2 = e / SW, k = e - g2 * SW;
            xch[e] = ld_relaxed(slots + ((long)par * (G + 1) + g2) * SLOT_WORDS + k);
        }
        __syncthreads();
        // ---- global winner over the G candidates (same in every workgroup)
        bv = -1.0;
        br = INT_MAX;
        int bg = -1;
        for (int g2 = t; g2 < G; g2 += NTL) {
            const double v = bitsd(xch[g2 * SW + 0]);
            const int r = (int)(unsigned)xch[g2 * SW + 1];
            if (better(v, r, bv, br)) { bv = v; br = r; bg = g2; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int orr = __shfl_xor(br, off);
            const int og = __shfl_xor(bg, off);
            if (better(ov, orr, bv, br)) { bv = ov; br = orr; bg = og; }
        }
        if (lane == 0) { red_v[wave] = bv; red_r[wave] = br; red_g[wave] = bg; }
        __syncthreads();
        if (t == 0) {
            double v = red_v[0];
            int r = red_r[0], g2 = red_g[0];
            for (int w2 = 1; w2 < NTL / 64; w2++)
                if (better(red_v[w2], red_r[w2], v, r)) { v = red_v[w2]; r = red_r[w2]; g2 = red_g[w2]; }
            // no comparable candidate (every |x| is NaN): pivot on the current
            // row, as BLAS idamax returns the first index for NaN input
            const bool none = (r < j || r >= n || g2 < 0);
            s_win[1] = none ? j : r;
            s_win[2] = none ? G : g2;
        }
        __syncthreads();
        const int p = s_win[1];
        const unsigned long long *pslot = xch + s_win[2] * SW;
        const unsigned long long *cslot = xch + G * SW;
        const int p_lorig = (int)(unsigned)(pslot[2] & 0xffffffffULL), p_borig = (int)(unsigned)(pslot[2] >> 32);
        const int c_lorig = (int)(unsigned)(cslot[2] & 0xffffffffULL), c_borig = (int)(unsigned)(cslot[2] >> 32);
#define PIVROW(k) ((T)bitsd(pslot[3 + (k)]))
#define CURROW(k) ((T)bitsd(cslot[3 + (k)]))
        const T d = PIVROW(jl);
        const T s = T(1) / d;
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
            if (r > j && r < n) {
                const T l = row[q][jl];
#pragma clang loop unroll(full)
                for (int k = jl + 1; k < W; k++) row[q][k] -= l * (PIVROW(k) * s);
            }
        }
        if (t == 0) {
            pivs[jl] = p;
            if (blockIdx.x == 0) {
                a.piv[j] = p;
                a.invd[j] = s;
            }
        }
    }
#undef PIVROW
#undef CURROW
    __syncthreads();

    // ---- write back rows (coalesced through LDS), origins and row pairs
#pragma clang loop unroll(full)
    for (int q = 0; q < RPT; q++) {
        const long base = (long)c + (long)blockIdx.x * NTL + (long)q * TT;
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
    if (blockIdx.x == 0) {
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
__global__ __launch_bounds__(256) void rowperm_kernel(PermArgs<T> a)
{
    extern __shared__ unsigned char smem_raw[];
    T *buf = (T *)smem_raw;
    const int cs = a.c0 + blockIdx.x * a.sw;
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

// pairs of the block's row permutation: (i+k, orig[i+k]) and (piv[i+k], orig[piv[i+k]])
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
// Back substitution on the unit-upper U (lu2.c:110-116), blocks of BSB rows
// from the bottom.  Step b: workgroup 0 first applies the solved block b+1 to
// the rows of block b and then solves block b in LDS; the other workgroups
// apply block b+1 to the rows above block b.
constexpr int BSB = 64;

template <typename T>
__global__ __launch_bounds__(256) void backsub_step_kernel(const T *A, long lda, int n, int b, T *x)
{
    __shared__ T U[BSB][BSB + 1];
    __shared__ T xs[BSB];
    const int r0 = b * BSB;
    const int r1 = min(n, r0 + BSB);
    const int nrb = r1 - r0;
    const int s0 = r1, s1 = min(n, r1 + BSB);   // previously solved block
    const int t = threadIdx.x;
    const int rows_per_wg = 64;
    // rows handled by this workgroup
    int rlo, rhi;
    if (blockIdx.x == 0) { rlo = r0; rhi = r1; }
    else { rlo = (blockIdx.x - 1) * rows_per_wg; rhi = min(r0, rlo + rows_per_wg); }
    if (s1 > s0) {
        if (t < s1 - s0) xs[t] = x[s0 + t];
        __syncthreads();
        // 4 threads per row, 16 columns each
        const int rr = rlo + (t >> 2), part = t & 3;
        T acc = T(0);
        if (rr < rhi) {
            const T *u = A + (long)rr * lda + s0;
            for (int k = part * 16; k < min(s1 - s0, part * 16 + 16); k++) acc += u[k] * xs[k];
        }
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (rr < rhi && part == 0) x[rr] -= acc;
    }
    if (blockIdx.x != 0) return;
    __syncthreads();
    // solve block b: x[r] -= sum_{k>r, k<r1} U[r][k] x[k], bottom up
    for (int e = t; e < BSB * BSB; e += 256) {
        const int rr = e / BSB, kk = e % BSB;
        U[rr][kk] = (rr < nrb && kk < nrb) ? A[(long)(r0 + rr) * lda + r0 + kk] : T(0);
    }
    if (t < BSB) xs[t] = (t < nrb) ? x[r0 + t] : T(0);
    __syncthreads();
    if (t < 64) {
        // one wave; lane t keeps x[r0+t] in a register, bottom-up
        T xt = xs[t];
        for (int rr = nrb - 1; rr >= 0; rr--) {
            T v = (t > rr && t < nrb) ? U[rr][t] * xt : T(0);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (t == rr) xt -= v;
        }
        if (t < nrb) x[r0 + t] = xt;
    }
}

// ---------------------------------------------------------------------------
// FP64 residual: per row |(A x)_i - b_i| and sum_j |a_ij|; per column sum |a_ij|.
__global__ __launch_bounds__(256) void resid_rows_kernel(const double *A, long lda, int n, const double *x,
                                                         double *rerr, double *rabs)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long r = (long)blockIdx.x * 4 + wave;
    if (r >= n) return;
    const double *a = A + r * lda;
    double s = 0.0, sa = 0.0;
    for (int c = lane; c < n; c += 64) {
        s = fma(a[c], x[c], s);
        sa += fabs(a[c]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        sa += __shfl_xor(sa, off);
    }
    if (lane == 0) {
        rerr[r] = fabs(s - a[n]);
        rabs[r] = sa;
    }
}

__global__ __launch_bounds__(256) void resid_cols_kernel(const double *A, long lda, int n, double *cabs)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    double s = 0.0;
    for (long r = 0; r < n; r++) s += fabs(A[r * lda + col]);
    cabs[col] = s;
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
        v[0] = fmax(v[0], rerr[r]);
        v[1] = fmax(v[1], rabs[r]);
        v[2] = fmax(v[2], cabs[r]);
        v[3] = fmax(v[3], fabs(x[r]));
        v[4] += fabs(x[r]);
        v[5] = fmax(v[5], fabs(A[(long)r * lda + n]));
    }
    for (int k = 0; k < 6; k++) sh[k][t] = v[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int k = 0; k < 6; k++) sh[k][t] = (k == 4) ? sh[k][t] + sh[k][t + st] : fmax(sh[k][t], sh[k][t + st]);
        __syncthreads();
    }
    if (t < 6) out[t] = sh[t][0];
}

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
    LU2MI_CHECK(hipMalloc(&ctx->d_pairs, sizeof(int) * 4 * (size_t)(nbcap > 64 ? nbcap : 64)));
    LU2MI_CHECK(hipMalloc(&ctx->d_linv, sizeof(double) * (size_t)nbcap * nbcap));
    LU2MI_CHECK(hipMalloc(&ctx->d_tmp, sizeof(double) * (size_t)nbcap * nbcap));
    LU2MI_CHECK(hipMalloc(&ctx->d_y, sizeof(double) * 4 * (size_t)ncap + 64));
    size_t wneed = sizeof(double) * (size_t)nbcap * (size_t)(ncap + 1);
    LU2MI_TRY(lu2mi_dev_buf(&ctx->d_w, &ctx->w_bytes, wneed));
    size_t sneed = sizeof(unsigned long long) * 2 * (size_t)(1024 + 1) * SLOT_WORDS;
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

template <typename T>
struct Factor {
    lu2mi_ctx *ctx;
    T *A;
    long lda;
    int n, nb, compat, W;
    T *linv, *tmp, *wbuf;
    int *pairs;

    int leaf(int i, int c)
    {
        const int R = n - c;
        // rows per thread: keep the grid at <= 128 workgroups where possible
        int rpt = (R + NTL * 128 - 1) / (NTL * 128);
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
        la.bar = ctx->d_bar; la.bar_base = ctx->bar_count; la.err = ctx->d_err;
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
        ctx->bar_count += (unsigned long long)W * G;
        // apply the leaf's swaps to the other panel columns; L part (columns
        // left of the leaf) of the leaf's diagonal rows scaled by 1/d
        return rowperm(pairs, pairs + 2 * W, 2 * W, i, i + nb, c, c + W, c, c, c + W, nullptr, 0, 0, 0);
    }

    int rowperm(const int *dst, const int *src, int np, int c0, int c1, int k0, int k1, int sce, int d0,
                int d1, T *Wb, long ldw, int w0, int w1)
    {
        if (c1 <= c0) return LU2MI_OK;
        int sw = (int)(16384 / (np * (int)sizeof(T)));
        if (sw > 128) sw = 128;
        if (sw < 8) sw = 8;
        PermArgs<T> p;
        p.A = A; p.lda = lda; p.dst = dst; p.src = src; p.np = np;
        p.c0 = c0; p.c1 = c1; p.k0 = k0; p.k1 = k1; p.sce = sce; p.d0 = d0; p.d1 = d1;
        p.invd = Prec<T>::invd(ctx); p.W = Wb; p.ldw = ldw; p.w0 = w0; p.w1 = w1; p.sw = sw;
        const int strips = (c1 - c0 + sw - 1) / sw;
        const size_t shm = (size_t)np * sw * sizeof(T);
        hipLaunchKernelGGL(rowperm_kernel<T>, dim3(strips), dim3(256), shm, ctx->stream, p);
        LU2MI_CHECK(hipGetLastError());
        return LU2MI_OK;
    }

    int panel(int i, int c, int w)
    {
        if (w <= W) return leaf(i, c);
        const int h = w / 2;
        LU2MI_TRY(panel(i, c, h));
        // right half of this node: rows [c, c+h) scaled by 1/d, then
        // X = inv(D^-1 L11) * scaled rows, then the rows below
        {
            const long e = (long)h * h;
            hipLaunchKernelGGL(scale_copy_kernel<T>, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, ctx->stream,
                               A, lda, c, c + h, h, h, Prec<T>::invd(ctx), wbuf);
            LU2MI_CHECK(hipGetLastError());
            const long off = (long)(c - i);
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, 1.0, linv + off * nb + off, nb, wbuf, h, 0.0,
                                   A + (long)c * lda + c + h, lda, 1));
            LU2MI_TRY(gemm_call<T>(ctx, n - c - h, h, h, -1.0, A + (long)(c + h) * lda + c, lda,
                                   A + (long)c * lda + c + h, lda, 1.0, A + (long)(c + h) * lda + c + h, lda, 0));
        }
        LU2MI_TRY(panel(i, c + h, h));
        // off-diagonal block of the inverse: X21 = -inv22 * L21 * inv11
        {
            const long off = (long)(c - i);
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, 1.0, A + (long)(c + h) * lda + c, lda, linv + off * nb + off, nb,
                                   0.0, tmp, h, 0));
            LU2MI_TRY(gemm_call<T>(ctx, h, h, h, -1.0, linv + (off + h) * nb + off + h, nb, tmp, h, 0.0,
                                   linv + (off + h) * nb + off, nb, 1));
        }
        return LU2MI_OK;
    }

    int run(int *dpiv_out, T *dx)
    {
        for (int i = 0; i < n; i += nb) {
            hipLaunchKernelGGL(fill_kernel<T>, dim3(256), dim3(256), 0, ctx->stream, linv, (long)nb * nb, T(0));
            LU2MI_CHECK(hipGetLastError());
            LU2MI_TRY(panel(i, i, nb));
            const int ncols = n + 1 - (i + nb);
            // block-level row permutation over [i+nb, n], 1/d scaling of the
            // block rows, block rows staged in w (lu2.c:502-509 awork copy)
            hipLaunchKernelGGL(block_pairs_kernel, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream,
                               ctx->d_piv, ctx->d_orig, i, nb, pairs, pairs + 2 * nb);
            LU2MI_CHECK(hipGetLastError());
            const long ldw = (ncols + 1) & ~1L;
            LU2MI_TRY(rowperm(pairs, pairs + 2 * nb, 2 * nb, i + nb, n + 1, 0, 0, n + 1, i, i + nb, wbuf, ldw, i,
                              i + nb));
            // U12 = inv(D^-1 L11) * staged rows   (lu2.c:511-512)
            LU2MI_TRY(gemm_call<T>(ctx, nb, ncols, nb, 1.0, linv, nb, wbuf, ldw, 0.0, A + (long)i * lda + i + nb,
                                   lda, 1));
            // A22 -= L21 * U12   (lu2.c:580)
            const int M = n - i - nb;
            if (M > 0) {
                KernelTimer &kt = ctx->tgemm;
                const bool tm = ctx->timing_on && kt.used < kt.cap;
                if (tm) LU2MI_CHECK(hipEventRecord(kt.ev[2 * kt.used], ctx->stream));
                LU2MI_TRY(gemm_call<T>(ctx, M, ncols, nb, -1.0, A + (long)(i + nb) * lda + i, lda,
                                       A + (long)i * lda + i + nb, lda, 1.0, A + (long)(i + nb) * lda + i + nb,
                                       lda, 0, 1));
                if (tm) {
                    LU2MI_CHECK(hipEventRecord(kt.ev[2 * kt.used + 1], ctx->stream));
                    kt.flops[kt.used] = 2.0 * M * (double)ncols * nb;
                    kt.used++;
                }
            }
        }
        // x = column n, then unit-upper back substitution
        hipLaunchKernelGGL(copy_col_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A, lda, n, n, dx);
        LU2MI_CHECK(hipGetLastError());
        const int nblk = (n + BSB - 1) / BSB;
        for (int b = nblk - 1; b >= 0; b--) {
            const int above = b * BSB;
            const int grid = 1 + (above + 63) / 64;
            hipLaunchKernelGGL(backsub_step_kernel<T>, dim3(grid), dim3(256), 0, ctx->stream, A, lda, n, b, dx);
            LU2MI_CHECK(hipGetLastError());
        }
        if (dpiv_out && dpiv_out != ctx->d_piv)
            LU2MI_CHECK(hipMemcpyAsync(dpiv_out, ctx->d_piv, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
        return LU2MI_OK;
    }
};

template <typename T>
static int factor_impl(lu2mi_ctx *ctx, int n, T *dA, long lda, int nb, int pivmode, int *dpiv_out, T *dx)
{
    if (!ctx || !dA || n <= 0 || lda < n + 1 || nb < 32 || (nb & (nb - 1)) || n % nb) {
        lu2mi_set_error_msg("lumcolumn: need nb power of two >= 32, n %% nb == 0, lda >= n+1 "
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
    ctx->tgemm.used = 0;
    LU2MI_CHECK(hipMemsetAsync(ctx->d_err, 0, sizeof(int), ctx->stream));
    LU2MI_CHECK(hipMemsetAsync(ctx->d_bar, 0, sizeof(unsigned long long), ctx->stream));
    ctx->bar_count = 0;
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
int lu2mi_factor_f32(lu2mi_ctx *ctx, int n, float *dA, long lda, int nb, int pivmode, int *dpiv, float *dx)
{
    return factor_impl<float>(ctx, n, dA, lda, nb, pivmode, dpiv, dx);
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

int lu2mi_residual_impl(lu2mi_ctx *ctx, int n, const double *dA0, long lda, const double *dx, double *out)
{
    if (!ctx || !dA0 || !dx || !out || n <= 0 || lda < n + 1) {
        lu2mi_set_error_msg("residual: bad arguments");
        return LU2MI_ERR_ARG;
    }
    double *buf = nullptr;
    LU2MI_CHECK(hipMalloc(&buf, sizeof(double) * (3 * (size_t)n + 8)));
    double *rerr = buf, *rabs = buf + n, *cabs = buf + 2 * (size_t)n, *dout = buf + 3 * (size_t)n;
    hipLaunchKernelGGL(resid_rows_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, dA0, lda, n, dx, rerr, rabs);
    hipLaunchKernelGGL(resid_cols_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dA0, lda, n, cabs);
    hipLaunchKernelGGL(resid_final_kernel, dim3(1), dim3(256), 0, ctx->stream, dA0, lda, n, dx, rerr, rabs, cabs,
                       dout);
    double h[6];
    hipError_t e = hipMemcpyAsync(h, dout, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) {
        lu2mi_set_error_msg("residual: %s", hipGetErrorString(e));
        return LU2MI_ERR_HIP;
    }
    const double eps = 1.1102230246251565e-16, eps_lu2 = 2e-16;
    out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = h[3]; out[4] = h[4]; out[5] = h[5];
    out[6] = h[0] / (eps * (h[1] * h[3] + h[5]) * n);
    out[7] = h[0] / (eps_lu2 * h[2] * n);
    out[8] = h[0] / (eps_lu2 * h[2] * h[4]);
    out[9] = h[0] / (eps_lu2 * h[1] * h[3] * n);
    return LU2MI_OK;
}
