// lu2mi_internal.h -- shared declarations of the MI355X LU library (not public).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "lu2mi.h"

#define LU2MI_CHECK(call)                                                          \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            lu2mi_set_error_msg("%s:%d %s -> %s", __FILE__, __LINE__, #call,       \
                                hipGetErrorString(e_));                            \
            return LU2MI_ERR_HIP;                                                  \
        }                                                                          \
    } while (0)

#define LU2MI_TRY(expr)                                                            \
    do {                                                                           \
        int rc_ = (expr);                                                          \
        if (rc_ != LU2MI_OK) return rc_;                                           \
    } while (0)

void lu2mi_set_error_msg(const char *fmt, ...);

// ---------------------------------------------------------------------------
// GEMM: C = alpha * A * B + beta * C, all row-major (lu2lib.c:2079 shape).
// Batched over gridDim.y with element strides sA/sB/sC.  When a_lower is set,
// A is lower triangular (columns > row are zero) and each row tile only runs
// the K range it needs.
struct GemmDesc {
    int M = 0, N = 0, K = 0, batch = 1;
    double alpha = 1.0, beta = 0.0;
    const void *A = nullptr; long lda = 0; long sA = 0;
    const void *B = nullptr; long ldb = 0; long sB = 0;
    void *C = nullptr; long ldc = 0; long sC = 0;
    int a_lower = 0;
    // 0: any product; 1: trailing update or U12 product in the direct-to-LDS
    // FP64 tile (gemm2_kernel); 2: trailing update in the register-staged
    // tile.  Both trailing forms are kernel symbols of their own (for
    // profiling).
    int tag = 0;
};
int lu2mi_gemm_f64(const GemmDesc &g, hipStream_t s);
int lu2mi_gemm_f32(const GemmDesc &g, hipStream_t s);

// ---------------------------------------------------------------------------
// Per-context device workspace for the GPU-resident factorization.
struct KernelTimer {
    // events around each launch group of one subsystem, read after sync
    hipEvent_t *ev;      // 2 * cap
    double *flops;       // cap: algorithmic flops of each group
    double *bytes;       // cap: algorithmic bytes of each group
    int cap, used;
};

struct lu2mi_ctx {
    int device;
    hipStream_t stream;
    int num_cu;
    // workspace sized for n_cap rows, nb_cap block
    int n_cap, nb_cap;
    int *d_piv;            // n   global pivot rows
    int *d_orig;           // n   block-start row id of the data held by each row
    double *d_invd;        // n   1/d per row (fp64 path)
    float *d_invd_f;       // n   1/d per row (fp32 path)
    int *d_pairs;          // 4*nb_cap (dst, src) lists
    void *d_linv;          // nb*nb inverse of the unit-lower diagonal block
    void *d_tmp;           // nb*nb scratch for the inverse recursion
    void *d_w;             // nb*(n+1) right-part staging (awork of lu2.c:1263)
    size_t w_bytes;
    void *d_slots;         // leaf kernel exchange slots
    size_t slots_bytes;
    unsigned long long *d_bar;  // back-substitution frontier (solved block count)
    unsigned long long bar_count;
    unsigned long long gbar_count;  // sub-panel grid barrier arrivals issued so far
    int sub_w;             // widest sub-panel factored by one persistent launch (0: off)
    int *d_err;            // device error flag
    double *d_y;           // back-substitution vector scratch
    // options
    int pivmode;
    int leaf_w;
    int timing_on;
    int lookahead;         // factor panel i+1 on stream2 during block i's trailing update
    int trail_kernel;      // trailing-update tile: 1 direct-to-LDS FP64, 0 register-staged
    int dgemm_as_trailing; // lu2mi_dgemm_dev runs its product as a trailing update
    int strict_sync;       // leaf and back substitution publish with release/acquire
    int wide16;            // 16-bit updates in 256 x 256 tiles where they fill the device
    int backsub128;        // back substitution in 128-row blocks (64 otherwise)
    hipStream_t stream2;
    hipEvent_t ev_next, ev_panel;
    KernelTimer tcat[LU2MI_T_COUNT];   // per subsystem (LU2MI_T_*)
    double t_last_ms[8];
    // mixed precision: FP32 working copy, kept per-block factors, refinement vectors
    void *m_af; size_t m_af_bytes;
    void *m_linv; size_t m_linv_bytes;
    void *m_pairs; size_t m_pairs_bytes;
    void *m_vec; size_t m_vec_bytes;
    // 16-bit trailing updates: 0 off, LU2MI_LOWP_FP16 or LU2MI_LOWP_BF16; the
    // scaled 16-bit copies of L21 and U12 and their block exponents
    int lowp;
    void *m_a16; size_t m_a16_bytes;
    void *m_b16; size_t m_b16_bytes;
    void *m_s16; size_t m_s16_bytes;
    // host-memory seams (mygdrdgemm, mydgemm, gdr_dgemm_, ...): device
    // buffers (two row chunks of A and C, all of B), pinned bounce buffers
    // (two for uploads, two for downloads), the two copy streams and the
    // events that chain chunk uploads, products and downloads
    void *h_dev_a; size_t h_dev_a_bytes;
    void *h_dev_b; size_t h_dev_b_bytes;
    void *h_dev_c; size_t h_dev_c_bytes;
    void *h_pin[4];        // [0,1] uploads, [2,3] downloads; h_pin_bytes each
    size_t h_pin_bytes;
    hipStream_t s_up, s_down;
    hipEvent_t ev_up[2], ev_done[2], ev_down[2], ev_bdone;
    double h_stats[4];     // last host-memory product: bytes up, bytes down, ms, chunks
    // this device's copy of the residual scratch array (lu2mi_residual_dev_noalloc)
    double *d_scratch;
};

int lu2mi_ctx_reserve(lu2mi_ctx *ctx, int n, int nb, int elem_bytes);

#define LU2MI_LOWP_FP16 1
#define LU2MI_LOWP_BF16 2
#define LU2MI_LOWP_FP16X2 3   // FP16 hi + lo operand split, three products
// C -= L U with the operands rounded to 16 bits under per-128-block
// power-of-two scales, FP32 accumulation (stream-ordered, no sync)
int lu2mi_update16(lu2mi_ctx *ctx, int fmt, int M, int N, int K, const float *L, long ldl, const float *U,
                   long ldu, float *C, long ldc);
int lu2mi_factor_f64(lu2mi_ctx *ctx, int n, double *dA, long lda, int nb, int pivmode,
                     int *dpiv_out, double *dx);
int lu2mi_factor_f32(lu2mi_ctx *ctx, int n, float *dA, long lda, int nb, int pivmode,
                     int *dpiv_out, float *dx, float *keep_linv, int *keep_pairs);
int lu2mi_factor_f64_keep(lu2mi_ctx *ctx, int n, double *dA, long lda, int nb, int pivmode, int *dpiv,
                          double *dx, double *keep_linv, int *keep_pairs);
int lu2mi_backsub_f32_f64(lu2mi_ctx *ctx, const float *A, long lda, int n, double *x);
int lu2mi_cm_panel_f64(lu2mi_ctx *ctx, double *dcm, long ld, int n, int m, int i, int pivmode, int *hpv);
int lu2mi_backsub_f64_f64(lu2mi_ctx *ctx, const double *A, long lda, int n, double *x);
int lu2mi_dev_buf(void **p, size_t *cur, size_t need);
int lu2mi_scratch_address(double **p);
int lu2mi_residual_noalloc_impl(lu2mi_ctx *ctx, int n, const double *dA0, long lda, const double *dx, double *dout);

// default context used by the lu2-shaped (reference-signature) entry points
lu2mi_ctx *lu2mi_default_ctx(void);
