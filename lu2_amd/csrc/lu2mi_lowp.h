// lu2mi_lowp.h -- the 16-bit operand formats of the trailing update
// (FP16, BF16, FP16 hi + lo) and the wide-tile product's launcher, shared by
// lu2mi_lowp.hip and lu2mi_lowp_wide.hip.
#pragma once
#include "lu2mi_internal.h"
#include "lu2mi_dev.h"

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));

namespace lu2mi_lowp {
constexpr int TB = 128;    // scaling block = output tile edge
constexpr int BKH = 64;    // K per LDS stage (16-bit elements): two 32-deep matrix-core steps
constexpr int PADH = 8;    // LDS row pad (elements): rows 144 bytes apart, conflict-free fragment reads

template <int FMT> struct H16;
template <> struct H16<LU2MI_LOWP_FP16> {
    typedef _Float16 t;
    typedef h8_t v8;
    static constexpr int NP = 1;   // operand planes
    static __device__ __forceinline__ f4_t mfma(v8 a, v8 b, f4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct H16<LU2MI_LOWP_BF16> {
    typedef __bf16 t;
    typedef b8_t v8;
    static constexpr int NP = 1;
    static __device__ __forceinline__ f4_t mfma(v8 a, v8 b, f4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <> struct H16<LU2MI_LOWP_FP16X2> {
    typedef _Float16 t;
    typedef h8_t v8;
    static constexpr int NP = 2;   // hi and lo planes
    static __device__ __forceinline__ f4_t mfma(v8 a, v8 b, f4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
}  // namespace lu2mi_lowp

// C[M x N] -= 2^-(sa+sb) A16 B16t^T in the wide (256 x 256) tile: the scaled
// 16-bit operand copies padded by one 128-row block past the last one, pa/pb
// the plane strides, nsa/nsb the numbers of scaling blocks.  Launched on the
// context's stream; same bits as the 128 x 128 tile.
int lu2mi_gemm16w_launch(lu2mi_ctx *ctx, int fmt, const void *A16, const void *B16t, int M, int N, int K, long pa,
                         long pb, const int *sa, const int *sb, int nsa, int nsb, float *C, long ldc);
