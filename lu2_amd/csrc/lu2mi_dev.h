// lu2mi_dev.h -- device helpers shared by the kernels of the MI355X LU library:
// the drand48 generator with jump-ahead and the matrix-core (MFMA) traits.
#pragma once
#include <hip/hip_runtime.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// drand48 (lu2.c:70-85 draws it row by row).  x_{k+1} = (a x_k + c) mod 2^48,
// x_0 = (seed << 16) | 0x330E, value = x / 2^48.  Jump-ahead composes the
// affine map, so any draw is reached in O(log k) steps and every element is
// bit-identical to glibc's sequential draws.
namespace lu2mi_rng {
constexpr unsigned long long M48 = (1ULL << 48) - 1;
struct Lcg {
    unsigned long long a, c;
};
__device__ __forceinline__ Lcg lcg_then(Lcg f, Lcg g)  // g after f
{
    return Lcg{(g.a * f.a) & M48, (g.a * f.c + g.c) & M48};
}
__device__ inline Lcg lcg_pow(unsigned long long steps)
{
    Lcg r{1, 0}, b{0x5DEECE66DULL, 0xBULL};
    while (steps) {
        if (steps & 1) r = lcg_then(r, b);
        b = lcg_then(b, b);
        steps >>= 1;
    }
    return r;
}
__host__ __device__ inline unsigned long long seed_state(unsigned long long seed)
{
    return ((seed << 16) | 0x330EULL) & M48;
}
__device__ __forceinline__ double to_unit(unsigned long long x)
{
    return (double)x * (1.0 / 281474976710656.0);
}
}  // namespace lu2mi_rng

// ---------------------------------------------------------------------------
// Matrix-core traits.  FP64: v_mfma_f64_16x16x4f64; FP32:
// v_mfma_f32_16x16x4f32.  In both, lane l supplies A[l%16][k0+l/16] and
// B[k0+l/16][l%16]; the 16x16 result lands in the lane's four accumulator
// registers at row out_row(l, r), column l%16 (layouts checked on gfx950).
// BK is one K stage of the tiled GEMM: 16 bytes per A row in both types.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4_t acc_t;
    typedef double2 vec_t;   // 16 bytes
    static constexpr int BK = 16, EPV = 2;
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef f4_t acc_t;
    typedef float4 vec_t;
    static constexpr int BK = 32, EPV = 4;
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ---------------------------------------------------------------------------
// Tile order of the GEMM grids.  Workgroups go round-robin over the 8 XCDs;
// xcd_remap gives each XCD a contiguous run of tile indices, and tile_coords
// walks GROUP_M-row bands so that the A rows and B columns a run shares stay
// in that XCD's L2.
namespace lu2mi_tiles {
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
}  // namespace lu2mi_tiles
