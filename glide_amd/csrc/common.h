// Device helpers shared by the gfx950 kernels: error plumbing, Philox noise,
// packed-record loads (float / bfloat16) and grid interpolation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/glide.h"

namespace glide {

// ---------------------------------------------------------------------------
// Host-side error plumbing (thread-local message, status codes of glide.h)
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

#define GLIDE_HIP(call)                                  \
    do {                                                 \
        hipError_t _e = (call);                          \
        if (_e != hipSuccess) return check_hip(_e, #call); \
    } while (0)

#define GLIDE_REQUIRE(cond, ...)                                       \
    do {                                                               \
        if (!(cond)) return set_error(GLIDE_ERR_INVALID, __VA_ARGS__); \
    } while (0)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the noise layout of the engine:
// key = (seed lo, seed hi), counter = (uid, step, stream, 0).
// ---------------------------------------------------------------------------
constexpr uint32_t kStreamMeander = 0x10000u;
constexpr uint32_t kStreamConvection = 0x20000u;

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

// Uniform in (0, 1): ((x >> 9) + 0.5) * 2^-23, exact in float (24 bits).
__device__ __forceinline__ float u01(uint32_t x) {
    return (static_cast<float>(x >> 9) + 0.5f) * 1.1920928955078125e-07f;
}

// Box-Muller pair from (ua, ub): (r cos(2 pi ub), r sin(2 pi ub)), on the
// hardware log2 / sqrt / sin / cos (the sin and cos take revolutions).
__device__ __forceinline__ void box_muller(float ua, float ub, float& n0, float& n1) {
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(ua));
    n0 = r * __builtin_amdgcn_cosf(ub);
    n1 = r * __builtin_amdgcn_sinf(ub);
}

// ---------------------------------------------------------------------------
// bfloat16 <-> float (round to nearest even)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

__device__ __forceinline__ uint32_t f32_to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// Packed record of 8 values: 32 B (float) or 16 B (bfloat16).
template <bool BF16>
struct Rec8;

template <>
struct Rec8<false> {
    using Store = float4;
    static constexpr int kStride = 2;  // float4 per record
    __device__ __forceinline__ static void load(const void* base, int64_t cell, float (&v)[8]) {
        const float4* p = static_cast<const float4*>(base) + cell * 2;
        const float4 a = p[0];
        const float4 b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
    __device__ __forceinline__ static void load4(const void* base, int64_t cell, float (&v)[4]) {
        const float4 a = static_cast<const float4*>(base)[cell * 2];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    __device__ __forceinline__ static void load2(const void* base, int64_t cell, float (&v)[2]) {
        const float2 a = static_cast<const float2*>(base)[cell * 4];
        v[0] = a.x; v[1] = a.y;
    }
    __device__ __forceinline__ static void store(void* base, int64_t cell, const float (&v)[8]) {
        float4* p = static_cast<float4*>(base) + cell * 2;
        p[0] = make_float4(v[0], v[1], v[2], v[3]);
        p[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
};

template <>
struct Rec8<true> {
    __device__ __forceinline__ static void load(const void* base, int64_t cell, float (&v)[8]) {
        const uint4 a = static_cast<const uint4*>(base)[cell];
        v[0] = bf16_lo(a.x); v[1] = bf16_hi(a.x);
        v[2] = bf16_lo(a.y); v[3] = bf16_hi(a.y);
        v[4] = bf16_lo(a.z); v[5] = bf16_hi(a.z);
        v[6] = bf16_lo(a.w); v[7] = bf16_hi(a.w);
    }
    __device__ __forceinline__ static void load4(const void* base, int64_t cell, float (&v)[4]) {
        const uint2 a = static_cast<const uint2*>(base)[cell * 2];
        v[0] = bf16_lo(a.x); v[1] = bf16_hi(a.x);
        v[2] = bf16_lo(a.y); v[3] = bf16_hi(a.y);
    }
    __device__ __forceinline__ static void load2(const void* base, int64_t cell, float (&v)[2]) {
        const uint32_t a = static_cast<const uint32_t*>(base)[cell * 4];
        v[0] = bf16_lo(a); v[1] = bf16_hi(a);
    }
    __device__ __forceinline__ static void store(void* base, int64_t cell, const float (&v)[8]) {
        uint4 a;
        a.x = f32_to_bf16_bits(v[0]) | (f32_to_bf16_bits(v[1]) << 16);
        a.y = f32_to_bf16_bits(v[2]) | (f32_to_bf16_bits(v[3]) << 16);
        a.z = f32_to_bf16_bits(v[4]) | (f32_to_bf16_bits(v[5]) << 16);
        a.w = f32_to_bf16_bits(v[6]) | (f32_to_bf16_bits(v[7]) << 16);
        static_cast<uint4*>(base)[cell] = a;
    }
};

// Surface record of 12 values (blh sp ustar shf t0 at start, then at end).
template <bool BF16>
__device__ __forceinline__ void load_surf(const void* base, int64_t cell, float (&v)[10]) {
    if constexpr (!BF16) {
        const float4* p = static_cast<const float4*>(base) + cell * 3;
        const float4 a = p[0], b = p[1], c = p[2];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x;
        v[5] = b.y; v[6] = b.z; v[7] = b.w; v[8] = c.x; v[9] = c.y;
    } else {
        const uint2* p = static_cast<const uint2*>(base) + cell * 3;
        const uint2 a = p[0], b = p[1], c = p[2];
        v[0] = bf16_lo(a.x); v[1] = bf16_hi(a.x); v[2] = bf16_lo(a.y); v[3] = bf16_hi(a.y);
        v[4] = bf16_lo(b.x); v[5] = bf16_hi(b.x); v[6] = bf16_lo(b.y); v[7] = bf16_hi(b.y);
        v[8] = bf16_lo(c.x); v[9] = bf16_hi(c.x);
    }
}

// torch.lerp(start, end, weight) with a scalar weight: two-sided form.
__device__ __forceinline__ float torch_lerp(float a, float b, float w) {
    return (fabsf(w) < 0.5f) ? a + w * (b - a) : b - (b - a) * (1.0f - w);
}

// ---------------------------------------------------------------------------
// Grid coordinates. The reference normalises to [-1, 1] and grid_sample
// (align_corners=True) maps back to [0, n-1]; here the fractional index is
// formed directly and clamped to the same range (clamp-to-edge sampling).
// ---------------------------------------------------------------------------
struct HGrid {
    float lon0, inv_dlon, lat0, inv_dlat;  // index = (coord - c0) * inv_span * (n - 1)
    int nx, ny, nz;
};

__device__ __forceinline__ float frac_index(float c, float c0, float scale, int n) {
    float f = (c - c0) * scale;
    return fminf(fmaxf(f, 0.0f), static_cast<float>(n - 1));
}

// Fractional ascending level index of z against levels[0..nz) (reference
// gpu_engine.py:264-292): count = #levels < z, upper = clamp(count, 1, nz-1),
// frac = clamp((z - L_lo) / max(L_hi - L_lo, 1e-6), 0, 1).
__device__ __forceinline__ float level_index(const float* __restrict__ lev, int nz, float z) {
    int cnt = 0;
    int step = 64;
    while (step > nz) step >>= 1;
    for (; step > 0; step >>= 1) {
        const int j = cnt + step;
        if (j <= nz && lev[j - 1] < z) cnt = j;
    }
    const int up = min(max(cnt, 1), nz - 1);
    const float lo = lev[up - 1];
    const float hi = lev[up];
    float fr = (z - lo) * __builtin_amdgcn_rcpf(fmaxf(hi - lo, 1e-6f));
    fr = fminf(fmaxf(fr, 0.0f), 1.0f);
    return static_cast<float>(up - 1) + fr;
}

struct Corner3 {
    int64_t c000, c001, c010, c011, c100, c101, c110, c111;
    float w000, w001, w010, w011, w100, w101, w110, w111;
};

// Trilinear corners at fractional indices (ix, iy, iz) (already clamped).
__device__ __forceinline__ Corner3 corners3(float ix, float iy, float iz, int nx, int ny, int nz) {
    const float fx0 = floorf(ix), fy0 = floorf(iy), fz0 = floorf(iz);
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0), z0 = static_cast<int>(fz0);
    const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
    const float fx = ix - fx0, fy = iy - fy0, fz = iz - fz0;
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    const int64_t sz = static_cast<int64_t>(ny) * nx;
    const int64_t b0 = z0 * sz, b1 = z1 * sz;
    const int64_t r0 = static_cast<int64_t>(y0) * nx, r1 = static_cast<int64_t>(y1) * nx;
    Corner3 c;
    c.c000 = b0 + r0 + x0; c.c001 = b0 + r0 + x1;
    c.c010 = b0 + r1 + x0; c.c011 = b0 + r1 + x1;
    c.c100 = b1 + r0 + x0; c.c101 = b1 + r0 + x1;
    c.c110 = b1 + r1 + x0; c.c111 = b1 + r1 + x1;
    c.w000 = gx * gy * gz; c.w001 = fx * gy * gz;
    c.w010 = gx * fy * gz; c.w011 = fx * fy * gz;
    c.w100 = gx * gy * fz; c.w101 = fx * gy * fz;
    c.w110 = gx * fy * fz; c.w111 = fx * fy * fz;
    return c;
}

template <bool BF16, int K>
__device__ __forceinline__ void accum8(const void* base, int64_t cell, float w, float (&acc)[K]) {
    float v[8];
    Rec8<BF16>::load(base, cell, v);
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += w * v[k];
}

// Trilinear sample of the first K (<= 8) channels of a record field.
template <bool BF16, int K>
__device__ __forceinline__ void sample_rec(const void* base, const Corner3& c, float (&acc)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;
    accum8<BF16, K>(base, c.c000, c.w000, acc);
    accum8<BF16, K>(base, c.c001, c.w001, acc);
    accum8<BF16, K>(base, c.c010, c.w010, acc);
    accum8<BF16, K>(base, c.c011, c.w011, acc);
    accum8<BF16, K>(base, c.c100, c.w100, acc);
    accum8<BF16, K>(base, c.c101, c.w101, acc);
    accum8<BF16, K>(base, c.c110, c.w110, acc);
    accum8<BF16, K>(base, c.c111, c.w111, acc);
}

// Trilinear sample of channels 0..1 only (the free-troposphere pair).
template <bool BF16>
__device__ __forceinline__ void sample_rec2(const void* base, const Corner3& c, float& a0, float& a1) {
    float v[2];
    a0 = 0.0f;
    a1 = 0.0f;
#define GLIDE_ACC2(cell, w)                  \
    Rec8<BF16>::load2(base, (cell), v);      \
    a0 += (w) * v[0];                        \
    a1 += (w) * v[1];
    GLIDE_ACC2(c.c000, c.w000) GLIDE_ACC2(c.c001, c.w001) GLIDE_ACC2(c.c010, c.w010)
    GLIDE_ACC2(c.c011, c.w011) GLIDE_ACC2(c.c100, c.w100) GLIDE_ACC2(c.c101, c.w101)
    GLIDE_ACC2(c.c110, c.w110) GLIDE_ACC2(c.c111, c.w111)
#undef GLIDE_ACC2
}

}  // namespace glide
