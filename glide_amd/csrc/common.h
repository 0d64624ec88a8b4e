// Device helpers shared by the gfx950 kernels: error plumbing, Philox noise,
// packed-record loads (float / bfloat16) and grid interpolation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/glide.h"

namespace glide {

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8), and each XCD has its own L2. This bijection of [0, nb) gives each
// XCD a contiguous eighth of the logical blocks, so neighbouring blocks that
// touch the same lines (particles binned by met cell and their met records,
// a permutation's nearby sources) share one L2 instead of eight.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    constexpr uint32_t kXcd = 8;
    const uint32_t xcd = b % kXcd, idx = b / kXcd;
    const uint32_t q = nb / kXcd, r = nb % kXcd;
    return xcd * q + (xcd < r ? xcd : r) + idx;
}

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
        // one 32 x 32 -> 64-bit product each (v_mad_u64_u32)
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t lo0 = static_cast<uint32_t>(p0), hi0 = static_cast<uint32_t>(p0 >> 32);
        const uint32_t lo1 = static_cast<uint32_t>(p1), hi1 = static_cast<uint32_t>(p1 >> 32);
        // three-input XORs: one v_bitop3 each (truth table 0x96 = a ^ b ^ c)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

// Uniform in (0, 1): ((x >> 9) + 0.5) * 2^-23, exact in float (24 bits).
// Formed as the float 1 + (x >> 9) * 2^-23 (the 23 bits as a mantissa) plus
// -(1 - 2^-24): the exact sum has 24 significant bits, so the one rounding
// returns it unchanged.
__device__ __forceinline__ float u01(uint32_t x) {
    return __uint_as_float((x >> 9) | 0x3f800000u) + (-0.99999994039535522f);
}

// Box-Muller pair from (ua, ub): (r cos(2 pi ub), r sin(2 pi ub)), on the
// hardware log2 / sqrt / sin / cos (the sin and cos take revolutions).
__device__ __forceinline__ void box_muller(float ua, float ub, float& n0, float& n1) {
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(ua));
    n0 = r * __builtin_amdgcn_cosf(ub);
    n1 = r * __builtin_amdgcn_sinf(ub);
}

// Element access at a 32-bit byte offset from a uniform base pointer (arrays
// below 4 GiB): keeps per-lane addresses 32-bit.
template <typename T>
__device__ __forceinline__ T ldo(const T* p, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T* p, uint32_t off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off) = v;
}

// Active-mask test of main.py:1396-1401 with a per-particle backward length
// when lo_off is given: r_off >= cursor and r_off - length <= cursor, alive.
// With GLIDE_F_RELEASE_TERMINUS, lo_off is the release terminus te and the
// test is r_off >= cursor > te (the release's own schedule).
__device__ __forceinline__ bool in_window(const GlideParticles& s, uint32_t i1, float cursor, float length,
                                          int32_t flags) {
    const uint32_t i4 = i1 * 4u;
    const float ro = ldo(s.r_off, i4);
    if (flags & GLIDE_F_RELEASE_TERMINUS) {
        return (ro >= cursor) && (cursor > ldo(s.lo_off, i4)) && (ldo(s.alive, i1) != 0);
    }
    const float lo = s.lo_off ? ldo(s.lo_off, i4) : ro - length;
    return (ro >= cursor) && (lo <= cursor) && (ldo(s.alive, i1) != 0);
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
    static constexpr uint32_t kBytes = 32;
    // Loads at a 32-bit byte offset from the (uniform) base: the scalar-base,
    // vector-offset addressing form, no 64-bit address arithmetic per lane.
    __device__ __forceinline__ static void load_at(const void* base, uint32_t off, float (&v)[8]) {
        const float4* p = reinterpret_cast<const float4*>(static_cast<const char*>(base) + off);
        const float4 a = p[0];
        const float4 b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
    // The first 6 values only (the records' last two slots are padding).
    __device__ __forceinline__ static void load6_at(const void* base, uint32_t off, float (&v)[6]) {
        const char* q = static_cast<const char*>(base) + off;
        const float4 a = *reinterpret_cast<const float4*>(q);
        const float2 b = *reinterpret_cast<const float2*>(q + 16);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y;
    }
    __device__ __forceinline__ static void load2_at(const void* base, uint32_t off, float (&v)[2]) {
        const float2 a = *reinterpret_cast<const float2*>(static_cast<const char*>(base) + off);
        v[0] = a.x; v[1] = a.y;
    }
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
    static constexpr uint32_t kBytes = 16;
    __device__ __forceinline__ static void load_at(const void* base, uint32_t off, float (&v)[8]) {
        const uint4 a = *reinterpret_cast<const uint4*>(static_cast<const char*>(base) + off);
        v[0] = bf16_lo(a.x); v[1] = bf16_hi(a.x);
        v[2] = bf16_lo(a.y); v[3] = bf16_hi(a.y);
        v[4] = bf16_lo(a.z); v[5] = bf16_hi(a.z);
        v[6] = bf16_lo(a.w); v[7] = bf16_hi(a.w);
    }
    __device__ __forceinline__ static void load6_at(const void* base, uint32_t off, float (&v)[6]) {
        const uint3 a = *reinterpret_cast<const uint3*>(static_cast<const char*>(base) + off);
        v[0] = bf16_lo(a.x); v[1] = bf16_hi(a.x);
        v[2] = bf16_lo(a.y); v[3] = bf16_hi(a.y);
        v[4] = bf16_lo(a.z); v[5] = bf16_hi(a.z);
    }
    __device__ __forceinline__ static void load2_at(const void* base, uint32_t off, float (&v)[2]) {
        const uint32_t a = *reinterpret_cast<const uint32_t*>(static_cast<const char*>(base) + off);
        v[0] = bf16_lo(a); v[1] = bf16_hi(a);
    }
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


// Surface record at a byte offset: 12 floats (48 B) in both met layouts. The
// surface scalars set the stability regime (air density from surface
// pressure, Obukhov length, w*), so they are not rounded to bfloat16; only the
// trilinear wind and support records are.
__device__ __forceinline__ void load_surf_at(const void* base, uint32_t off, float (&v)[10]) {
    const float4* p = reinterpret_cast<const float4*>(static_cast<const char*>(base) + off);
    const float4 a = p[0], c1 = p[1], c2 = p[2];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = c1.x;
    v[5] = c1.y; v[6] = c1.z; v[7] = c1.w; v[8] = c2.x; v[9] = c2.y;
}
constexpr uint32_t kSurfBytes = 48u;

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

// The same for a position held as grid point c plus an offset (GlideParticles.cell).
__device__ __forceinline__ float cell_index(int c, float off, float scale, int n) {
    const float f = static_cast<float>(c) + off * scale;
    return fminf(fmaxf(f, 0.0f), static_cast<float>(n - 1));
}

// Level ladders in LDS hold kLadder entries: levels[0..nz), then +inf, so a
// branchless binary search may read past nz (a padded entry is never < z).
constexpr int kLadder = 128;

// First step of the level search: the smallest power of two t with
// 2t - 1 >= nz - 1 (the steps t, t/2, .., 1 reach every count that the clamp
// to nz - 1 distinguishes; the last entry read, 2t - 2, is below kLadder for
// nz <= kLadder).
__host__ __device__ inline int ladder_top(int nz) {
    int t = 1;
    while (2 * t - 1 < nz - 1) t *= 2;
    return t;
}

// Copies levels[0..nz) into the LDS ladder, +inf after; every thread of the
// block calls it, a barrier must follow.
__device__ __forceinline__ void fill_ladder(float* lev, const float* levels, int nz) {
    for (int k = threadIdx.x; k < kLadder; k += blockDim.x) lev[k] = k < nz ? levels[k] : __builtin_inff();
}

// Fractional ascending level index of z against the ladder (reference
// gpu_engine.py:264-292): count = #levels < z, upper = clamp(count, 1, nz-1),
// frac = clamp((z - L_lo) / max(L_hi - L_lo, 1e-6), 0, 1). The count is a
// branchless binary search from step top (ladder_top(nz)); counts at or above
// nz - 1 give the same upper level.
__device__ __forceinline__ float level_index(const float* __restrict__ lev, int nz, int top, float z) {
    int cnt = 0;
    for (int step = top; step > 0; step >>= 1) cnt = (lev[cnt + step - 1] < z) ? cnt + step : cnt;
    const int up = min(max(cnt, 1), nz - 1);
    const float lo = lev[up - 1];
    const float hi = lev[up];
    float fr = (z - lo) * __builtin_amdgcn_rcpf(fmaxf(hi - lo, 1e-6f));
    fr = fminf(fmaxf(fr, 0.0f), 1.0f);
    return static_cast<float>(up - 1) + fr;
}

// level_index for a height near one already looked up: u is that lookup's
// upper level (1 <= u <= nz - 1) and becomes this one's. One step up or down
// from u is tried first (two LDS reads), verified to bracket z as the count
// of levels below z would (the clamped end intervals included), and the
// binary search runs only when it does not. The result equals level_index's
// for non-decreasing levels.
__device__ __forceinline__ float level_index_near(const float* __restrict__ lev, int nz, int top, float z, int& u) {
    const int u0 = u;
    int up = u0 + (lev[u0] < z ? 1 : 0) - (lev[u0 - 1] >= z ? 1 : 0);
    up = min(max(up, 1), nz - 1);
    float lo = lev[up - 1];
    float hi = lev[up];
    if (!((up == 1 || lo < z) && (up == nz - 1 || z <= hi))) {
        int cnt = 0;
        for (int step = top; step > 0; step >>= 1) cnt = (lev[cnt + step - 1] < z) ? cnt + step : cnt;
        up = min(max(cnt, 1), nz - 1);
        lo = lev[up - 1];
        hi = lev[up];
    }
    u = up;
    float fr = (z - lo) * __builtin_amdgcn_rcpf(fmaxf(hi - lo, 1e-6f));
    fr = fminf(fmaxf(fr, 0.0f), 1.0f);
    return static_cast<float>(up - 1) + fr;
}

// level_index that also returns the upper level it found (for level_index_near).
__device__ __forceinline__ float level_index_up(const float* __restrict__ lev, int nz, int top, float z, int& u) {
    int cnt = 0;
    for (int step = top; step > 0; step >>= 1) cnt = (lev[cnt + step - 1] < z) ? cnt + step : cnt;
    const int up = min(max(cnt, 1), nz - 1);
    u = up;
    const float lo = lev[up - 1];
    const float hi = lev[up];
    float fr = (z - lo) * __builtin_amdgcn_rcpf(fmaxf(hi - lo, 1e-6f));
    fr = fminf(fmaxf(fr, 0.0f), 1.0f);
    return static_cast<float>(up - 1) + fr;
}

// Byte offsets and weights of the 8 trilinear corners (order x fastest, then
// y, then z: 000 001 010 011 100 101 110 111).
struct Corner3 {
    uint32_t o[8];
    float w[8];
};

// Trilinear corners at fractional indices (ix, iy, iz) (already clamped) of a
// [nz][ny][nx] array of rb-byte records (the whole array below 4 GiB).
__device__ __forceinline__ Corner3 corners3(float ix, float iy, float iz, int nx, int ny, int nz, uint32_t rb) {
    const float fx0 = floorf(ix), fy0 = floorf(iy), fz0 = floorf(iz);
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0), z0 = static_cast<int>(fz0);
    const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
    const float fx = ix - fx0, fy = iy - fy0, fz = iz - fz0;
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    // column-major records: (row, column, level), a column's levels contiguous
    const uint32_t colb = static_cast<uint32_t>(nz) * rb;
    const uint32_t row = static_cast<uint32_t>(nx) * colb;
    const uint32_t dx = static_cast<uint32_t>(x1 - x0) * colb;
    const uint32_t dy = static_cast<uint32_t>(y1 - y0) * row;
    const uint32_t dz = static_cast<uint32_t>(z1 - z0) * rb;
    const uint32_t o = static_cast<uint32_t>(y0) * row + static_cast<uint32_t>(x0) * colb + static_cast<uint32_t>(z0) * rb;
    Corner3 c;
    c.o[0] = o;
    c.o[1] = o + dx;
    c.o[2] = o + dy;
    c.o[3] = o + dy + dx;
    c.o[4] = o + dz;
    c.o[5] = o + dz + dx;
    c.o[6] = o + dz + dy;
    c.o[7] = o + dz + dy + dx;
    c.w[0] = gx * gy * gz; c.w[1] = fx * gy * gz;
    c.w[2] = gx * fy * gz; c.w[3] = fx * fy * gz;
    c.w[4] = gx * gy * fz; c.w[5] = fx * gy * fz;
    c.w[6] = gx * fy * fz; c.w[7] = fx * fy * fz;
    return c;
}

// Trilinear sample of the first K (<= 8) channels of a record field (K <= 6
// loads the first 6 values of each record only).
template <bool BF16, int K>
__device__ __forceinline__ void sample_rec(const void* base, const Corner3& c, float (&acc)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if constexpr (K <= 6) {
            float v[6];
            Rec8<BF16>::load6_at(base, c.o[q], v);
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += c.w[q] * v[k];
        } else {
            float v[8];
            Rec8<BF16>::load_at(base, c.o[q], v);
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += c.w[q] * v[k];
        }
    }
}

// Packed window layouts (GlideWindow): bfloat16 records of 8 values, 16 B per
// cell; float32 in two planes over the same cell index, plane A (16 B per
// cell) followed by plane B (8 B per cell, from byte ncell * 16). Corner
// offsets (corners3 with rb = kCellUnit) are cell * 16 B for bfloat16 (the
// record) and cell * 8 B for float32 (plane B's entry; plane A's is at twice
// that). A float32 gather thus reads 16 B and 8 B entries of a column's
// levels, which share fewer cache lines than 32 B records would.
template <bool BF16>
struct WinRec {
    static constexpr uint32_t kCellUnit = BF16 ? 16u : 8u;
};
__device__ __forceinline__ float4 plane_a(const void* base, uint32_t o) {
    return *reinterpret_cast<const float4*>(static_cast<const char*>(base) + (o << 1));
}
__device__ __forceinline__ float plane_a0(const void* base, uint32_t o) {
    return *reinterpret_cast<const float*>(static_cast<const char*>(base) + (o << 1));
}
__device__ __forceinline__ float2 plane_b(const void* base, uint32_t planeb, uint32_t o) {
    return *reinterpret_cast<const float2*>(static_cast<const char*>(base) + planeb + o);
}

// Trilinear sample of the wind channels u_s v_s w_s u_e v_e w_e (float32
// plane A = u_s v_s u_e v_e, plane B = w_s w_e).
template <bool BF16>
__device__ __forceinline__ void sample_wind(const void* base, uint32_t planeb, const Corner3& c, float (&acc)[6]) {
    if constexpr (BF16) {
        sample_rec<true, 6>(base, c, acc);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 A = plane_a(base, c.o[q]);
            const float2 B = plane_b(base, planeb, c.o[q]);
            acc[0] += c.w[q] * A.x;
            acc[1] += c.w[q] * A.y;
            acc[2] += c.w[q] * B.x;
            acc[3] += c.w[q] * A.z;
            acc[4] += c.w[q] * A.w;
            acc[5] += c.w[q] * B.y;
        }
    }
}

// Trilinear sample of the support channels sigma_w_ft T_Lw_ft rho drho_dz
// (float32 plane A) into acc[0..3]; bfloat16 records also give the meander
// pair sigma_mu sigma_mv in acc[4..5] (float32: plane B, sample_aux_b).
template <bool BF16>
__device__ __forceinline__ void sample_aux_a(const void* base, const Corner3& c, float (&acc)[6]) {
    if constexpr (BF16) {
        sample_rec<true, 6>(base, c, acc);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 A = plane_a(base, c.o[q]);
            acc[0] += c.w[q] * A.x;
            acc[1] += c.w[q] * A.y;
            acc[2] += c.w[q] * A.z;
            acc[3] += c.w[q] * A.w;
        }
    }
}

// float32 support plane B (the meander pair sigma_mu sigma_mv).
__device__ __forceinline__ void sample_aux_b(const void* base, uint32_t planeb, const Corner3& c, float& m0, float& m1) {
    m0 = 0.0f;
    m1 = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float2 B = plane_b(base, planeb, c.o[q]);
        m0 += c.w[q] * B.x;
        m1 += c.w[q] * B.y;
    }
}

// sigma_w_ft (the support record's channel 0) at one corner offset.
template <bool BF16>
__device__ __forceinline__ float sample_sigma_ft_at(const void* base, uint32_t o) {
    if constexpr (BF16) {
        float v[2];
        Rec8<true>::load2_at(base, o, v);
        return v[0];
    } else {
        return plane_a0(base, o);
    }
}

// Trilinear sample of sigma_w_ft alone (the free-troposphere probes).
template <bool BF16>
__device__ __forceinline__ float sample_sigma_ft(const void* base, const Corner3& c) {
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += c.w[q] * sample_sigma_ft_at<BF16>(base, c.o[q]);
    return acc;
}

// Trilinear sample of channels 0..1 only (the free-troposphere pair).
template <bool BF16>
__device__ __forceinline__ void sample_rec2(const void* base, const Corner3& c, float& a0, float& a1) {
    a0 = 0.0f;
    a1 = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        float v[2];
        Rec8<BF16>::load2_at(base, c.o[q], v);
        a0 += c.w[q] * v[0];
        a1 += c.w[q] * v[1];
    }
}

}  // namespace glide
