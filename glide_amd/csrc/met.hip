// Hourly met kernels: terrain-following resample (one thread per column and
// target level, coalesced along x) and the per-window packed-record build.
#include "closure.h"

namespace glide {

constexpr int kColBlock = 256;
constexpr double kMPerDegLatD = 110540.0;

// ---------------------------------------------------------------------------
// Resample: vertical_grid.py:166-316 + met_reader.py:641-693
// ---------------------------------------------------------------------------

// Orientation of the source levels: top-of-atmosphere first when the bbox-mean
// height of level 0 exceeds that of the last level (vertical_grid.py:195-196).
// kOrientBlocks partial sums in float64, then one thread adds them in order.
constexpr int kOrientBlocks = 64;

__global__ __launch_bounds__(256) void level_orientation_kernel(const float* __restrict__ h, int32_t zp,
                                                                int64_t ncol, double* __restrict__ part) {
    __shared__ double s0[256], s1[256];
    double a = 0.0, b = 0.0;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; c < ncol; c += 256 * kOrientBlocks) {
        a += static_cast<double>(h[c]);
        b += static_cast<double>(h[static_cast<int64_t>(zp - 1) * ncol + c]);
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] += s1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

__global__ void level_orientation_final(const double* __restrict__ part, int32_t* __restrict__ flipped) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < kOrientBlocks; ++k) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    *flipped = (a > b) ? 1 : 0;
}

struct ResampleArgs {
    const float* src3;
    int32_t c3, zp, ny, nx;
    const float* h;
    const double* agl;
    int32_t za;
    const float* terrain;
    const double* mlon;
    double dlat, dlon;
    const double* p_hpa;
    int32_t iu, iv, iw;
    float* dst3;
    double* p_part;  // [n_blocks][za] per-block sums of the regridded pressure
    const int32_t* flipped;
};

// Terrain slope per degree as numpy.gradient computes it on a float32 field
// with a uniform float64 coordinate step: float32 differences, float64
// division, result rounded to float32.
__device__ __forceinline__ float grad1(const float* __restrict__ f, int64_t stride, int i, int n, double d) {
#pragma clang fp contract(off)
    if (i == 0) return static_cast<float>(static_cast<double>(f[stride] - f[0]) / d);
    if (i == n - 1) return static_cast<float>(static_cast<double>(f[0] - f[-stride]) / d);
    return static_cast<float>(static_cast<double>(f[stride] - f[-stride]) / (2.0 * d));
}

// One thread per (column, target level): blockIdx.y is the level, so each
// block's pressure partial covers the same 256 columns at one level.
__global__ __launch_bounds__(kColBlock) void resample_kernel(const ResampleArgs a) {
#pragma clang fp contract(off)
    __shared__ double psum[kColBlock];
    const int64_t ncol = static_cast<int64_t>(a.ny) * a.nx;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const bool live = col < ncol;
    const int zp = a.zp;
    const bool flip = *a.flipped != 0;
    // ascending index -> source index
    auto src_idx = [&](int asc) { return flip ? (zp - 1 - asc) : asc; };
    auto hasc = [&](int asc) { return static_cast<double>(a.h[static_cast<int64_t>(src_idx(asc)) * ncol + col]); };

    double pval = 0.0;
    if (live) {
        const double tg = a.agl[j];
        // k0 = number of sub-surface levels; idx = searchsorted(h, tg,
        // side="left"), which on the ascending heights is the number of levels
        // below tg: one pass of independent loads (a scan that stops at the
        // first level >= tg makes every load wait for the one before)
        int k0 = 0, idx = 0;
        for (int k = 0; k < zp; ++k) {
            const double hk = hasc(k);
            k0 += (hk < 0.0) ? 1 : 0;
            idx += (hk < tg) ? 1 : 0;
        }
        const int y = static_cast<int>(col / a.nx);
        const int x = static_cast<int>(col % a.nx);
        const float* tc = a.terrain + col;
        const float dlat_g = grad1(tc, a.nx, y, a.ny, a.dlat);
        const float dlon_g = grad1(tc, 1, x, a.nx, a.dlon);
        const double dhdx = static_cast<double>(dlon_g) / a.mlon[y];
        const float dhdy = dlat_g / static_cast<float>(kMPerDegLatD);  // float32, as numpy keeps it
        const double ztop = a.agl[a.za - 1];
        const double ztop_c = ztop > 1.0 ? ztop : 1.0;
        int lower = idx - 1;
        if (lower < 0) lower = 0;
        if (lower < k0) lower = k0;
        if (lower > zp - 2) lower = zp - 2;
        const double hlo = hasc(lower), hhi = hasc(lower + 1);
        double den = hhi - hlo;
        if (den < 1e-6) den = 1e-6;
        double fr = (tg - hlo) / den;
        fr = fr < 0.0 ? 0.0 : (fr > 1.0 ? 1.0 : fr);
        const int slo = src_idx(lower), shi = src_idx(lower + 1);
        const float wf = static_cast<float>(fr);
        float uu = 0.0f, vv = 0.0f, ww = 0.0f;
        for (int c = 0; c < a.c3; ++c) {
            const float* fc = a.src3 + static_cast<int64_t>(c) * zp * ncol;
            const float f0 = fc[static_cast<int64_t>(slo) * ncol + col];
            const float f1 = fc[static_cast<int64_t>(shi) * ncol + col];
            const float r = (fabsf(wf) < 0.5f) ? f0 + wf * (f1 - f0) : f1 - (f1 - f0) * (1.0f - wf);
            if (c == a.iu) uu = r;
            if (c == a.iv) vv = r;
            if (c == a.iw) ww = r;
            if (c != a.iw) a.dst3[(static_cast<int64_t>(c) * a.za + j) * ncol + col] = r;
        }
        // Slope correction of w in float64, rounded once (vertical_grid.py:296-316).
        double taper = 1.0 - tg / ztop_c;
        taper = taper < 0.0 ? 0.0 : (taper > 1.0 ? 1.0 : taper);
        const double slope = static_cast<double>(uu) * dhdx + static_cast<double>(vv * dhdy);
        a.dst3[(static_cast<int64_t>(a.iw) * a.za + j) * ncol + col] =
            static_cast<float>(static_cast<double>(ww) - taper * slope);
        const double p0 = a.p_hpa[slo], p1 = a.p_hpa[shi];
        pval = (fabs(fr) < 0.5) ? p0 + fr * (p1 - p0) : p1 - (p1 - p0) * (1.0 - fr);
    }
    psum[threadIdx.x] = pval;
    __syncthreads();
    for (int off = kColBlock / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) psum[threadIdx.x] += psum[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.p_part[static_cast<int64_t>(blockIdx.x) * a.za + j] = psum[0];
}

constexpr int kMeanStage = 1024;  // partials staged in LDS per pass

// One workgroup per level: the per-block partials are loaded in parallel into
// LDS, then added in block order by one thread (the order the float64 sum
// has always had), so the loads no longer wait on one another.
__global__ __launch_bounds__(256) void pressure_mean_kernel(const double* __restrict__ part, int64_t nblk,
                                                            int32_t za, int64_t ncol, double* __restrict__ out) {
    __shared__ double stage[kMeanStage];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int64_t b0 = 0; b0 < nblk; b0 += kMeanStage) {
        const int64_t nb = nblk - b0 < kMeanStage ? nblk - b0 : kMeanStage;
        for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) stage[b] = part[(b0 + b) * za + j];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t b = 0; b < nb; ++b) s += stage[b];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = s / static_cast<double>(ncol);
}

// ---------------------------------------------------------------------------
// Window build: hanna.py:747-853 staging + hanna.py:1694-1830 support stacks
// ---------------------------------------------------------------------------
struct WindowArgs {
    const float* s3;
    const float* e3;
    const float* s2;
    const float* e2;
    const double* p_hpa;
    const float* lev;
    int32_t nz, ny, nx, radius;
    float coeff;
    void* wind;
    void* aux;
    void* surf;
};

__device__ __forceinline__ float safe_den(float d) { return fabsf(d) < 1e-6f ? 1e-6f : d; }

// Levels per workgroup of the window build: one 128 B run of records per
// column (bfloat16: 8 records of 16 B; float32: 8 plane-A entries of 16 B, and
// 8 plane-B entries of 8 B), so the column-major writes go out in whole or
// half cache lines.
template <bool BF16>
constexpr int window_levels() { return 8; }

// One thread per (column, level); a workgroup takes kColBlock / kTz
// consecutive columns (the fastest index, so the field reads coalesce) times
// kTz consecutive levels (blockIdx.y). The records are staged in LDS and
// written column by column. Level 0's threads also write the surface record.
template <bool BF16>
__global__ __launch_bounds__(kColBlock) void window_kernel(const WindowArgs a) {
    constexpr int kTz = window_levels<BF16>();
    constexpr int kTc = kColBlock / kTz;
    // staged [column][level]: bfloat16 records (wind, aux), or float32 plane-A
    // entries (wind, aux) and plane-B entries (wind, aux)
    __shared__ uint4 stage[2][kColBlock];
    __shared__ uint2 stage_b[BF16 ? 1 : 2][BF16 ? 1 : kColBlock];
    const int64_t ncol = static_cast<int64_t>(a.ny) * a.nx;
    const int tc = static_cast<int>(threadIdx.x) % kTc;
    const int tz = static_cast<int>(threadIdx.x) / kTc;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * kTc + tc;
    const int nz = a.nz;
    const int z = static_cast<int>(blockIdx.y) * kTz + tz;
    const bool ok = col < ncol && z < nz;
    const int y = static_cast<int>((ok ? col : 0) / a.nx);
    const int x = static_cast<int>((ok ? col : 0) % a.nx);
    auto at = [&](const float* f, int ch, int zz) {
        return f[(static_cast<int64_t>(ch) * nz + zz) * ncol + col];
    };
    auto pres = [&](int zz) { return static_cast<float>(a.p_hpa[zz]) * 100.0f; };
    auto theta = [&](int zz) {
        const float tm = 0.5f * (at(a.s3, 3, zz) + at(a.e3, 3, zz));
        return potential_temperature(tm, pres(zz));
    };
    auto umid = [&](int zz) { return 0.5f * (at(a.s3, 0, zz) + at(a.e3, 0, zz)); };
    auto vmid = [&](int zz) { return 0.5f * (at(a.s3, 1, zz) + at(a.e3, 1, zz)); };
    auto rho = [&](int zz) {
        const float tm = 0.5f * (at(a.s3, 3, zz) + at(a.e3, 3, zz));
        return pres(zz) / (287.05f * fmaxf(tm, 1.0f));
    };
    if (ok) {
        const int zl = z == 0 ? 0 : z - 1;
        const int zh = z == nz - 1 ? nz - 1 : z + 1;
        const float dh = safe_den(a.lev[zh] - a.lev[zl]);
        // vertical gradients (central inside, one-sided at the ends)
        const float th = theta(z);
        const float dth = (theta(zh) - theta(zl)) / dh;
        const float du = (umid(zh) - umid(zl)) / dh;
        const float dv = (vmid(zh) - vmid(zl)) / dh;
        const float drho = (rho(zh) - rho(zl)) / dh;
        const float n2 = brunt_vaisala_squared(th, dth);
        const float shear_sq = du * du + dv * dv;
        const float shear_mag = sqrtf(fmaxf(shear_sq, 1e-8f));
        const float ri = gradient_richardson(n2, shear_sq);
        const float kz = free_trop_diffusivity(a.lev[z], shear_mag, ri);
        float sw, tl;
        free_trop_sigma_tl(kz, n2, sw, tl);
        const int64_t zoff = static_cast<int64_t>(z) * ncol;
        const float smu = a.coeff * local_std(a.s3 + zoff, a.e3 + zoff, y, x, a.ny, a.nx, a.radius);
        const float smv = a.coeff * local_std(a.s3 + (static_cast<int64_t>(nz) + z) * ncol,
                                              a.e3 + (static_cast<int64_t>(nz) + z) * ncol, y, x, a.ny, a.nx,
                                              a.radius);
        const int sl = tc * kTz + tz;
        if constexpr (BF16) {
            float wr[8] = {at(a.s3, 0, z), at(a.s3, 1, z), at(a.s3, 2, z), at(a.e3, 0, z),
                           at(a.e3, 1, z), at(a.e3, 2, z), 0.0f, 0.0f};
            float xr[8] = {sw, tl, rho(z), drho, smu, smv, 0.0f, 0.0f};
            Rec8<true>::store(stage[0], sl, wr);
            Rec8<true>::store(stage[1], sl, xr);
        } else {
            stage[0][sl] = make_uint4(__float_as_uint(at(a.s3, 0, z)), __float_as_uint(at(a.s3, 1, z)),
                                      __float_as_uint(at(a.e3, 0, z)), __float_as_uint(at(a.e3, 1, z)));
            stage_b[0][sl] = make_uint2(__float_as_uint(at(a.s3, 2, z)), __float_as_uint(at(a.e3, 2, z)));
            stage[1][sl] = make_uint4(__float_as_uint(sw), __float_as_uint(tl), __float_as_uint(rho(z)),
                                      __float_as_uint(drho));
            stage_b[1][sl] = make_uint2(__float_as_uint(smu), __float_as_uint(smv));
        }
    }
    __syncthreads();
    // column-major: a column's levels are contiguous, so each column's kTz
    // staged entries are one run in the window (float32: in each plane)
    const int64_t ncell = ncol * nz;
    for (int q = static_cast<int>(threadIdx.x); q < kColBlock; q += kColBlock) {
        const int64_t c = static_cast<int64_t>(blockIdx.x) * kTc + q / kTz;
        const int zz = static_cast<int>(blockIdx.y) * kTz + q % kTz;
        if (c < ncol && zz < nz) {
            const int64_t g = c * nz + zz;
            static_cast<uint4*>(a.wind)[g] = stage[0][q];
            static_cast<uint4*>(a.aux)[g] = stage[1][q];
            if constexpr (!BF16) {
                reinterpret_cast<uint2*>(static_cast<uint4*>(a.wind) + ncell)[g] = stage_b[0][q];
                reinterpret_cast<uint2*>(static_cast<uint4*>(a.aux) + ncell)[g] = stage_b[1][q];
            }
        }
    }
    if (!ok || z != 0) return;
    // Surface record: blh sp ustar shf t(level 0) at start then end.
    float sr[12];
    for (int c = 0; c < 4; ++c) {
        sr[c] = a.s2[static_cast<int64_t>(c) * ncol + col];
        sr[5 + c] = a.e2[static_cast<int64_t>(c) * ncol + col];
    }
    sr[4] = at(a.s3, 3, 0);
    sr[9] = at(a.e3, 3, 0);
    sr[10] = 0.0f;
    sr[11] = 0.0f;
    // float in both layouts (the surface scalars are not rounded to bfloat16)
    float4* p = static_cast<float4*>(a.surf) + col * 3;
    p[0] = make_float4(sr[0], sr[1], sr[2], sr[3]);
    p[1] = make_float4(sr[4], sr[5], sr[6], sr[7]);
    p[2] = make_float4(sr[8], sr[9], sr[10], sr[11]);
}

}  // namespace glide

using namespace glide;

extern "C" int glide_resample_hour(const float* src3, int32_t c3, int32_t zp, int32_t ny, int32_t nx,
                                   const float* height_agl, const double* agl_levels, int32_t za,
                                   const float* terrain, const double* m_per_deg_lon, double dlat_deg,
                                   double dlon_deg, const double* p_level_hpa, int32_t iu, int32_t iv,
                                   int32_t iw, float* dst3, double* p_out_hpa, void* stream) {
    GLIDE_REQUIRE(src3 && height_agl && agl_levels && terrain && m_per_deg_lon && p_level_hpa && dst3 && p_out_hpa,
                  "glide_resample_hour: NULL argument");
    GLIDE_REQUIRE(zp >= 2 && za >= 1 && ny >= 2 && nx >= 2 && c3 >= 1, "bad resample dimensions");
    GLIDE_REQUIRE(iu >= 0 && iu < c3 && iv >= 0 && iv < c3 && iw >= 0 && iw < c3, "bad u/v/w channel index");
    GLIDE_REQUIRE(dlat_deg != 0.0 && dlon_deg != 0.0, "grid spacing must be non-zero");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t ncol = static_cast<int64_t>(ny) * nx;
    const int64_t nblk = (ncol + kColBlock - 1) / kColBlock;
    // Small device scratch: orientation flag + per-block pressure sums.
    static thread_local void* scratch = nullptr;
    static thread_local size_t scratch_bytes = 0;
    const size_t need = 256 + sizeof(double) * 2 * kOrientBlocks + sizeof(double) * static_cast<size_t>(nblk) * za;
    if (scratch_bytes < need) {
        if (scratch) GLIDE_HIP(hipFree(scratch));
        scratch = nullptr;
        scratch_bytes = 0;
        hipError_t e = hipMalloc(&scratch, need);
        if (e != hipSuccess) return set_error(GLIDE_ERR_MEMORY, "hipMalloc(%zu) failed for resample scratch", need);
        scratch_bytes = need;
    }
    int32_t* flipped = static_cast<int32_t*>(scratch);
    double* opart = reinterpret_cast<double*>(static_cast<char*>(scratch) + 256);
    double* part = opart + 2 * kOrientBlocks;
    hipLaunchKernelGGL(level_orientation_kernel, dim3(kOrientBlocks), dim3(256), 0, st, height_agl, zp, ncol, opart);
    hipLaunchKernelGGL(level_orientation_final, dim3(1), dim3(1), 0, st, opart, flipped);
    ResampleArgs a{src3, c3, zp, ny, nx, height_agl, agl_levels, za, terrain, m_per_deg_lon, dlat_deg, dlon_deg,
                   p_level_hpa, iu, iv, iw, dst3, part, flipped};
    hipLaunchKernelGGL(resample_kernel, dim3(static_cast<unsigned>(nblk), static_cast<unsigned>(za)), dim3(kColBlock),
                       0, st, a);
    hipLaunchKernelGGL(pressure_mean_kernel, dim3(static_cast<unsigned>(za)), dim3(256), 0, st, part, nblk, za, ncol,
                       p_out_hpa);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_build_window(const float* h3_s, const float* h3_e, const float* h2_s, const float* h2_e,
                                  const double* p_level_hpa, const float* levels, int32_t nz, int32_t ny,
                                  int32_t nx, double meander_coeff, int32_t meander_radius, int32_t dtype,
                                  void* wind, void* aux, void* surf, void* stream) {
    GLIDE_REQUIRE(h3_s && h3_e && h2_s && h2_e && p_level_hpa && levels && wind && aux && surf,
                  "glide_build_window: NULL argument");
    GLIDE_REQUIRE(nz >= 2 && ny >= 1 && nx >= 1, "bad window dimensions");
    GLIDE_REQUIRE(meander_radius >= 1, "meander_stencil_radius must be >= 1");
    GLIDE_REQUIRE(meander_coeff > 0.0, "meander_coefficient must be > 0");
    GLIDE_REQUIRE(dtype == GLIDE_MET_F32 || dtype == GLIDE_MET_BF16, "unknown met dtype");
    WindowArgs a{h3_s, h3_e, h2_s, h2_e, p_level_hpa, levels, nz, ny, nx, meander_radius,
                 static_cast<float>(meander_coeff), wind, aux, surf};
    const int64_t ncol = static_cast<int64_t>(ny) * nx;
    const bool bf16 = dtype == GLIDE_MET_BF16;
    const int tz = bf16 ? window_levels<true>() : window_levels<false>();
    const int tc = kColBlock / tz;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 grid(static_cast<unsigned>((ncol + tc - 1) / tc), static_cast<unsigned>((nz + tz - 1) / tz));
    if (bf16) {
        hipLaunchKernelGGL(window_kernel<true>, grid, dim3(kColBlock), 0, st, a);
    } else {
        hipLaunchKernelGGL(window_kernel<false>, grid, dim3(kColBlock), 0, st, a);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
