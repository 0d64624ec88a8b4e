// Hourly met kernels: terrain-following resample (one thread per column and
// target level, coalesced along x) and the per-window packed-record build.
#include "closure.h"

namespace glide {

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
    double* p_part;  // [n_groups][za] per-wavefront sums of the regridded pressure (64 columns each)
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

// Target levels per thread of the resample, and 64-column wavefronts per block.
constexpr int kResLevels = 4;
constexpr int kResWaves = 4;

// One thread per (column, kResLevels consecutive target levels): the column's
// source heights are read once for all of its levels, in batches of
// independent loads, and the terrain slopes once per thread. A wavefront
// takes 64 consecutive columns, so every load and store is coalesced along x;
// each wavefront leaves one float64 pressure partial per target level.
__global__ __launch_bounds__(64 * kResWaves) void resample_kernel(const ResampleArgs a) {
#pragma clang fp contract(off)
    const int64_t ncol = static_cast<int64_t>(a.ny) * a.nx;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    // logical block: column blocks in order, each with its level groups next
    // to one another, and each XCD a contiguous range (the column block's
    // heights and fields stay in one L2 for all its level groups)
    const int n_lg = (a.za + kResLevels - 1) / kResLevels;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t cb = lb / n_lg;
    const int lg = static_cast<int>(lb % n_lg);
    const int64_t grp = cb * kResWaves + __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);  // uniform
    const int64_t col = grp * 64 + lane;
    const bool live = col < ncol;
    const int zp = a.zp;
    const int j0 = lg * kResLevels;
    const bool flip = *a.flipped != 0;
    auto src_idx = [&](int asc) { return flip ? (zp - 1 - asc) : asc; };

    // idx = searchsorted(h, tg, side="left") on the ascending heights, i.e. the
    // number of levels below tg, and k0 the number below ground; counting is
    // order-free, so the source order serves. For a float h and a double tg,
    // h < tg exactly when h < tg rounded up to a float.
    double tg[kResLevels];
    float tgu[kResLevels];
#pragma unroll
    for (int g = 0; g < kResLevels; ++g) {
        tg[g] = a.agl[min(j0 + g, a.za - 1)];
        tgu[g] = __double2float_ru(tg[g]);
    }
    double pval[kResLevels];
#pragma unroll
    for (int g = 0; g < kResLevels; ++g) pval[g] = 0.0;
    if (live) {
        int k0 = 0;
        int idx[kResLevels];
#pragma unroll
        for (int g = 0; g < kResLevels; ++g) idx[g] = 0;
        const float* hp = a.h + col;
        auto count = [&](float hk) {
            k0 += (hk < 0.0f) ? 1 : 0;
#pragma unroll
            for (int g = 0; g < kResLevels; ++g) idx[g] += (hk < tgu[g]) ? 1 : 0;
        };
        int k = 0;
        for (; k + 8 <= zp; k += 8) {
            float hv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) hv[u] = hp[static_cast<int64_t>(k + u) * ncol];
#pragma unroll
            for (int u = 0; u < 8; ++u) count(hv[u]);
        }
        for (; k < zp; ++k) count(hp[static_cast<int64_t>(k) * ncol]);

        const int y = static_cast<int>(col / a.nx);
        const int x = static_cast<int>(col % a.nx);
        const float* tc = a.terrain + col;
        const float dlat_g = grad1(tc, a.nx, y, a.ny, a.dlat);
        const float dlon_g = grad1(tc, 1, x, a.nx, a.dlon);
        const double dhdx = static_cast<double>(dlon_g) / a.mlon[y];
        const float dhdy = dlat_g / static_cast<float>(kMPerDegLatD);  // float32, as numpy keeps it
        const double ztop = a.agl[a.za - 1];
        const double ztop_c = ztop > 1.0 ? ztop : 1.0;
        // bracketing source levels and fractions of the thread's targets; every
        // level's loads go out together, a channel at a time
        int slo[kResLevels], shi[kResLevels];
        double fr[kResLevels];
        float wf[kResLevels];
#pragma unroll
        for (int g = 0; g < kResLevels; ++g) {
            int lower = idx[g] - 1;
            if (lower < 0) lower = 0;
            if (lower < k0) lower = k0;
            if (lower > zp - 2) lower = zp - 2;
            slo[g] = src_idx(lower);
            shi[g] = src_idx(lower + 1);
        }
        double hlo[kResLevels], hhi[kResLevels];
#pragma unroll
        for (int g = 0; g < kResLevels; ++g) {
            hlo[g] = static_cast<double>(a.h[static_cast<int64_t>(slo[g]) * ncol + col]);
            hhi[g] = static_cast<double>(a.h[static_cast<int64_t>(shi[g]) * ncol + col]);
        }
#pragma unroll
        for (int g = 0; g < kResLevels; ++g) {
            double den = hhi[g] - hlo[g];
            if (den < 1e-6) den = 1e-6;
            double q = (tg[g] - hlo[g]) / den;
            fr[g] = q < 0.0 ? 0.0 : (q > 1.0 ? 1.0 : q);
            wf[g] = static_cast<float>(fr[g]);
        }
        float uu[kResLevels], vv[kResLevels], ww[kResLevels];
        for (int c = 0; c < a.c3; ++c) {
            const float* fc = a.src3 + static_cast<int64_t>(c) * zp * ncol;
            float f0[kResLevels], f1[kResLevels];
#pragma unroll
            for (int g = 0; g < kResLevels; ++g) {
                f0[g] = fc[static_cast<int64_t>(slo[g]) * ncol + col];
                f1[g] = fc[static_cast<int64_t>(shi[g]) * ncol + col];
            }
#pragma unroll
            for (int g = 0; g < kResLevels; ++g) {
                const int j = j0 + g;
                const float r = (fabsf(wf[g]) < 0.5f) ? f0[g] + wf[g] * (f1[g] - f0[g])
                                                      : f1[g] - (f1[g] - f0[g]) * (1.0f - wf[g]);
                if (c == a.iu) uu[g] = r;
                if (c == a.iv) vv[g] = r;
                if (c == a.iw) ww[g] = r;
                if (c != a.iw && j < a.za) a.dst3[(static_cast<int64_t>(c) * a.za + j) * ncol + col] = r;
            }
        }
#pragma unroll
        for (int g = 0; g < kResLevels; ++g) {
            const int j = j0 + g;
            if (j >= a.za) break;
            // Slope correction of w in float64, rounded once (vertical_grid.py:296-316).
            double taper = 1.0 - tg[g] / ztop_c;
            taper = taper < 0.0 ? 0.0 : (taper > 1.0 ? 1.0 : taper);
            const double slope = static_cast<double>(uu[g]) * dhdx + static_cast<double>(vv[g] * dhdy);
            a.dst3[(static_cast<int64_t>(a.iw) * a.za + j) * ncol + col] =
                static_cast<float>(static_cast<double>(ww[g]) - taper * slope);
            const double p0 = a.p_hpa[slo[g]], p1 = a.p_hpa[shi[g]];
            pval[g] = (fabs(fr[g]) < 0.5) ? p0 + fr[g] * (p1 - p0) : p1 - (p1 - p0) * (1.0 - fr[g]);
        }
    }
    // per-wavefront pressure sums (a fixed butterfly: the same bits every run)
#pragma unroll
    for (int g = 0; g < kResLevels; ++g) {
        double v = pval[g];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        pval[g] = v;
    }
    const int64_t ngrp = (ncol + 63) / 64;
    if (lane == 0 && grp < ngrp) {
#pragma unroll
        for (int g = 0; g < kResLevels; ++g)
            if (j0 + g < a.za) a.p_part[grp * a.za + j0 + g] = pval[g];
    }
}

constexpr int kMeanStage = 1024;  // partials staged in LDS per pass

// One workgroup per level: the per-wavefront partials are loaded in parallel
// into LDS, then added in column order by one thread, so the float64 sum has
// a fixed order and the loads do not wait on one another.
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

// Window build tile: a workgroup takes kWinX consecutive columns of one row
// (the lanes of a wavefront, so the field reads coalesce) and kWinZ
// consecutive levels (one wavefront each). A column's kWinZ records are one
// 128 B run in the window (bfloat16: 8 records of 16 B; float32: 8 plane-A
// entries of 16 B and 8 plane-B entries of 8 B), written whole from LDS.
constexpr int kWinX = 64;
constexpr int kWinZ = 8;
constexpr int kWinT = kWinX + 2;  // stencil tile width: one column either side

// local_std's 3 x 3 stencil (two-pass population std, edge-clipped) over a
// staged tile of hourly means whose row 1, column lane + 1 is the centre;
// the flags say which neighbouring rows and columns exist. Same values, same
// order of the sums.
__device__ __forceinline__ float stencil_std3(const float (&t)[3][kWinT], int lane, bool ylo, bool yhi, bool xlo,
                                              bool xhi) {
    float v[9];
    bool in[9];
    int cnt = 0;
    float sum = 0.0f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int dy = q / 3, dx = q % 3;
        in[q] = (dy != 0 || ylo) && (dy != 2 || yhi) && (dx != 0 || xlo) && (dx != 2 || xhi);
        v[q] = t[dy][lane + dx];
        if (in[q]) {
            sum += v[q];
            ++cnt;
        }
    }
    const float mean = sum / static_cast<float>(cnt);
    float ss = 0.0f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        if (in[q]) {
            const float d = v[q] - mean;
            ss += d * d;
        }
    }
    return sqrtf(ss / static_cast<float>(cnt));
}

// One thread per (column, level). Everything a thread needs from its
// neighbours is evaluated once per workgroup into LDS first: the hourly-mean
// u and v over the 3 x 3 stencil rows of each level (the meander spread), and
// theta, rho, mean u and mean v of the row at the kWinZ + 2 levels the
// vertical gradients touch. The values and the order of every sum are those
// of the per-thread evaluation, so the records do not depend on the tiling.
template <bool BF16>
__global__ __launch_bounds__(kWinX * kWinZ) void window_kernel(const WindowArgs a) {
    // One LDS block serves two phases: first the neighbour tiles, then (after
    // every thread has its values in registers) the staged records.
    constexpr int kTileFloats = kWinZ * 2 * 3 * kWinT;  // [level][u, v][row y-1, y, y+1][column x0-1 ...]
    constexpr int kRowFloats = (kWinZ + 2) * 4 * kWinX;  // [level z0-1 ...][theta, rho, mean u, mean v][column]
    constexpr int kNeighbourBytes = 4 * (kTileFloats + kRowFloats);
    constexpr int kStageBytes = kWinX * kWinZ * (BF16 ? 32 : 48);  // [column][level] records
    constexpr int kSmemBytes = kNeighbourBytes > kStageBytes ? kNeighbourBytes : kStageBytes;
    __shared__ uint4 smem[(kSmemBytes + 15) / 16];
    auto tile = reinterpret_cast<float (*)[2][3][kWinT]>(smem);
    auto rowv = reinterpret_cast<float (*)[4][kWinX]>(reinterpret_cast<float*>(smem) + kTileFloats);
    uint4* const stage0 = smem;                       // bfloat16 records, or float32 plane-A entries
    uint4* const stage1 = smem + kWinX * kWinZ;
    uint2* const stage_b0 = reinterpret_cast<uint2*>(smem + 2 * kWinX * kWinZ);  // float32 plane-B entries
    uint2* const stage_b1 = stage_b0 + kWinX * kWinZ;
    const int64_t ncol = static_cast<int64_t>(a.ny) * a.nx;
    const int nz = a.nz, nx = a.nx, ny = a.ny;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    // the wavefront's level (uniform: its plane offsets are scalar arithmetic)
    const int wz = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int nxb = (nx + kWinX - 1) / kWinX;
    const int nzb = (nz + kWinZ - 1) / kWinZ;
    // logical block: rows in order, each row's segments and level groups next
    // to one another, and each XCD a contiguous range of rows (the stencil's
    // neighbouring rows stay in one L2)
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int zb = static_cast<int>(lb % nzb);
    const int rs = static_cast<int>(lb / nzb);
    const int y = rs / nxb;
    const int x0 = (rs % nxb) * kWinX;
    const int x = x0 + lane;
    const int z0 = zb * kWinZ;
    const int z = z0 + wz;
    const bool ok = x < nx && z < nz;
    const int64_t col = static_cast<int64_t>(y) * nx + (x < nx ? x : nx - 1);
    auto at = [&](const float* f, int ch, int zz, int64_t c) {
        return f[(static_cast<int64_t>(ch) * nz + zz) * ncol + c];
    };
    auto pres = [&](int zz) { return static_cast<float>(a.p_hpa[zz]) * 100.0f; };
    // the record's own winds, in flight while the neighbours are staged
    float su = 0.0f, sv = 0.0f, sw3 = 0.0f, eu = 0.0f, ev = 0.0f, ew3 = 0.0f;
    if (ok) {
        su = at(a.s3, 0, z, col);
        sv = at(a.s3, 1, z, col);
        sw3 = at(a.s3, 2, z, col);
        eu = at(a.e3, 0, z, col);
        ev = at(a.e3, 1, z, col);
        ew3 = at(a.e3, 2, z, col);
    }
    const bool tiled = a.radius == 1;
    if (tiled && z < nz) {
        // hourly-mean u and v over rows y-1..y+1, columns x0-1..x0+kWinX
#pragma unroll
        for (int it = 0; it < (3 * kWinT + 63) / 64; ++it) {
            const int q = lane + 64 * it;
            if (q < 3 * kWinT) {
                const int r = q / kWinT, t = q % kWinT;
                const int yy = y + r - 1, xx = x0 + t - 1;
                float mu = 0.0f, mv = 0.0f;
                if (yy >= 0 && yy < ny && xx >= 0 && xx < nx) {
                    const int64_t c = static_cast<int64_t>(yy) * nx + xx;
                    mu = 0.5f * (at(a.s3, 0, z, c) + at(a.e3, 0, z, c));
                    mv = 0.5f * (at(a.s3, 1, z, c) + at(a.e3, 1, z, c));
                }
                tile[wz][0][r][t] = mu;
                tile[wz][1][r][t] = mv;
            }
        }
    }
    // theta, rho, mean u, mean v of the row at levels z0-1 .. z0+kWinZ
#pragma unroll
    for (int it = 0; it < ((kWinZ + 2) * kWinX + kWinX * kWinZ - 1) / (kWinX * kWinZ); ++it) {
        const int lz = wz + it * kWinZ, t = lane;  // wavefront-uniform level slot
        const int zz = z0 - 1 + lz;
        if (lz < kWinZ + 2 && zz >= 0 && zz < nz && x0 + t < nx) {
            const int64_t c = static_cast<int64_t>(y) * nx + x0 + t;
            const float tm = 0.5f * (at(a.s3, 3, zz, c) + at(a.e3, 3, zz, c));
            const float pz = pres(zz);
            rowv[lz][0][t] = potential_temperature(tm, pz);
            rowv[lz][1][t] = pz / (287.05f * fmaxf(tm, 1.0f));
            rowv[lz][2][t] = 0.5f * (at(a.s3, 0, zz, c) + at(a.e3, 0, zz, c));
            rowv[lz][3][t] = 0.5f * (at(a.s3, 1, zz, c) + at(a.e3, 1, zz, c));
        }
    }
    __syncthreads();
    float sw = 0.0f, tl = 0.0f, rho_z = 0.0f, drho = 0.0f, smu = 0.0f, smv = 0.0f;
    if (ok) {
        const int zl = z == 0 ? 0 : z - 1;
        const int zh = z == nz - 1 ? nz - 1 : z + 1;
        const int ll = zl - z0 + 1, lh = zh - z0 + 1, lc = wz + 1;
        const float dh = safe_den(a.lev[zh] - a.lev[zl]);
        // vertical gradients (central inside, one-sided at the ends)
        const float th = rowv[lc][0][lane];
        const float dth = (rowv[lh][0][lane] - rowv[ll][0][lane]) / dh;
        const float du = (rowv[lh][2][lane] - rowv[ll][2][lane]) / dh;
        const float dv = (rowv[lh][3][lane] - rowv[ll][3][lane]) / dh;
        rho_z = rowv[lc][1][lane];
        drho = (rowv[lh][1][lane] - rowv[ll][1][lane]) / dh;
        const float n2 = brunt_vaisala_squared(th, dth);
        const float shear_sq = du * du + dv * dv;
        const float shear_mag = sqrtf(fmaxf(shear_sq, 1e-8f));
        const float ri = gradient_richardson(n2, shear_sq);
        const float kz = free_trop_diffusivity(a.lev[z], shear_mag, ri);
        free_trop_sigma_tl(kz, n2, sw, tl);
        if (tiled) {
            const bool ylo = y > 0, yhi = y < ny - 1, xlo = x > 0, xhi = x < nx - 1;
            smu = a.coeff * stencil_std3(tile[wz][0], lane, ylo, yhi, xlo, xhi);
            smv = a.coeff * stencil_std3(tile[wz][1], lane, ylo, yhi, xlo, xhi);
        } else {
            const int64_t zoff = static_cast<int64_t>(z) * ncol;
            smu = a.coeff * local_std(a.s3 + zoff, a.e3 + zoff, y, x, ny, nx, a.radius);
            smv = a.coeff * local_std(a.s3 + (static_cast<int64_t>(nz) + z) * ncol,
                                      a.e3 + (static_cast<int64_t>(nz) + z) * ncol, y, x, ny, nx, a.radius);
        }
    }
    __syncthreads();  // every neighbour read is done: the records take the LDS over
    const int sl = lane * kWinZ + wz;
    if (ok) {
        if constexpr (BF16) {
            float wr[8] = {su, sv, sw3, eu, ev, ew3, 0.0f, 0.0f};
            float xr[8] = {sw, tl, rho_z, drho, smu, smv, 0.0f, 0.0f};
            Rec8<true>::store(stage0, sl, wr);
            Rec8<true>::store(stage1, sl, xr);
        } else {
            stage0[sl] = make_uint4(__float_as_uint(su), __float_as_uint(sv), __float_as_uint(eu),
                                    __float_as_uint(ev));
            stage_b0[sl] = make_uint2(__float_as_uint(sw3), __float_as_uint(ew3));
            stage1[sl] = make_uint4(__float_as_uint(sw), __float_as_uint(tl), __float_as_uint(rho_z),
                                    __float_as_uint(drho));
            stage_b1[sl] = make_uint2(__float_as_uint(smu), __float_as_uint(smv));
        }
    }
    __syncthreads();
    // column-major: a column's levels are contiguous, so each column's kWinZ
    // staged entries are one run in the window (float32: in each plane)
    const int64_t ncell = ncol * nz;
    {
        const int q = static_cast<int>(threadIdx.x);
        const int xx = x0 + q / kWinZ;
        const int zz = z0 + q % kWinZ;
        if (xx < nx && zz < nz) {
            const int64_t g = (static_cast<int64_t>(y) * nx + xx) * nz + zz;
            static_cast<uint4*>(a.wind)[g] = stage0[q];
            static_cast<uint4*>(a.aux)[g] = stage1[q];
            if constexpr (!BF16) {
                reinterpret_cast<uint2*>(static_cast<uint4*>(a.wind) + ncell)[g] = stage_b0[q];
                reinterpret_cast<uint2*>(static_cast<uint4*>(a.aux) + ncell)[g] = stage_b1[q];
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
    sr[4] = at(a.s3, 3, 0, col);
    sr[9] = at(a.e3, 3, 0, col);
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
    const int64_t nblk = (ncol + 63) / 64;  // pressure partials: one per 64-column wavefront
    // Small device scratch: orientation flag + per-wavefront pressure sums.
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
    const int64_t n_rb = ((nblk + kResWaves - 1) / kResWaves) * ((za + kResLevels - 1) / kResLevels);
    GLIDE_REQUIRE(n_rb <= 0x7fffffff, "resample grid too large");
    hipLaunchKernelGGL(resample_kernel, dim3(static_cast<unsigned>(n_rb)), dim3(64 * kResWaves), 0, st, a);
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
    const bool bf16 = dtype == GLIDE_MET_BF16;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nxb = (static_cast<int64_t>(nx) + kWinX - 1) / kWinX;
    const int64_t n_wb = nxb * ny * ((nz + kWinZ - 1) / kWinZ);
    GLIDE_REQUIRE(n_wb <= 0x7fffffff, "window grid too large");
    const dim3 grid(static_cast<unsigned>(n_wb));
    if (bf16) {
        hipLaunchKernelGGL(window_kernel<true>, grid, dim3(kWinX * kWinZ), 0, st, a);
    } else {
        hipLaunchKernelGGL(window_kernel<false>, grid, dim3(kWinX * kWinZ), 0, st, a);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
