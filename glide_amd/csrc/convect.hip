// Per-column reduced-Emanuel convection, backward (transpose) sampling.
//
// The reference builds one non-divergent mass-flux matrix for the bbox-mean
// column (emanuel.py:429-506) and samples each active particle's destination
// from the renormalised transposed column of its host level
// (emanuel.py:720-805). Here every met column carries its own cloud base/top
// and cloud-base flux; the matrix is structured (rank-1 updraft + sub-diagonal
// subsidence + diagonal closure), so each particle evaluates the entries of its
// column on the fly in O(nz) instead of materialising Z x Z per column.
#include "footprint_key.h"

namespace glide {

constexpr int kConvLevels = 128;

struct ColumnArgs {
    const float* conv;  // [4][ny][nx]: cape, lcl_h, top_h, dtv
    const float* t_s;
    const float* t_e;  // [nz][ny][nx]
    float alpha;
    const double* p_hpa;
    const float* lev;
    int nz, ny, nx;
    float closure_c, trigger_dtv, min_cape, min_depth;
    int32_t* i_lcl;
    int32_t* i_lnb;
    float* m_b;
};

// Column descriptors (emanuel.py:629-675 applied per column).
__global__ void convect_columns_kernel(const ColumnArgs a) {
    const int64_t ncol = static_cast<int64_t>(a.ny) * a.nx;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const float cape = a.conv[c];
    const float lcl_h = a.conv[ncol + c];
    const float top_h = a.conv[2 * ncol + c];
    const float dtv = a.conv[3 * ncol + c];
    int lcl = -1, lnb = a.nz - 1;
    for (int k = 0; k < a.nz; ++k) {
        if (a.lev[k] >= lcl_h) {
            lcl = k;
            break;
        }
    }
    for (int k = 0; k < a.nz; ++k) {
        if (a.lev[k] >= top_h) {
            lnb = k;
            break;
        }
    }
    bool fires = lcl >= 0 && lnb > lcl && cape >= a.min_cape && dtv >= a.trigger_dtv;
    if (fires && (a.lev[lnb] - a.lev[lcl]) < a.min_depth) fires = false;
    float mb = 0.0f;
    if (fires) {
        const float ts = a.t_s[static_cast<int64_t>(lcl) * ncol + c];
        const float te = a.t_e[static_cast<int64_t>(lcl) * ncol + c];
        const float t_lcl = ts * (1.0f - a.alpha) + te * a.alpha;
        const float p_lcl = static_cast<float>(a.p_hpa[lcl]) * 100.0f;
        const float rho = p_lcl / (287.05f * fmaxf(t_lcl, 150.0f));
        const float wb = fminf(sqrtf(2.0f * fmaxf(cape, 0.0f)), 5.0f);
        mb = a.closure_c * rho * wb * 3600.0f;
    }
    a.i_lcl[c] = fires ? lcl : -1;
    a.i_lnb[c] = fires ? lnb : -1;
    a.m_b[c] = mb;
}

// ---------------------------------------------------------------------------
// Surface-parcel lift per column (compute_lcl_lnb_cape, emanuel.py:282-399,
// with the free functions of emanuel.py:124-255), in float64. One wavefront
// per column: lanes take levels for the fixed 30-step moist-adiabat bisection,
// then lane 0 scans the levels in order (LCL, LNB, the CAPE sum), so the
// result does not depend on the launch.
// ---------------------------------------------------------------------------
namespace thermo {
constexpr double kG = 9.80665, kRd = 287.05, kRv = 461.5, kCp = 1005.0, kLv = 2.5e6, kP0 = 100000.0;
constexpr double kEps = kRd / kRv, kKappa = kRd / kCp, kPmin = 1.0, kTmin = 150.0;

__device__ __forceinline__ double esat(double t) {  // Bolton (1980) eq. 10, Pa
    const double tc = t - 273.15;
    return 611.2 * exp(17.67 * tc / (tc + 243.5));
}
__device__ __forceinline__ double q_to_e(double q, double p) { return q * p / (kEps + (1.0 - kEps) * q); }
__device__ __forceinline__ double e_to_q(double e, double p) { return kEps * e / (fmax(p, kPmin) - (1.0 - kEps) * e); }
__device__ __forceinline__ double tv(double t, double q) { return t * (1.0 + (1.0 / kEps - 1.0) * q); }
__device__ __forceinline__ double t_lcl(double t, double q, double p) {  // Bolton (1980) eq. 22
    const double e_hpa = fmax(q_to_e(q, p), 1.0) / 100.0;
    return 2840.0 / fmax(3.5 * log(t) - log(e_hpa) - 4.805, 1e-3) + 55.0;
}
__device__ __forceinline__ double theta(double t, double p) { return t * pow(kP0 / fmax(p, kPmin), kKappa); }
__device__ __forceinline__ double theta_e(double t, double q, double p) {
    return theta(t, p) * exp(kLv * q / (kCp * fmax(t_lcl(t, q, p), kTmin)));
}
__device__ __forceinline__ double qsat(double t, double p) { return e_to_q(esat(t), p); }
// Temperature on the moist adiabat of theta_e at pressure p: 30 halvings of
// [180, 350] K, no early exit (emanuel.py:215-255).
__device__ double lift_moist(double the, double p) {
    double lo = 180.0, hi = 350.0;
    for (int it = 0; it < 30; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (theta_e(mid, qsat(mid, p), p) > the) {
            hi = mid;
        } else {
            lo = mid;
        }
    }
    return 0.5 * (lo + hi);
}
}  // namespace thermo

constexpr int kCapeWaves = 4;

struct CapeArgs {
    const float *t_s, *t_e, *q_s, *q_e;  // [nz][ncol]
    double alpha;
    const double* p_hpa;
    const float* lev;
    int nz;
    int64_t ncol;
    int32_t* i_lcl;
    int32_t* i_lnb;
    double* cape;
    double* dtv;
    float* conv;  // [4][ncol] or NULL
};

__global__ __launch_bounds__(64 * kCapeWaves) void column_cape_kernel(const CapeArgs a) {
    using namespace thermo;
    __shared__ double s_t[kCapeWaves][kConvLevels], s_lp[kCapeWaves][kConvLevels];
    __shared__ double s_b[kCapeWaves][kConvLevels], s_d[kCapeWaves][kConvLevels];
    __shared__ uint8_t s_below[kCapeWaves][kConvLevels];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * kCapeWaves + wave;
    if (col >= a.ncol) return;  // whole waves leave together; no block barrier below
    const double al = a.alpha;
    auto lerp_at = [&](const float* s, const float* e, int k) {
        const int64_t o = static_cast<int64_t>(k) * a.ncol + col;
        return static_cast<double>(s[o]) * (1.0 - al) + static_cast<double>(e[o]) * al;
    };
    // surface parcel (level 0), evaluated by every lane
    const double t0 = lerp_at(a.t_s, a.t_e, 0);
    const double q0 = fmax(lerp_at(a.q_s, a.q_e, 0), 0.0);
    const double p0 = a.p_hpa[0] * 100.0;
    const double tl = t_lcl(t0, q0, p0);
    const double p_lcl = p0 * pow(tl / fmax(t0, kTmin), kCp / kRd);
    const double the = theta_e(t0, q0, p0);
    const double th0 = theta(t0, p0);
    for (int k = lane; k < a.nz; k += 64) {
        const double t = lerp_at(a.t_s, a.t_e, k);
        const double q = fmax(lerp_at(a.q_s, a.q_e, k), 0.0);
        const double p = a.p_hpa[k] * 100.0;
        const bool below = p >= p_lcl;
        const double tp = below ? th0 * pow(p / kP0, kKappa) : lift_moist(the, p);
        const double qp = below ? q0 : qsat(tp, p);
        const double tve = tv(t, q), tvp = tv(tp, qp);
        s_t[wave][k] = t;
        s_lp[wave][k] = log(fmax(p, kPmin));
        s_b[wave][k] = (tvp - tve) / fmax(tve, kTmin);
        s_d[wave][k] = tvp - tve;
        s_below[wave][k] = below ? 1 : 0;
    }
    // the wave's level rows are in LDS before lane 0 reads them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane != 0) return;
    int lcl = -1;
    for (int k = 0; k < a.nz; ++k) {
        if (!s_below[wave][k]) {
            lcl = k;
            break;
        }
    }
    int lnb = -1;
    double cape = 0.0, dtv = -1.0;
    if (lcl >= 0) {
        if (lcl + 1 < a.nz) dtv = s_d[wave][lcl + 1];
        // sum over the positive-buoyancy layers from the LCL up, dz from the
        // hypsometric equation on the environment temperature
        for (int k = (lcl > 1 ? lcl : 1); k < a.nz; ++k) {
            const double dz = kRd * 0.5 * (s_t[wave][k - 1] + s_t[wave][k]) * (s_lp[wave][k - 1] - s_lp[wave][k]) / kG;
            cape += kG * fmax(s_b[wave][k], 0.0) * dz;
        }
        lnb = a.nz - 1;
        for (int k = lcl + 1; k < a.nz; ++k) {
            if (s_b[wave][k] <= 0.0) {
                lnb = k;
                break;
            }
        }
    }
    if (a.i_lcl) a.i_lcl[col] = lcl;
    if (a.i_lnb) a.i_lnb[col] = lnb;
    if (a.cape) a.cape[col] = cape;
    if (a.dtv) a.dtv[col] = dtv;
    if (a.conv) {
        a.conv[col] = static_cast<float>(cape);
        a.conv[a.ncol + col] = lcl >= 0 ? a.lev[lcl] : __builtin_huge_valf();
        a.conv[2 * a.ncol + col] = lnb >= 0 ? a.lev[lnb] : __builtin_huge_valf();
        a.conv[3 * a.ncol + col] = static_cast<float>(dtv);
    }
}

// Layer masses |dp|/g with half layers at the ends (emanuel.py:263-279).
__global__ void layer_mass_kernel(const double* p_hpa, int nz, float* lm) {
    const int k = threadIdx.x;
    if (k >= nz) return;
    auto p = [&](int i) { return static_cast<float>(p_hpa[i]) * 100.0f; };
    float dp;
    if (k == 0) {
        dp = fabsf(p(0) - p(1)) * 0.5f;
    } else if (k == nz - 1) {
        dp = fabsf(p(nz - 2) - p(nz - 1)) * 0.5f;
    } else {
        dp = 0.5f * fabsf(p(k - 1) - p(k + 1));
    }
    lm[k] = dp / 9.80665f;
}

// Structured column of one cloud: e (entrainment, BL), d (detrainment, cloud),
// phi (subsidence), CFL scale.
struct Cloud {
    int lcl, lnb;
    float mb, bl_total, w_total, scale;
};

__device__ __forceinline__ float cloud_w(const Cloud& c, int k) {
    if (k < c.lcl || k > c.lnb) return 0.0f;
    return 1.0f - static_cast<float>(k - c.lcl) / static_cast<float>(c.lnb - c.lcl);
}
__device__ __forceinline__ float cloud_e(const Cloud& c, const float* m, int i) {
    return i < c.lcl ? c.mb * m[i] / c.bl_total : 0.0f;
}
__device__ __forceinline__ float cloud_d(const Cloud& c, int k) { return c.mb * cloud_w(c, k) / c.w_total; }

__device__ Cloud make_cloud(int lcl, int lnb, float mb, const float* m, int nz) {
    Cloud c{lcl, lnb, mb, 0.0f, 0.0f, 1.0f};
    float bl = 0.0f;
    for (int i = 0; i < lcl; ++i) bl += m[i];
    c.bl_total = fmaxf(bl, 1e-12f);
    float wt = 0.0f;
    for (int k = lcl; k <= lnb; ++k) wt += cloud_w(c, k);
    c.w_total = fmaxf(wt, 1e-12f);
    float sum_d = 0.0f;
    for (int k = 0; k < nz; ++k) sum_d += cloud_d(c, k);
    // CFL cap on the per-layer outflux fraction (emanuel.py:500-504); row i of
    // the unscaled matrix sums to e_i * sum(d) / mb + phi_{i-1}.
    float ce = 0.0f, cd = 0.0f, phi_prev = 0.0f, mx = 0.0f;
    for (int i = 0; i < nz; ++i) {
        const float row = cloud_e(c, m, i) * sum_d / mb + phi_prev;
        mx = fmaxf(mx, row / fmaxf(m[i], 1e-12f));
        ce += cloud_e(c, m, i);
        cd += cloud_d(c, i);
        phi_prev = (i < nz - 1) ? fmaxf(ce - cd, 0.0f) : 0.0f;
    }
    if (mx > 0.9f) c.scale = 0.9f / mx;
    return c;
}

struct ConvectArgs {
    GlideStepParams p;
    GlideParticles s;
    float lon0, lat0, sx, sy;
    float cdlon, cdlat;  // grid spacing (cell positions: the cell's point is lon0 + cx * cdlon)
    int nx, ny, nz;
    const float* levels;
    const int32_t* i_lcl;
    const int32_t* i_lnb;
    const float* m_b;
    const float* lm;
    const float4* cloud;  // per column: bl_total, w_total, scale, sum of d (NULL: derive per particle)
    FpDev fp;
    GlideEmit e;
};

// Per column: the cloud totals every particle of the column needs (the same
// make_cloud and detrainment sum the particles would each repeat).
__global__ void cloud_prepare_kernel(const int32_t* i_lcl, const int32_t* i_lnb, const float* m_b, const float* lm,
                                     int nz, int64_t ncol, float4* cloud) {
    const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int lcl = i_lcl[col], lnb = i_lnb[col];
    const float mb = m_b[col];
    float4 v = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    if (lcl >= 0 && lnb > lcl && mb > 0.0f) {
        const Cloud c = make_cloud(lcl, lnb, mb, lm, nz);
        float sum_d = 0.0f;
        for (int k = 0; k < nz; ++k) sum_d += cloud_d(c, k);
        v = make_float4(c.bl_total, c.w_total, c.scale, sum_d);
    }
    cloud[col] = v;
}

// Convection of particle i. With GLIDE_F_FIXED its footprint key and
// fixed-point value go to fkey/fval (the caller's dropped key and 0 stay for an
// inactive particle) instead of memory.
__device__ __forceinline__ void convect_particle(const ConvectArgs& a, const float* lev, const float* lm,
                                                 int64_t i, uint64_t& fkey, long long& fval) {
    const GlideParticles& s = a.s;
    const GlideStepParams& p = a.p;
    const bool emit = (p.flags & GLIDE_F_EMIT) != 0;
    const bool fixed = (p.flags & GLIDE_F_FIXED) != 0;
    const float co = static_cast<float>(p.cursor_off);
    bool active;
    if (p.flags & GLIDE_F_MASK_INPUT) {
        active = s.active[i] != 0;
    } else {
        active = in_window(s, static_cast<uint32_t>(i), co, static_cast<float>(p.length_s), p.flags);
    }
    if (!active) {
        if (emit && !fixed) {
            a.e.keys[i] = a.fp.n_keys;
            a.e.vals[i] = 0.0f;
        }
        return;
    }
    const int64_t ps = s.pos_stride;
    float lon = s.lon[i * ps];
    float lat = s.lat[i * ps];
    float z = s.z[i * ps];
    const int nz = a.nz;
    // Nearest column (cell positions: from the cell's point and the offsets,
    // then the absolute coordinates for the footprint and domain formulas).
    float gx, gy;
    if (s.cell != nullptr) {
        const int32_t c = s.cell[i];
        const int pcy = c >> 16, pcx = c & 0xffff;
        gx = cell_index(pcx, lon, a.sx, a.nx);
        gy = cell_index(pcy, lat, a.sy, a.ny);
        lon = __builtin_fmaf(static_cast<float>(pcx), a.cdlon, a.lon0) + lon;
        lat = __builtin_fmaf(static_cast<float>(pcy), a.cdlat, a.lat0) + lat;
    } else {
        gx = frac_index(lon, a.lon0, a.sx, a.nx);
        gy = frac_index(lat, a.lat0, a.sy, a.ny);
    }
    const int cx = static_cast<int>(floorf(gx + 0.5f));
    const int cy = static_cast<int>(floorf(gy + 0.5f));
    const int64_t col = static_cast<int64_t>(min(cy, a.ny - 1)) * a.nx + min(cx, a.nx - 1);
    const int lcl = a.i_lcl[col];
    const int lnb = a.i_lnb[col];
    const float mb = a.m_b[col];
    if (lcl >= 0 && lnb > lcl && mb > 0.0f) {
        // Host = nearest level (emanuel.py:765-771).
        int cnt = 0;
        for (int k = 0; k < nz; ++k) cnt += (lev[k] < z) ? 1 : 0;
        const int up = min(max(cnt, 1), nz - 1);
        const int lo = up - 1;
        const int host = ((z - lev[lo]) < (lev[up] - z)) ? lo : up;
        // Column `host` of fmassfrac, over m[host] (emanuel.py:756-779).
        Cloud c;
        float sum_d = 0.0f;
        if (a.cloud != nullptr) {
            const float4 cv = a.cloud[col];
            c = Cloud{lcl, lnb, mb, cv.x, cv.y, cv.z};
            sum_d = cv.w;
        } else {
            c = make_cloud(lcl, lnb, mb, lm, nz);
            for (int k = 0; k < nz; ++k) sum_d += cloud_d(c, k);
        }
        const float d_h = cloud_d(c, host);
        float ce = 0.0f, cd = 0.0f;
        for (int k = 0; k <= host && k < nz; ++k) {
            ce += cloud_e(c, lm, k);
            cd += cloud_d(c, k);
        }
        const float phi_h = (host < nz - 1) ? fmaxf(ce - cd, 0.0f) : 0.0f;  // f[host+1][host]
        float phi_hm1 = 0.0f;  // f[host][host-1]
        if (host >= 1) {
            const float ce1 = ce - cloud_e(c, lm, host), cd1 = cd - cloud_d(c, host);
            phi_hm1 = fmaxf(ce1 - cd1, 0.0f);
        }
        const float row_h = c.scale * (cloud_e(c, lm, host) * sum_d / mb + phi_hm1);
        const float lmh = fmaxf(lm[host], 1e-6f);
        auto colv = [&](int j) {
            float f = c.scale * (cloud_e(c, lm, j) * d_h / mb + ((j == host + 1) ? phi_h : 0.0f));
            if (j == host) f += fmaxf(lmh - row_h, 0.0f);
            return fmaxf(f / lmh, 0.0f);
        };
        float tot = 0.0f;
        for (int j = 0; j < nz; ++j) tot += colv(j);
        tot = fmaxf(tot, 1e-12f);
        const uint32_t uid = s.uid ? s.uid[i] : static_cast<uint32_t>(i);
        const U4 r = philox4x32_10(uid, static_cast<uint32_t>(p.step_index), kStreamConvection, p.noise_hi,
                                   static_cast<uint32_t>(p.seed), static_cast<uint32_t>(p.seed >> 32));
        const float u = u01(r.x);
        float cum = 0.0f;
        int dest = nz - 1;
        for (int j = 0; j < nz; ++j) {
            cum += colv(j) / tot;
            if (cum >= u) {
                dest = j;
                break;
            }
        }
        if (dest != host) {
            const float zl = lev[dest > 0 ? dest - 1 : 0];
            const float zh = lev[dest];
            z = zl + u01(r.y) * (zh - zl);
            s.z[i * ps] = z;
        }
    }
    if (emit) {
        const int64_t tb = time_bin(s.we_off[i], co, a.fp.n_t);
        const int64_t rb = s.rel_idx ? s.rel_idx[i] : 0;
        const float val = s.weight[i] * static_cast<float>(p.dt);
        if (fixed) {
            fkey = fixed_cell(a.fp, lon, lat, z, tb, rb);
            fval = fkey != kNoCell ? to_fixed(val, a.e.fixed_exp[rb]) : 0;
        } else {
            a.e.keys[i] = footprint_key(a.fp, true, lon, lat, z, tb, rb);
            a.e.vals[i] = val;
        }
        const bool esc = !within_domain(a.fp, lon, lat, z);
        if (esc) s.alive[i] = 0;
        count_flags(esc, a.e.escaped);
    }
}

__global__ __launch_bounds__(256) void convect_kernel(const ConvectArgs a) {
    __shared__ float lev[kConvLevels];
    __shared__ float lm[kConvLevels];
    for (int k = threadIdx.x; k < a.nz; k += blockDim.x) {
        lev[k] = a.levels[k];
        lm[k] = a.lm[k];
    }
    __syncthreads();
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t fkey = kNoCell;
    long long fval = 0;
    if (i < a.s.n) convect_particle(a, lev, lm, i, fkey, fval);
    if ((a.p.flags & (GLIDE_F_EMIT | GLIDE_F_FIXED)) == (GLIDE_F_EMIT | GLIDE_F_FIXED))
        wave_fixed_add(fkey, fval, fixed_copy(a.e, a.fp.acc_cells));  // every lane is here
}

// Mass conservation per column in float64: the float entries
// f[i][j] = scale (e_i d_j / M_b + [i == j+1] phi_j), summed in float64 along
// rows and columns, give max_i |row_i - col_i| / m_i (non-divergence,
// emanuel.py:437-446). One wavefront per column. The cloud profile is the one
// make_cloud builds (same per-level terms, sums in the same order), with the
// per-level terms spread over the lanes and the running sums on lane 0.
constexpr int kColsPerBlock = 4;
constexpr int kEntryCap = 1024;  // cloud-block entries staged per column (rows below base x cloud levels)

__global__ __launch_bounds__(64 * kColsPerBlock) void colsum_check_kernel(const int32_t* i_lcl, const int32_t* i_lnb,
                                                                         const float* m_b, const float* lm_g, int nz,
                                                                         int64_t ncol, double* err, float4* cloud_out) {
    __shared__ float lm[kConvLevels];
    __shared__ float es[kColsPerBlock][kConvLevels];
    __shared__ float dsh[kColsPerBlock][kConvLevels];
    __shared__ float ph[kColsPerBlock][kConvLevels];
    __shared__ float rowv[kColsPerBlock][kConvLevels];
    __shared__ float ent[kColsPerBlock][kEntryCap];
    __shared__ float totals[kColsPerBlock][4];  // bl_total, w_total, sum_d, phi above cloud top
    for (int k = threadIdx.x; k < nz; k += blockDim.x) lm[k] = lm_g[k];
    // the wavefront's column (uniform: its descriptors and loop bounds are scalar)
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * kColsPerBlock + wave;
    const bool have = col < ncol;
    int lcl = -1, lnb = -1;
    float mb = 0.0f;
    if (have) {
        lcl = i_lcl[col];
        lnb = i_lnb[col];
        mb = m_b[col];
    }
    const bool cloud = have && lcl >= 0 && lnb > lcl && mb > 0.0f;
    Cloud c{lcl, lnb, mb, 0.0f, 0.0f, 1.0f};
    float* e = es[wave];
    float* d = dsh[wave];
    float* phi = ph[wave];
    float* rv = rowv[wave];
    if (cloud) {
        for (int k = lane; k < nz; k += 64) d[k] = cloud_w(c, k);  // detrainment shape first
    }
    __syncthreads();
    if (cloud && lane == 0) {
        float bl = 0.0f;
#pragma unroll 8
        for (int i = 0; i < lcl; ++i) bl += lm[i];
        float wt = 0.0f;
#pragma unroll 8
        for (int k = lcl; k <= lnb; ++k) wt += d[k];
        totals[wave][0] = fmaxf(bl, 1e-12f);
        totals[wave][1] = fmaxf(wt, 1e-12f);
    }
    __syncthreads();
    if (cloud) {
        c.bl_total = totals[wave][0];
        c.w_total = totals[wave][1];
        for (int k = lane; k < nz; k += 64) {
            e[k] = cloud_e(c, lm, k);
            d[k] = cloud_d(c, k);
        }
    }
    __syncthreads();
    if (cloud && lane == 0) {
        // d is +0 outside [lcl, lnb], and adding +0 leaves the sum as it is
        float sum_d = 0.0f;
#pragma unroll 8
        for (int k = lcl; k <= lnb; ++k) sum_d += d[k];
        totals[wave][2] = sum_d;
    }
    __syncthreads();
    const float sum_d = cloud ? totals[wave][2] : 0.0f;
    if (cloud) {
        for (int k = lane; k < nz; k += 64) rv[k] = e[k] * sum_d / mb;
    }
    __syncthreads();
    if (cloud && lane == 0) {
        // Above cloud top e and d are +0, so the running sums stop changing and
        // every level there has phi = max(sum e - sum d, 0) (0 on the top
        // level): the walk ends at lnb and the lanes fill in the rest.
        float ce = 0.0f, cd = 0.0f, phi_prev = 0.0f;
#pragma unroll 8
        for (int i = 0; i <= lnb; ++i) {
            rv[i] = rv[i] + phi_prev;  // unscaled row sum of row i
            ce += e[i];
            cd += d[i];
            phi_prev = (i < nz - 1) ? fmaxf(ce - cd, 0.0f) : 0.0f;
            phi[i] = phi_prev;
        }
        totals[wave][3] = fmaxf(ce - cd, 0.0f);
    }
    __syncthreads();
    if (cloud) {
        const float p_top = totals[wave][3];
        for (int i = lane; i < nz; i += 64) {
            if (i > lnb) {
                rv[i] = rv[i] + p_top;  // phi_{i-1}: level lnb's or the constant above it
                phi[i] = (i < nz - 1) ? p_top : 0.0f;
            }
        }
    }
    // e_i is zero at and above cloud base (i >= lcl) and d_j outside
    // [lcl, lnb], so f[i][j] is +0 except for i < lcl with lcl <= j <= lnb
    // and on the sub-diagonal j = i - 1. Each lane visits only those entries,
    // in ascending j as a full sweep would; adding +0 leaves a float64 sum of
    // non-negative terms unchanged, so both sums are the full sweep's bit for
    // bit. The block i < lcl, lcl <= j <= lnb is evaluated once, spread over
    // the lanes, when it fits the staging area.
    const int nb = lnb - lcl + 1;
    const bool staged = cloud && lcl * nb <= kEntryCap;
    float* en = ent[wave];
    const auto f = [&](int r, int q) { return c.scale * (e[r] * d[q] / mb + ((r == q + 1) ? phi[q] : 0.0f)); };
    if (cloud) {
        // CFL cap on the per-layer outflux fraction (emanuel.py:500-504).
        float mx = 0.0f;
        for (int i = lane; i < nz; i += 64) mx = fmaxf(mx, rv[i] / fmaxf(lm[i], 1e-12f));
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
        if (mx > 0.9f) c.scale = 0.9f / mx;
        if (staged) {
            // entry t = i * nb + jj on lane t mod 64, (i, jj) stepped without a division per entry
            const int n_ent = lcl * nb;
            const int qs = 64 / nb, rs = 64 % nb;
            int i = lane / nb, jj = lane % nb;
            for (int t = lane; t < n_ent; t += 64) {
                en[t] = f(i, lcl + jj);
                i += qs;
                jj += rs;
                if (jj >= nb) {
                    jj -= nb;
                    ++i;
                }
            }
        }
    }
    __syncthreads();
    double worst = 0.0;
    if (cloud) {
        for (int i = lane; i < nz; i += 64) {
            double row = 0.0, cols = 0.0;
            if (i >= 1) row += static_cast<double>(f(i, i - 1));  // i - 1 < lcl whenever i < lcl
            if (staged) {
                if (i < lcl)
                    for (int jj = 0; jj < nb; ++jj) row += static_cast<double>(en[i * nb + jj]);
                if (i >= lcl && i <= lnb)
                    for (int j = 0; j < lcl; ++j) cols += static_cast<double>(en[j * nb + i - lcl]);
            } else {
                if (i < lcl)
                    for (int j = lcl; j <= lnb; ++j) row += static_cast<double>(f(i, j));
                if (i >= lcl && i <= lnb)
                    for (int j = 0; j < lcl; ++j) cols += static_cast<double>(f(j, i));
            }
            if (i + 1 < nz) cols += static_cast<double>(f(i + 1, i));  // i + 1 > lcl - 1 whenever i >= lcl
            const double r = fabs(row - cols) / fmax(static_cast<double>(lm[i]), 1e-12);
            worst = r > worst ? r : worst;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(worst, off, 64);
        worst = o > worst ? o : worst;
    }
    if (have && lane == 0) {
        err[col] = worst;
        if (cloud_out != nullptr)
            cloud_out[col] = cloud ? make_float4(c.bl_total, c.w_total, c.scale, sum_d)
                                   : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    }
}

}  // namespace glide

using namespace glide;

extern "C" int glide_convect_columns(const float* conv, const float* t_s, const float* t_e, double alpha,
                                     const double* p_level_hpa, const float* levels, int32_t nz, int32_t ny,
                                     int32_t nx, double closure_c, double trigger_dtv_k, double min_cape_j_kg,
                                     double min_cloud_depth_m, int32_t* i_lcl, int32_t* i_lnb, float* m_b,
                                     float* layer_mass, void* stream) {
    GLIDE_REQUIRE(conv && t_s && t_e && p_level_hpa && levels && i_lcl && i_lnb && m_b && layer_mass,
                  "glide_convect_columns: NULL argument");
    GLIDE_REQUIRE(nz >= 2 && nz <= kConvLevels && ny >= 1 && nx >= 1, "bad convection dimensions");
    GLIDE_REQUIRE(closure_c > 0, "closure_c must be > 0");
    ColumnArgs a{conv, t_s, t_e, static_cast<float>(alpha), p_level_hpa, levels, nz, ny, nx,
                 static_cast<float>(closure_c), static_cast<float>(trigger_dtv_k),
                 static_cast<float>(min_cape_j_kg), static_cast<float>(min_cloud_depth_m), i_lcl, i_lnb, m_b};
    const int64_t ncol = static_cast<int64_t>(ny) * nx;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(convect_columns_kernel, dim3(static_cast<unsigned>((ncol + 255) / 256)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(layer_mass_kernel, dim3(1), dim3(kConvLevels), 0, st, p_level_hpa, nz, layer_mass);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_column_cape(const float* t_s, const float* t_e, const float* q_s, const float* q_e, double alpha,
                                 const double* p_level_hpa, const float* levels, int32_t nz, int32_t ny, int32_t nx,
                                 int32_t* i_lcl, int32_t* i_lnb, double* cape, double* dtv, float* conv,
                                 void* stream) {
    GLIDE_REQUIRE(t_s && t_e && q_s && q_e && p_level_hpa, "glide_column_cape: NULL input");
    GLIDE_REQUIRE(i_lcl || i_lnb || cape || dtv || conv, "glide_column_cape: no output");
    GLIDE_REQUIRE(conv == nullptr || levels != nullptr, "glide_column_cape: conv needs levels");
    GLIDE_REQUIRE(nz >= 2 && nz <= kConvLevels && ny >= 1 && nx >= 1, "bad column dimensions");
    GLIDE_REQUIRE(alpha >= 0.0 && alpha <= 1.0, "alpha must be in [0, 1]");
    CapeArgs a{t_s, t_e, q_s, q_e, alpha, p_level_hpa, levels, nz, static_cast<int64_t>(ny) * nx,
               i_lcl, i_lnb, cape, dtv, conv};
    const int64_t blocks = (a.ncol + kCapeWaves - 1) / kCapeWaves;
    hipLaunchKernelGGL(column_cape_kernel, dim3(static_cast<unsigned>(blocks)), dim3(64 * kCapeWaves), 0,
                       static_cast<hipStream_t>(stream), a);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_convect(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
                             const int32_t* i_lcl, const int32_t* i_lnb, const float* m_b,
                             const float* layer_mass, double* colsum_err, float* cloud,
                             const GlideFootprintGrid* fg, const GlideEmit* e, void* stream) {
    GLIDE_REQUIRE(p && s && g && i_lcl && i_lnb && m_b && layer_mass, "glide_convect: NULL argument");
    GLIDE_REQUIRE(g->nz >= 2 && g->nz <= kConvLevels && g->nx >= 2 && g->ny >= 2, "bad grid");
    GLIDE_REQUIRE(g->lon1 != g->lon0 && g->lat1 != g->lat0, "Grid dimension has 0 span (first == last)");
    GLIDE_REQUIRE(s->pos_stride >= 1 && s->n >= 0, "bad particles");
    if (p->flags & GLIDE_F_MASK_INPUT) {
        GLIDE_REQUIRE(s->active, "active mask is NULL");
    } else {
        GLIDE_REQUIRE(s->r_off && s->alive, "r_off/alive are NULL (no explicit mask)");
        GLIDE_REQUIRE(!(p->flags & GLIDE_F_RELEASE_TERMINUS) || s->lo_off, "GLIDE_F_RELEASE_TERMINUS needs lo_off");
    }
    ConvectArgs a{};
    a.p = *p;
    a.s = *s;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.lon0 = static_cast<float>(g->lon0);
    a.lat0 = static_cast<float>(g->lat0);
    a.sx = static_cast<float>((g->nx - 1) / (g->lon1 - g->lon0));
    a.sy = static_cast<float>((g->ny - 1) / (g->lat1 - g->lat0));
    a.cdlon = static_cast<float>((g->lon1 - g->lon0) / (g->nx - 1));
    a.cdlat = static_cast<float>((g->lat1 - g->lat0) / (g->ny - 1));
    GLIDE_REQUIRE(s->cell == nullptr || (s->pos_stride == 1 && g->nx <= 65535 && g->ny <= 65535),
                  "cell positions need pos_stride 1 and nx, ny <= 65535");
    a.levels = g->levels;
    a.i_lcl = i_lcl;
    a.i_lnb = i_lnb;
    a.m_b = m_b;
    a.lm = layer_mass;
    if (p->flags & GLIDE_F_EMIT) {
        if (p->flags & GLIDE_F_FIXED) {
            GLIDE_REQUIRE(fg && e && e->fixed && e->fixed_exp && e->escaped, "fixed-point emit outputs are NULL");
            GLIDE_REQUIRE(fg->t_bin == nullptr, "GLIDE_F_FIXED takes flat keys (t_bin NULL)");
            GLIDE_REQUIRE(e->fixed_copies >= 0 && e->fixed_copies <= 1024, "fixed_copies must be in [0, 1024]");
        } else {
            GLIDE_REQUIRE(fg && e && e->keys && e->vals && e->escaped, "emit outputs are NULL");
        }
        GLIDE_REQUIRE(s->we_off && s->weight && s->alive, "we_off/weight/alive needed to emit");
        const uint64_t cells = static_cast<uint64_t>(fg->n_r) * fg->n_t * fg->n_z * fg->n_y * fg->n_x;
        GLIDE_REQUIRE(cells > 0, "bad footprint size");
        if (p->flags & GLIDE_F_FIXED) {
            GLIDE_REQUIRE(fg->slab_off == nullptr || (fg->slab_nt != nullptr && fg->slab_cells > 0),
                          "per-release slabs need slab_nt and slab_cells");
        } else {
            GLIDE_REQUIRE(cells < 0xffffffffull, "footprint has too many cells for 32-bit keys");
            GLIDE_REQUIRE(fg->slab_off == nullptr, "per-release slabs need GLIDE_F_FIXED");
        }
        a.fp = make_fp(*fg);
        a.e = *e;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t ncol = static_cast<int64_t>(g->ny) * g->nx;
    float4* cloud4 = (cloud != nullptr && s->n > 0) ? reinterpret_cast<float4*>(cloud) : nullptr;
    if (colsum_err != nullptr) {
        // the check derives each column's cloud totals on the way (bit for bit
        // make_cloud's), so it also writes the per-column records
        hipLaunchKernelGGL(colsum_check_kernel, dim3(static_cast<unsigned>((ncol + kColsPerBlock - 1) / kColsPerBlock)),
                           dim3(64 * kColsPerBlock), 0, st, i_lcl, i_lnb, m_b, layer_mass, g->nz, ncol, colsum_err,
                           cloud4);
    } else if (cloud4 != nullptr) {
        hipLaunchKernelGGL(cloud_prepare_kernel, dim3(static_cast<unsigned>((ncol + 255) / 256)), dim3(256), 0, st,
                           i_lcl, i_lnb, m_b, layer_mass, g->nz, ncol, cloud4);
    }
    a.cloud = cloud4;
    if (s->n > 0) {
        hipLaunchKernelGGL(convect_kernel, dim3(static_cast<unsigned>((s->n + 255) / 256)), dim3(256), 0, st, a);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
