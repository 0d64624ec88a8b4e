// Fused per-particle backward step (one thread per particle).
//
// Restates the reference's static step core in one pass over the particle:
// RK2 advection through the time-interpolated trilinear winds
// (hanna.py:855-896), surface bilinear samples and stability (hanna.py:955-967),
// sigma/T_L at z_eval and the +-1 m probes (hanna.py:983-991, 1500-1596), the
// density drift (hanna.py:993-1006), the substepped exact-OU update with
// clipping, displacement and smooth-wall reflection (hanna.py:1224-1368),
// meander (hanna.py:1028-1054) and the active-mask gate (hanna.py:1056-1068).
// With GLIDE_F_EMIT the same pass writes the footprint key/value of the new
// position and applies the escape kill (main.py:1459-1477).
//
// Inactive particles are not touched at all (equivalent to the reference's
// torch.where gate). Each active particle runs only its own k substeps: with
// the counter-based noise a sub_dt = 0 substep is the identity, so this equals
// the fixed max_substeps loop (hanna.py:1245-1254).
#include <hip/hip_ext.h>

#include <vector>

#include "footprint_key.h"
#include "physics.h"

namespace glide {

struct StepArgs {
    GlideStepParams p;
    GlideParticles s;
    float lon0, lat0, sx, sy;  // met grid: ix = (lon - lon0) * sx
    float cdlon, cdlat;  // grid spacing (GlideParticles.cell: the cell's point is lon0 + cx * cdlon)
    float blon0, blat0, bsx, bsy;  // surface bilinear mapping (same grid)
    int nx, ny, nz;
    int ltop;  // first step of the level search (ladder_top(nz))
    uint32_t planeb;  // float32 window: byte offset of plane B (ncell * 16)
    const float* levels;
    const void* wind;
    const void* aux;
    const void* surf;
    FpDev fp;
    GlideEmit e;
    // Table mode: the step's scalars come from table[row[0]] on the device, the
    // escape count goes to e.escaped[row[0]], and the last workgroup to finish
    // advances row[0] (row[1..] count finished workgroups).
    const GlideStepParams* table;
    int32_t n_rows;  // rows of the table; a cursor outside [0, n_rows) steps nothing
    int32_t* row;
    unsigned long long* n_active;  // optional per-step active count (e->active)
    // Served slots (glide_step_slots): per-wavefront slot, scalars and window.
    GlideSlot* slots;
    const int32_t* wave_slot;
    GlideMetRing ring;
};

constexpr int kMaxLevels = kLadder;
constexpr int kStepBlock = 256;     // default threads per workgroup
constexpr int kStepBlockMax = 512;  // GlideStepParams.block may choose 64 .. 512
constexpr uint32_t kRowGroups = GLIDE_ROW_WORDS - 2;  // finished-workgroup counters of the step cursor

// Row r of the table, read into scalar registers (one readfirstlane per
// 32-bit word), so everything derived from it stays scalar.
__device__ __forceinline__ GlideStepParams load_row(const GlideStepParams* table, int r) {
    static_assert(sizeof(GlideStepParams) % 4 == 0, "GlideStepParams is a whole number of words");
    constexpr int kWords = sizeof(GlideStepParams) / 4;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(table + r);
    uint32_t buf[kWords];
#pragma unroll
    for (int k = 0; k < kWords; ++k) buf[k] = __builtin_amdgcn_readfirstlane(w[k]);
    GlideStepParams p;
    __builtin_memcpy(&p, buf, sizeof(p));
    return p;
}

// Per-launch (or, for served slots, per-wavefront) inputs of the step body.
struct StepCtx {
    const void* wind;
    const void* aux;
    const void* surf;
    unsigned long long* escaped;
    int32_t slot;  // >= 0: footprint release index of a served slot; -1: rel_idx
    uint32_t noise_hi;  // Philox counter word 3 (a served slot's; 0 otherwise)
};

// The step's scalars in the precision the particle update uses, 32-bit
// words only: a table row or served slot read from memory then holds a dozen
// scalar registers instead of the row's doubles (the conversions are the ones
// each use would make, so results do not change).
struct StepScalars {
    float dt, al, co, len, subc, zu, tm;
    int32_t kmax, flags;
    uint32_t k0, k1, step;
};
__device__ __forceinline__ StepScalars step_scalars(const GlideStepParams& p) {
    StepScalars q;
    q.dt = static_cast<float>(p.dt);
    q.al = static_cast<float>(p.alpha);
    q.co = static_cast<float>(p.cursor_off);
    q.len = static_cast<float>(p.length_s);
    q.subc = static_cast<float>(p.substep_c);
    q.zu = static_cast<float>(p.z_ubl_m);
    q.tm = static_cast<float>(p.meander_timescale_s);
    q.kmax = p.max_substeps;
    q.flags = p.flags;
    q.k0 = static_cast<uint32_t>(p.seed);
    q.k1 = static_cast<uint32_t>(p.seed >> 32);
    q.step = static_cast<uint32_t>(p.step_index);
    return q;
}

// Advances particle i; returns whether it was active. With GLIDE_F_FIXED the
// footprint key and fixed-point value go to fkey/fval (left as the caller set
// them, the dropped key and 0, for an inactive particle) instead of memory.
// CELL: the precision layout (GlideParticles.cell): lon and lat below are the
// offsets from the particle's cell point (rlon, rlat); every displacement
// moves them, and the absolute coordinate rlon + lon is formed only where a
// formula needs it (map factors, Coriolis, footprint cell, domain test).
//
// RELOAD: the state stores and the footprint epilogue read their pointers from
// the kernel arguments afresh (through an opaque copy of the kernarg segment
// address), so those pointers are not held in scalar registers through the
// whole particle update. Without it the table-row, served-slot, cell and
// bfloat16 instantiations run out of scalar registers and spill about 200 of
// them to vector lanes; the float32 plain launch, which holds them without
// spilling, is 2-4 us faster without the reloads.
template <bool BF16, bool FIXED, bool CELL, bool RELOAD>
__device__ __forceinline__ bool step_particle(const StepArgs& a, const float* lev, int64_t i,
                                              const StepScalars& p, const StepCtx& x, uint64_t& fkey,
                                              long long& fval) {
    // 32-bit byte offsets of this particle (the host checks the arrays fit)
    const uint32_t i1 = static_cast<uint32_t>(i);
    const uint32_t i4 = i1 * 4u;
    const uint32_t ip = static_cast<uint32_t>(i * a.s.pos_stride) * 4u;
    const GlideParticles& s = a.s;
    const bool emit = (p.flags & GLIDE_F_EMIT) != 0;
    constexpr bool fixed = FIXED;
    const float co = p.co;

    // The position loads go out with the mask's: one memory round trip before
    // the first gather instead of two (i < n, so they are in bounds).
    float lon = ldo(s.lon, ip);
    float lat = ldo(s.lat, ip);
    float z = ldo(s.z, ip);
    bool active;
    if (p.flags & GLIDE_F_MASK_INPUT) {
        active = ldo(s.active, i1) != 0;
    } else {
        active = in_window(s, i1, co, p.len, p.flags);
    }
    if (!active) {
        if (emit && !fixed) {
            sto(a.e.keys, i4, a.fp.n_keys);
            sto(a.e.vals, i4, 0.0f);
        }
        return false;
    }
    const float dt = p.dt;
    const float al = p.al;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    int cx = 0, cy = 0;
    float rlon = 0.0f, rlat = 0.0f;
    if constexpr (CELL) {
        const int32_t c = ldo(s.cell, i4);
        cy = c >> 16;
        cx = c & 0xffff;
        rlon = __builtin_fmaf(static_cast<float>(cx), a.cdlon, a.lon0);
        rlat = __builtin_fmaf(static_cast<float>(cy), a.cdlat, a.lat0);
    }
    // grid index and absolute coordinate of a position (an offset with CELL)
    auto gix = [&](float lo) { return CELL ? cell_index(cx, lo, a.sx, nx) : frac_index(lo, a.lon0, a.sx, nx); };
    auto giy = [&](float la) { return CELL ? cell_index(cy, la, a.sy, ny) : frac_index(la, a.lat0, a.sy, ny); };
    auto alat = [&](float la) { return CELL ? rlat + la : la; };

    // ---- RK2 advection (backward) ----
    // Level lookups after the first start from the last one's interval: the
    // heights looked up in a step are close together.
    int lvl = 1;
    bool first_lookup = true;
    auto wind_at = [&](float lo, float la, float zz, float& dlo, float& dla, float& dzz) {
        const float ix = gix(lo);
        const float iy = giy(la);
        const float iz = first_lookup ? level_index_up(lev, nz, a.ltop, zz, lvl)
                                      : level_index_near(lev, nz, a.ltop, zz, lvl);
        first_lookup = false;
        const Corner3 c = corners3(ix, iy, iz, nx, ny, nz, WinRec<BF16>::kCellUnit);
        float v[6];
        sample_wind<BF16>(x.wind, a.planeb, c, v);
        const float u = v[0] * (1.0f - al) + v[3] * al;
        const float vv = v[1] * (1.0f - al) + v[4] * al;
        const float w = v[2] * (1.0f - al) + v[5] * al;
        dlo = fdiv(u, m_per_deg_lon(alat(la)));
        dla = vv * (1.0f / kMPerDegLat);
        dzz = w;
    };
    {
        float d0, d1, d2;
        wind_at(lon, lat, z, d0, d1, d2);
        const float h = 0.5f * dt;  // float(0.5 * dt) exactly: halving commutes with the rounding
        float e0, e1, e2;
        wind_at(lon - h * d0, lat - h * d1, z - h * d2, e0, e1, e2);
        lon = lon - dt * e0;
        lat = lat - dt * e1;
        z = z - dt * e2;
    }

    float up = ldo(s.up, i4), vp = ldo(s.vp, i4), wp = ldo(s.wp, i4);
    const bool meander = (p.flags & GLIDE_F_MEANDER) != 0;
    float um = meander ? ldo(s.um, i4) : 0.0f, vm = meander ? ldo(s.vm, i4) : 0.0f;

    if (!(p.flags & GLIDE_F_NO_TURB)) {
        // ---- surface fields at the post-advection position ----
        Surface sf;
        float sp, shf, tsf;
        {
            const float ix = CELL ? cell_index(cx, lon, a.bsx, nx) : frac_index(lon, a.blon0, a.bsx, nx);
            const float iy = CELL ? cell_index(cy, lat, a.bsy, ny) : frac_index(lat, a.blat0, a.bsy, ny);
            const float fx0 = floorf(ix), fy0 = floorf(iy);
            const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
            const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1);
            const float fx = ix - fx0, fy = iy - fy0;
            const float w00 = (1.0f - fx) * (1.0f - fy), w01 = fx * (1.0f - fy);
            const float w10 = (1.0f - fx) * fy, w11 = fx * fy;
            float acc[5] = {0, 0, 0, 0, 0};
            constexpr uint32_t sb = kSurfBytes;
            const uint32_t r0 = static_cast<uint32_t>(y0 * nx), r1 = static_cast<uint32_t>(y1 * nx);
            const uint32_t offs[4] = {(r0 + x0) * sb, (r0 + x1) * sb, (r1 + x0) * sb, (r1 + x1) * sb};
            const float ws[4] = {w00, w01, w10, w11};
            float r[4][10];
#pragma unroll
            for (int q = 0; q < 4; ++q) load_surf_at(x.surf, offs[q], r[q]);
            // torch_lerp's two forms, chosen once: alpha is the same for the
            // whole launch
            if (fabsf(al) < 0.5f) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int ch = 0; ch < 5; ++ch) acc[ch] += ws[q] * (r[q][ch] + al * (r[q][5 + ch] - r[q][ch]));
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int ch = 0; ch < 5; ++ch)
                        acc[ch] += ws[q] * (r[q][5 + ch] - (r[q][5 + ch] - r[q][ch]) * (1.0f - al));
            }
            sf.blh = fmaxf(acc[0], kBlhMin);
            sp = acc[1];
            sf.ustar = fmaxf(acc[2], kUstarMin);
            shf = acc[3];
            tsf = acc[4];
        }
        float L, wstar;
        stability(sf.blh, sp, sf.ustar, shf, tsf, L, wstar);
        sf.L = L;
        sf.wstar = wstar;
        sf.hol = fdiv(sf.blh, L);
        sf.fcor = 2.0f * kOmega * fsin_deg(alat(lat));

        const bool tl_floors = (p.flags & GLIDE_F_TL_FLOORS) != 0;
        const bool prescribed = (p.flags & GLIDE_F_PRESCRIBED) != 0;  // the support record at every height
        const bool sl_over = (p.flags & GLIDE_F_SL_OVERRIDE) != 0;
        const float zu = p.zu;
        const float z_eval = fmaxf(z, zu);
        const float ix = gix(lon);
        const float iy = giy(lat);

        // support record at z_eval: FT pair, density pair (and, bfloat16, the
        // meander pair; float32 samples its plane after the substeps, when
        // it is needed, so its loads are not in flight with these)
        float ax[6];
        const float iz_eval = level_index_near(lev, nz, a.ltop, z_eval, lvl);
        {
            const Corner3 c = corners3(ix, iy, iz_eval, nx, ny, nz, WinRec<BF16>::kCellUnit);
            sample_aux_a<BF16>(x.aux, c, ax);
        }

        // Column turbulence at z_eval (hanna.py:1500-1596): above the
        // boundary layer (or prescribed) the support record's pair, else the
        // Hanna profiles (a lane takes one branch, not both).
        SigTL t;
        if (z_eval > sf.blh || prescribed) {
            t.su = t.sv = t.sw = ax[0];
            t.tu = t.tv = t.tw = ax[1];
        } else {
            t = in_bl(z_eval, sf);
            if (sl_over && z_eval < 0.1f * sf.blh) {
                const float z_sl = 0.1f * sf.blh;
                const SigTL tc = in_bl(fmaxf(z_eval, z_sl), sf);
                t.su = tc.su;
                t.sv = tc.sv;
                t.sw = surface_layer_sw(z_eval, sf.ustar, sf.L);
                const float zt = fmaxf(z_eval, kZMin);
                t.tu = kKarman * zt / fmaxf(t.su, kSigmaMin);
                t.tv = kKarman * zt / fmaxf(t.sv, kSigmaMin);
                t.tw = kKarman * zt / fmaxf(t.sw, kSigmaMin);
            }
            if (tl_floors) {
                t.tu = fmaxf(t.tu, 10.0f);
                t.tv = fmaxf(t.tv, 10.0f);
                t.tw = fmaxf(t.tw, 30.0f);
            }
        }
        t.su = fmaxf(t.su, kSigmaMin);
        t.sv = fmaxf(t.sv, kSigmaMin);
        t.sw = fmaxf(t.sw, kSigmaMin);
        t.tu = fmaxf(t.tu, kTLMin);
        t.tv = fmaxf(t.tv, kTLMin);
        t.tw = fmaxf(t.tw, kTLMin);

        // sigma_w at the +-1 m probes.
        // Free-troposphere sigma_w at both probes from one gather: z_lo's
        // stencil columns at its levels l, l+1, l+2 (clamped). z_hi, 2 m
        // above, falls in level l or l+1 unless the levels are closer than
        // that (then its own stencil is read). Same weights and order of
        // sums as a trilinear sample of each probe.
        const float z_hi = z_eval + 1.0f;
        const float z_lo = fmaxf(z_eval - 1.0f, zu);
        float ft_lo = 0.0f, ft_hi = 0.0f;
        if (z_hi > sf.blh || prescribed) {
            constexpr uint32_t ru = WinRec<BF16>::kCellUnit;
            int lp = lvl;
            const float izl = level_index_near(lev, nz, a.ltop, z_lo, lp);
            const float izh = level_index_near(lev, nz, a.ltop, z_hi, lp);
            const Corner3 cl = corners3(ix, iy, izl, nx, ny, nz, ru);
            const Corner3 ch = corners3(ix, iy, izh, nx, ny, nz, ru);
            const int zl0 = static_cast<int>(floorf(izl)), zh0 = static_cast<int>(floorf(izh));
            const uint32_t d1 = cl.o[4] - cl.o[0];  // level l to min(l + 1, nz - 1)
            const uint32_t d2 = zl0 + 2 <= nz - 1 ? ru : 0u;
            float sg[4][3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sg[q][0] = sample_sigma_ft_at<BF16>(x.aux, cl.o[q]);
                sg[q][1] = sample_sigma_ft_at<BF16>(x.aux, cl.o[q] + d1);
                sg[q][2] = sample_sigma_ft_at<BF16>(x.aux, cl.o[q] + d1 + d2);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) ft_lo += cl.w[q] * sg[q][0];
#pragma unroll
            for (int q = 0; q < 4; ++q) ft_lo += cl.w[q + 4] * sg[q][1];
            const int dh = zh0 - zl0;
            if (dh == 0 || dh == 1) {
#pragma unroll
                for (int q = 0; q < 4; ++q) ft_hi += ch.w[q] * (dh == 0 ? sg[q][0] : sg[q][1]);
#pragma unroll
                for (int q = 0; q < 4; ++q) ft_hi += ch.w[q + 4] * (dh == 0 ? sg[q][1] : sg[q][2]);
            } else {
                ft_hi = sample_sigma_ft<BF16>(x.aux, ch);
            }
        }
        auto sw_probe = [&](float zq, float ft) {
            float sw;
            if (zq > sf.blh || prescribed) {
                sw = ft;
            } else {
                sw = in_bl_sw(zq, sf);
                if (sl_over && zq < 0.1f * sf.blh) sw = surface_layer_sw(zq, sf.ustar, sf.L);
            }
            return fmaxf(sw, kSigmaMin);
        };
        const float sw_hi = sw_probe(z_hi, ft_hi);
        const float sw_lo = sw_probe(z_lo, ft_lo);
        const float dsig2 = fdiv(sw_hi * sw_hi - sw_lo * sw_lo, fmaxf(z_hi - z_lo, 1.0f));
        const float sw_sq = fmaxf(t.sw * t.sw, kSigmaMin * kSigmaMin);

        const float rho = fmaxf(ax[2], 1e-3f);
        const float dens_b = -fdiv(sw_sq * ax[3], rho);
        const float half_b = -0.5f * dsig2;

        // Substep schedule (hanna.py:1162-1186).
        // IEEE division: the ratio sits exactly on integers for the T_L floors.
        const float ratio = dt / (p.subc * fmaxf(t.tw, 1e-3f));
        // (a device table row's max_substeps below 1 still takes one substep)
        int k = max(p.kmax, 1);
        if (ratio < static_cast<float>(k)) k = max(static_cast<int>(ceilf(ratio)), 1);
        const float sdt = dt / static_cast<float>(k);

        const float su2 = t.su * t.su, sv2 = t.sv * t.sv, sw2 = t.sw * t.sw;
        const float umax = 4.0f * t.su, vmax = 4.0f * t.sv, wmax = 4.0f * t.sw;
        const uint32_t uid = s.uid ? ldo(s.uid, i4) : static_cast<uint32_t>(i);
        const uint32_t k0 = p.k0, k1 = p.k1;
        const uint32_t step = p.step;
        // sub_dt, sigma and T_L hold for every substep (hanna.py:1300-1320), so
        // the OU decay and noise amplitude are formed once per particle.
        const OuCoef cu = ou_coef(t.tu, su2, sdt), cv = ou_coef(t.tv, sv2, sdt), cw = ou_coef(t.tw, sw2, sdt);
        const float inv_sw_sq = __builtin_amdgcn_rcpf(sw_sq);
        for (int q = 0; q < k; ++q) {
            const U4 r = philox4x32_10(uid, step, static_cast<uint32_t>(q), x.noise_hi, k0, k1);
            float n0, n1, n2, n3;
            box_muller(u01(r.x), u01(r.y), n0, n1);
            box_muller(u01(r.z), u01(r.w), n2, n3);
            (void)n3;
            const float factor = 1.0f + (wp * wp) * inv_sw_sq;
            const float drift = factor * half_b + dens_b;
            float un = ou_apply(up, cu, n0, 0.0f);
            float vn = ou_apply(vp, cv, n1, 0.0f);
            float wn = ou_apply(wp, cw, n2, drift * sdt);
            // clamps to [-4 sigma, 4 sigma] (sigma > 0): one v_med3 each
            wn = __builtin_amdgcn_fmed3f(wn, -wmax, wmax);
            un = __builtin_amdgcn_fmed3f(un, -umax, umax);
            vn = __builtin_amdgcn_fmed3f(vn, -vmax, vmax);
            z = z + (-wn) * sdt;
            const float ml = m_per_deg_lon(alat(lat));
            lon = lon + fdiv((-un) * sdt, ml);
            lat = lat + (-vn) * sdt * (1.0f / kMPerDegLat);
            if (z < 0.0f) {
                z = 0.0f - z;
                wn = -wn;
            }
            up = un;
            vp = vn;
            wp = wn;
        }

        if (meander) {
            if constexpr (!BF16) {
                const Corner3 c = corners3(ix, iy, iz_eval, nx, ny, nz, WinRec<BF16>::kCellUnit);
                sample_aux_b(x.aux, a.planeb, c, ax[4], ax[5]);
            }
            const U4 r = philox4x32_10(uid, step, kStreamMeander, x.noise_hi, k0, k1);
            float n0, n1;
            box_muller(u01(r.x), u01(r.y), n0, n1);
            const float tm = p.tm;
            const OuCoef mu = ou_coef(tm, ax[4] * ax[4], dt);
            const OuCoef mv{mu.a, fsqrt(fmaxf(ax[5] * ax[5] * (1.0f - mu.a * mu.a), 0.0f))};  // same decay
            um = ou_apply(um, mu, n0, 0.0f);
            vm = ou_apply(vm, mv, n1, 0.0f);
            const float ml = m_per_deg_lon(alat(lat));
            lon = lon + fdiv((-um) * dt, ml);
            lat = lat + (-vm) * dt * (1.0f / kMPerDegLat);
        }
    }

    if constexpr (CELL) {
        // to the nearest grid point: the offsets stay within about half a cell
        const int ncx = min(max(static_cast<int>(floorf(static_cast<float>(cx) + lon * a.sx + 0.5f)), 0), nx - 1);
        const int ncy = min(max(static_cast<int>(floorf(static_cast<float>(cy) + lat * a.sy + 0.5f)), 0), ny - 1);
        lon = lon - static_cast<float>(ncx - cx) * a.cdlon;
        lat = lat - static_cast<float>(ncy - cy) * a.cdlat;
        cx = ncx;
        cy = ncy;
        rlon = __builtin_fmaf(static_cast<float>(cx), a.cdlon, a.lon0);
        rlat = __builtin_fmaf(static_cast<float>(cy), a.cdlat, a.lat0);
        sto(s.cell, i4, (cy << 16) | cx);
    }
    // The stores and the footprint epilogue, with their pointers from B: the
    // kernel arguments as passed (a) or read afresh (RELOAD).
#define GLIDE_STEP_FINISH(B)                                                                    \
    do {                                                                                        \
        sto(B.s.lon, ip, lon);                                                                  \
        sto(B.s.lat, ip, lat);                                                                  \
        sto(B.s.z, ip, z);                                                                      \
        sto(B.s.up, i4, up);                                                                    \
        sto(B.s.vp, i4, vp);                                                                    \
        sto(B.s.wp, i4, wp);                                                                    \
        if (meander) {                                                                          \
            sto(B.s.um, i4, um);                                                                \
            sto(B.s.vm, i4, vm);                                                                \
        }                                                                                       \
        if (emit) {                                                                             \
            if constexpr (CELL) { /* the absolute coordinates of the footprint and domain */    \
                lon = rlon + lon;                                                               \
                lat = rlat + lat;                                                               \
            }                                                                                   \
            const int64_t tb = time_bin(ldo(B.s.we_off, i4), co, a.fp.n_t);                     \
            const int64_t rb = x.slot >= 0 ? x.slot : (B.s.rel_idx ? ldo(B.s.rel_idx, i4) : 0); \
            const float val = ldo(B.s.weight, i4) * dt;                                         \
            if (fixed) {                                                                        \
                fkey = fixed_cell(a.fp, lon, lat, z, tb, rb);                                   \
                fval = fkey != kNoCell ? to_fixed(val, B.e.fixed_exp[rb]) : 0;                  \
            } else {                                                                            \
                sto(B.e.keys, i4, footprint_key(a.fp, true, lon, lat, z, tb, rb));              \
                sto(B.e.vals, i4, val);                                                         \
            }                                                                                   \
            const bool esc = !within_domain(a.fp, lon, lat, z);                                 \
            if (esc) sto(B.s.alive, i1, static_cast<uint8_t>(0));                               \
            count_flags(esc, x.escaped);                                                        \
        }                                                                                       \
    } while (0)
    if constexpr (RELOAD) {
        // the kernel's only argument sits at offset 0 of the kernarg segment;
        // the empty asm hides the address, so the loads happen here
        using KArgs = const __attribute__((address_space(4))) StepArgs;
        KArgs* kp = (KArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
        asm volatile("" : "+s"(kp));
        KArgs& b = *kp;
        GLIDE_STEP_FINISH(b);
    } else {
        GLIDE_STEP_FINISH(a);
    }
#undef GLIDE_STEP_FINISH
    return true;
}

// Counters the step's active count is spread over (GlideEmit.active_copies).
__device__ __forceinline__ uint32_t active_copies(const GlideEmit& e) {
    return e.active_copies > 1 ? static_cast<uint32_t>(e.active_copies) : 1u;
}

// Launch modes, fixed at compile time so each instantiation keeps only its
// own scalars live (the kernel is otherwise short of scalar registers).
constexpr int kModePlain = 0;  // scalars from the launch
constexpr int kModeTable = 1;  // scalars from the device table row (captured step body)
constexpr int kModeSlots = 2;  // served slots: scalars and window per wavefront


// FIXED: the fixed-point footprint (GLIDE_F_FIXED) as its own instantiation,
// so neither variant keeps the other's pointers live (the kernel is short of
// scalar registers; the spare ones spill to vector lanes).
//
// Occupancy: the bfloat16 kernels fit 96 VGPRs, 5 wavefronts per SIMD
// (measured 1-8 us per launch faster at 2^20 particles than at 4); the
// float32 ones keep 4 (held to 96 they spill 12-48 bytes per lane and run no
// faster).
template <bool BF16, int MODE, bool FIXED, bool CELL = false>
__global__ __launch_bounds__(kStepBlockMax) __attribute__((amdgpu_waves_per_eu(BF16 ? 5 : 4)))
void step_kernel(const StepArgs a) {
    __shared__ float lev[kLadder];
    fill_ladder(lev, a.levels, a.nz);
    __syncthreads();
    StepScalars p = step_scalars(a.p);
    StepCtx x{a.wind, a.aux, a.surf, a.e.escaped, -1, a.p.noise_hi};
    unsigned long long* n_active = a.n_active;
    const int64_t i =
        static_cast<int64_t>(MODE == kModeSlots ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) * blockDim.x +
        threadIdx.x;
    bool run = i < a.s.n;
    bool in_table = true;  // table mode: the cursor names a row
    if constexpr (MODE == kModeTable) {
        // every workgroup reads row[0] before any can advance it: the same r
        const int r = __builtin_amdgcn_readfirstlane(*a.row);
        in_table = r >= 0 && r < a.n_rows;
        if (in_table) {
            const GlideStepParams row = load_row(a.table, r);
            p = step_scalars(row);
            p.flags = a.p.flags;  // the template's, which the host checked the pointers for
            x.noise_hi = row.noise_hi;
            x.escaped += r;
            if (n_active != nullptr) n_active += static_cast<int64_t>(r) * active_copies(a.e);
        } else {
            if (!FIXED && run && (p.flags & GLIDE_F_EMIT)) {
                const uint32_t i4 = static_cast<uint32_t>(i) * 4u;
                sto(a.e.keys, i4, a.fp.n_keys);
                sto(a.e.vals, i4, 0.0f);
            }
            run = false;
            n_active = nullptr;
        }
    } else if constexpr (MODE == kModeSlots) {
        // Served slots: one slot per wavefront (64-particle blocks).
        const int32_t slot = __builtin_amdgcn_readfirstlane(run ? a.wave_slot[i >> 6] : -1);
        bool stepping = false;
        if (slot >= 0 && slot < a.fp.n_r) {  // the slots array holds fp.n_r entries
            GlideSlot* sl = a.slots + slot;
            stepping = __builtin_amdgcn_readfirstlane(sl->stepping) != 0;
            const int32_t hb = __builtin_amdgcn_readfirstlane(sl->bracket) - a.ring.first_hour;
            stepping = stepping && hb >= 0 && hb < a.ring.n_hours;  // never outside the ring
            if (stepping) {
                p.dt = static_cast<float>(sl->dt);
                p.al = static_cast<float>(sl->alpha);
                p.co = static_cast<float>(sl->cursor_off);
                p.step = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(sl->noise_step));
                const int64_t h = hb;
                x.wind = static_cast<const char*>(a.ring.wind) + h * a.ring.wind_stride;
                x.aux = static_cast<const char*>(a.ring.aux) + h * a.ring.aux_stride;
                x.surf = static_cast<const char*>(a.ring.surf) + h * a.ring.surf_stride;
                x.escaped = &sl->escaped;
                x.slot = slot;
                x.noise_hi = __builtin_amdgcn_readfirstlane(sl->noise_hi);
                n_active = &sl->active_steps;
                // a convecting slot's emission and escape kill follow its convection
                if (FIXED && __builtin_amdgcn_readfirstlane(sl->convect) != 0) p.flags &= ~GLIDE_F_EMIT;
            }
        }
        if (!stepping) {
            if (!FIXED && run && (p.flags & GLIDE_F_EMIT)) {
                const uint32_t i4 = static_cast<uint32_t>(i) * 4u;
                sto(a.e.keys, i4, a.fp.n_keys);
                sto(a.e.vals, i4, 0.0f);
            }
            run = false;
            n_active = nullptr;
        }
    }
    bool act = false;
    uint64_t fkey = kNoCell;
    long long fval = 0;
    constexpr bool kReload = BF16 || CELL || MODE != kModePlain;
    if (run) act = step_particle<BF16, FIXED, CELL, kReload>(a, lev, i, p, x, fkey, fval);
    if (FIXED && (p.flags & GLIDE_F_EMIT))
        wave_fixed_add(fkey, fval, fixed_copy(a.e, a.fp.acc_cells));  // every lane is here
    if constexpr (MODE == kModeSlots) {
        if (n_active != nullptr) count_flags(act, n_active);  // per wavefront: its slot's counter
    } else if (n_active != nullptr) {
        // active count: wave ballots summed in LDS, one atomic per workgroup at the end
        __shared__ uint32_t wact[kStepBlockMax / 64];
        const unsigned long long m = __ballot(act);
        if ((threadIdx.x & 63) == 0) wact[threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) c += wact[w];
            if (c) atomicAdd(n_active + blockIdx.x % active_copies(a.e), static_cast<unsigned long long>(c));
        }
    }
    if (MODE == kModeTable && in_table) {
        // Every workgroup has used row[0] before it counts itself finished, so
        // the last one to finish can advance the cursor (no waiting). The count
        // is split over kRowGroups counters and a top counter, so no single
        // address takes more than gridDim / kRowGroups atomics.
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t nb = gridDim.x;
            const uint32_t g = blockIdx.x % kRowGroups;
            const uint32_t members = (nb - g + kRowGroups - 1) / kRowGroups;
            uint32_t* cnt = reinterpret_cast<uint32_t*>(a.row + 2);
            if (atomicAdd(cnt + g, 1u) == members - 1) {
                atomicExch(cnt + g, 0u);
                const uint32_t groups = nb < kRowGroups ? nb : kRowGroups;
                if (atomicAdd(reinterpret_cast<uint32_t*>(a.row + 1), 1u) == groups - 1) {
                    atomicExch(reinterpret_cast<uint32_t*>(a.row + 1), 0u);
                    atomicAdd(a.row, 1);
                }
            }
        }
    }
}

__global__ void footprint_key_kernel(int64_t n, int64_t ps, const float* lon, const float* lat,
                                     const float* z, const uint8_t* active, const float* weight,
                                     const int64_t* t_idx, const int64_t* rel_idx, float dt,
                                     FpDev g, uint32_t* keys, float* vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool act = active[i] != 0;
    const uint32_t k = footprint_key(g, act, lon[i * ps], lat[i * ps], z[i * ps], t_idx[i], rel_idx[i]);
    keys[i] = k;
    vals[i] = (k == g.n_keys) ? 0.0f : weight[i] * dt;
}

// Table-mode step over no particles: the cursor advances by one when it
// names a row, as a step kernel's last workgroup would advance it.
__global__ void row_advance_kernel(int32_t* row, int32_t n_rows) {
    if (threadIdx.x != 0) return;
    const int32_t r = row[0];
    if (r >= 0 && r < n_rows) row[0] = r + 1;
}

static int fill_grid(StepArgs& a, const GlideGrid& g, const GlideWindow& w) {
    GLIDE_REQUIRE(g.nx >= 2 && g.ny >= 2 && g.nz >= 2, "met grid needs >= 2 points per axis");
    GLIDE_REQUIRE(g.nz <= kMaxLevels, "at most %d met levels supported (got %d)", kMaxLevels, g.nz);
    GLIDE_REQUIRE(g.lon1 != g.lon0 && g.lat1 != g.lat0, "Grid dimension has 0 span (first == last)");
    GLIDE_REQUIRE(g.levels != nullptr, "levels pointer is NULL");
    GLIDE_REQUIRE(w.wind && w.aux && w.surf, "met window pointers are NULL");
    GLIDE_REQUIRE(static_cast<uint64_t>(g.nz) * g.ny * g.nx * 32u < (1ull << 32),
                  "met window too large for 32-bit record offsets (%d x %d x %d)", g.nz, g.ny, g.nx);
    a.nx = g.nx;
    a.ny = g.ny;
    a.nz = g.nz;
    a.ltop = ladder_top(g.nz);
    a.planeb = static_cast<uint32_t>(static_cast<uint64_t>(g.nz) * g.ny * g.nx * 16u);
    a.lon0 = static_cast<float>(g.lon0);
    a.lat0 = static_cast<float>(g.lat0);
    a.sx = static_cast<float>((g.nx - 1) / (g.lon1 - g.lon0));
    a.sy = static_cast<float>((g.ny - 1) / (g.lat1 - g.lat0));
    a.cdlon = static_cast<float>((g.lon1 - g.lon0) / (g.nx - 1));
    a.cdlat = static_cast<float>((g.lat1 - g.lat0) / (g.ny - 1));
    a.blon0 = a.lon0;
    a.blat0 = a.lat0;
    a.bsx = a.sx;
    a.bsy = a.sy;
    a.levels = g.levels;
    a.wind = w.wind;
    a.aux = w.aux;
    a.surf = w.surf;
    return GLIDE_OK;
}

}  // namespace glide

using namespace glide;

// Launch timing (glide_step_timing_arm/read): the next armed step launches of
// the calling thread record their own start and end (events of the launch).
struct StepTiming {
    std::vector<hipEvent_t> ev;  // start, stop per recorded launch
    int armed = 0;
    int used = 0;
};
static thread_local StepTiming g_timing;

template <typename K>
static void launch_one(K kernel, dim3 grid, dim3 block, hipStream_t st, const StepArgs& a, hipEvent_t e0,
                       hipEvent_t e1) {
    if (e0 != nullptr) hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, a);
    else hipLaunchKernelGGL(kernel, grid, block, 0, st, a);
}

// One launch of the step kernel variant for the met precision and the footprint mode.
template <int MODE, bool CELL>
static void launch_variants(bool bf16, bool fixed, dim3 grid, dim3 block, hipStream_t st, const StepArgs& a,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (fixed) {
        if (bf16) launch_one(step_kernel<true, MODE, true, CELL>, grid, block, st, a, e0, e1);
        else launch_one(step_kernel<false, MODE, true, CELL>, grid, block, st, a, e0, e1);
    } else {
        if (bf16) launch_one(step_kernel<true, MODE, false, CELL>, grid, block, st, a, e0, e1);
        else launch_one(step_kernel<false, MODE, false, CELL>, grid, block, st, a, e0, e1);
    }
}

static int launch_step(const GlideStepParams* p, const GlideStepParams* table, int32_t n_rows, int32_t* row,
                       const GlideParticles* s, const GlideGrid* g, const GlideWindow* w,
                       const GlideFootprintGrid* fg, const GlideEmit* e, void* stream,
                       GlideSlot* slots = nullptr, const int32_t* wave_slot = nullptr,
                       const GlideMetRing* ring = nullptr) {
    GLIDE_REQUIRE(p && s && g && w, "glide_step: NULL argument");
    GLIDE_REQUIRE(s->n >= 0, "n must be >= 0");
    GLIDE_REQUIRE(s->lon && s->lat && s->z && s->up && s->vp && s->wp, "particle arrays are NULL");
    GLIDE_REQUIRE(s->pos_stride >= 1, "pos_stride must be >= 1");
    GLIDE_REQUIRE(static_cast<uint64_t>(s->n) * static_cast<uint64_t>(s->pos_stride) * 4u < (1ull << 32),
                  "particle arrays too large for 32-bit offsets");
    GLIDE_REQUIRE(p->max_substeps >= 1, "max_substeps must be >= 1");
    GLIDE_REQUIRE(p->substep_c > 0, "substep_c must be > 0");
    GLIDE_REQUIRE(p->dt >= 0, "dt must be >= 0");
    const int bs = p->block == 0 ? kStepBlock : p->block;
    GLIDE_REQUIRE(bs >= 64 && bs <= kStepBlockMax && bs % 64 == 0,
                  "step block %d: must be a multiple of 64 in [64, %d] (0 = %d)", p->block, kStepBlockMax, kStepBlock);
    GLIDE_REQUIRE(w->dtype == GLIDE_MET_F32 || w->dtype == GLIDE_MET_BF16, "unknown met dtype");
    if (p->flags & GLIDE_F_MEANDER) GLIDE_REQUIRE(s->um && s->vm, "meander state arrays are NULL");
    if (p->flags & GLIDE_F_MASK_INPUT) {
        GLIDE_REQUIRE(s->active, "active mask is NULL");
    } else {
        GLIDE_REQUIRE(s->r_off && s->alive, "r_off/alive are NULL (no explicit mask)");
        GLIDE_REQUIRE(!(p->flags & GLIDE_F_RELEASE_TERMINUS) || s->lo_off, "GLIDE_F_RELEASE_TERMINUS needs lo_off");
    }
    const bool cell = s->cell != nullptr;
    if (cell) {
        GLIDE_REQUIRE(slots == nullptr, "cell positions are not taken by glide_step_slots");
        GLIDE_REQUIRE(s->pos_stride == 1, "cell positions need pos_stride 1");
        GLIDE_REQUIRE(g->nx <= 65535 && g->ny <= 65535, "cell positions need nx, ny <= 65535");
    }
    StepArgs a{};
    a.p = *p;
    a.s = *s;
    a.table = table;
    a.n_rows = n_rows;
    a.row = row;
    a.n_active = e ? e->active : nullptr;
    if (e != nullptr) {
        GLIDE_REQUIRE(e->active_copies >= 0 && e->active_copies <= 1024, "active_copies must be in [0, 1024]");
        a.e.active_copies = e->active_copies;  // read by the step kernel with or without GLIDE_F_EMIT
    }
    a.slots = slots;
    a.wave_slot = wave_slot;
    if (ring != nullptr) a.ring = *ring;
    int rc = fill_grid(a, *g, *w);
    if (rc) return rc;
    if (p->flags & GLIDE_F_EMIT) {
        if (p->flags & GLIDE_F_FIXED) {
            GLIDE_REQUIRE(fg && e && e->fixed && e->fixed_exp && e->escaped, "fixed-point emit outputs are NULL");
            GLIDE_REQUIRE(fg->t_bin == nullptr, "GLIDE_F_FIXED takes flat keys (t_bin NULL)");
            GLIDE_REQUIRE(e->fixed_copies >= 0 && e->fixed_copies <= 1024, "fixed_copies must be in [0, 1024]");
        } else {
            GLIDE_REQUIRE(fg && e && e->keys && e->vals && e->escaped, "emit outputs are NULL");
        }
        GLIDE_REQUIRE(s->we_off && s->weight && s->alive, "we_off/weight/alive needed to emit");
        GLIDE_REQUIRE(fg->n_r >= 1 && fg->n_t >= 1 && fg->n_z >= 1 && fg->n_y >= 1 && fg->n_x >= 1,
                      "All footprint dimensions must be > 0");
        const uint64_t cells = static_cast<uint64_t>(fg->n_r) * fg->n_t * fg->n_z * fg->n_y * fg->n_x;
        if (p->flags & GLIDE_F_FIXED) {
            GLIDE_REQUIRE(fg->slab_off == nullptr || (fg->slab_nt != nullptr && fg->slab_cells > 0),
                          "per-release slabs need slab_nt and slab_cells");
        } else {
            GLIDE_REQUIRE(cells < 0xffffffffull, "footprint has too many cells for 32-bit keys");
            GLIDE_REQUIRE(fg->slab_off == nullptr, "per-release slabs need GLIDE_F_FIXED");
        }
        GLIDE_REQUIRE(fg->n_z <= 64, "at most 64 footprint z bins");
        a.fp = make_fp(*fg);
        a.e = *e;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (s->n == 0) {
        // No particle to step, but a table-mode call still advances its cursor
        // once (a row in range), so the replays after it read the right rows.
        if (table != nullptr) {
            hipLaunchKernelGGL(row_advance_kernel, dim3(1), dim3(64), 0, st, row, n_rows);
            GLIDE_HIP(hipGetLastError());
        }
        return GLIDE_OK;
    }
    const int64_t blocks = (s->n + bs - 1) / bs;
    GLIDE_REQUIRE(blocks < (1ll << 31), "too many particles");
    const dim3 grid(static_cast<unsigned>(blocks)), block(static_cast<unsigned>(bs));
    const bool bf16 = w->dtype == GLIDE_MET_BF16;
    const bool fixed = (p->flags & (GLIDE_F_EMIT | GLIDE_F_FIXED)) == (GLIDE_F_EMIT | GLIDE_F_FIXED);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_timing.armed > 0 && slots == nullptr) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GLIDE_HIP(hipStreamIsCapturing(st, &cs));
        if (cs == hipStreamCaptureStatusNone) {  // never inside a captured graph
            e0 = g_timing.ev[2 * g_timing.used];
            e1 = g_timing.ev[2 * g_timing.used + 1];
            ++g_timing.used;
            --g_timing.armed;
        }
    }
    if (table != nullptr) {
        if (cell) launch_variants<kModeTable, true>(bf16, fixed, grid, block, st, a, e0, e1);
        else launch_variants<kModeTable, false>(bf16, fixed, grid, block, st, a, e0, e1);
    } else if (slots != nullptr) {
        if (fixed) {
            if (bf16) hipLaunchKernelGGL((step_kernel<true, kModeSlots, true>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((step_kernel<false, kModeSlots, true>), grid, block, 0, st, a);
        } else {
            if (bf16) hipLaunchKernelGGL((step_kernel<true, kModeSlots, false>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((step_kernel<false, kModeSlots, false>), grid, block, 0, st, a);
        }
    } else {
        if (cell) launch_variants<kModePlain, true>(bf16, fixed, grid, block, st, a, e0, e1);
        else launch_variants<kModePlain, false>(bf16, fixed, grid, block, st, a, e0, e1);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_step_timing_arm(int32_t n) {
    GLIDE_REQUIRE(n >= 0 && n <= (1 << 20), "glide_step_timing_arm: n must be in [0, 2^20]");
    const size_t need = 2 * static_cast<size_t>(n);
    while (g_timing.ev.size() < need) {
        hipEvent_t e;
        GLIDE_HIP(hipEventCreate(&e));
        g_timing.ev.push_back(e);
    }
    g_timing.armed = n;
    g_timing.used = 0;
    return GLIDE_OK;
}

extern "C" int glide_step_timing_read(float* ms, int32_t n) {
    GLIDE_REQUIRE(ms != nullptr || n == 0, "glide_step_timing_read: NULL output");
    const int m = g_timing.used < n ? g_timing.used : n;
    for (int k = 0; k < m; ++k) {
        GLIDE_HIP(hipEventSynchronize(g_timing.ev[2 * k + 1]));
        GLIDE_HIP(hipEventElapsedTime(ms + k, g_timing.ev[2 * k], g_timing.ev[2 * k + 1]));
    }
    g_timing.armed = 0;
    return m;
}

extern "C" int glide_step(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
                          const GlideWindow* w, const GlideFootprintGrid* fg, const GlideEmit* e,
                          void* stream) {
    return launch_step(p, nullptr, 0, nullptr, s, g, w, fg, e, stream);
}

__global__ void slots_advance_kernel(GlideSlot* slots, int32_t n, double dt, int32_t flags) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    GlideSlot sl = slots[j];
    sl.stepping = 0;
    sl.convect = 0;
    if (sl.state == GLIDE_SLOT_RUNNING) {
        if (sl.k > 0) sl.t -= sl.last_d;
        if (sl.t > sl.term) {
            // the reference's cursor loop for a batch of this release alone
            const double d = fmin(dt, sl.t - sl.term);
            const double bh = floor(sl.t / 3600.0);
            double al = ((sl.t - 0.5 * d) - 3600.0 * bh) / 3600.0;
            al = al < 0.0 ? 0.0 : (al > 1.0 ? 1.0 : al);
            sl.last_d = d;
            sl.dt = static_cast<float>(d);
            sl.alpha = static_cast<float>(al);
            sl.cursor_off = static_cast<float>(sl.t - sl.rel_time);
            sl.noise_step = static_cast<int32_t>(floor(sl.t / dt));
            // the reference convects on a new bracket (main.py:1446-1458); a slot's first step is one
            if ((flags & GLIDE_SLOTS_CONVECT) && (sl.k == 0 || sl.bracket != static_cast<int32_t>(bh))) sl.convect = 1;
            sl.bracket = static_cast<int32_t>(bh);
            sl.stepping = 1;
            sl.k += 1;
        } else {
            sl.state = GLIDE_SLOT_FINISHED;
        }
    }
    slots[j] = sl;
}

extern "C" int glide_slots_advance(GlideSlot* slots, int32_t n_slots, double dt, int32_t flags, void* stream) {
    GLIDE_REQUIRE(slots != nullptr && n_slots >= 0, "glide_slots_advance: bad slots");
    GLIDE_REQUIRE(dt > 0, "dt must be > 0");
    GLIDE_REQUIRE((flags & ~GLIDE_SLOTS_CONVECT) == 0, "unknown slots_advance flags 0x%x", flags);
    if (n_slots == 0) return GLIDE_OK;
    hipLaunchKernelGGL(slots_advance_kernel, dim3((n_slots + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), slots, n_slots, dt, flags);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_step_slots(const GlideStepParams* tmpl, GlideSlot* slots, const int32_t* wave_slot,
                                const GlideParticles* s, const GlideGrid* g, const GlideMetRing* ring,
                                const GlideFootprintGrid* fg, const GlideEmit* e, void* stream) {
    GLIDE_REQUIRE(tmpl && slots && wave_slot && s && ring, "glide_step_slots: NULL argument");
    GLIDE_REQUIRE(s->n % 64 == 0, "the particle pool must hold whole 64-particle blocks");
    GLIDE_REQUIRE(tmpl->flags & GLIDE_F_RELEASE_TERMINUS, "served slots step with GLIDE_F_RELEASE_TERMINUS");
    GLIDE_REQUIRE(tmpl->flags & GLIDE_F_EMIT, "served slots emit footprint keys (GLIDE_F_EMIT)");
    GLIDE_REQUIRE(!(tmpl->flags & GLIDE_F_MASK_INPUT), "served slots compute their own mask");
    GLIDE_REQUIRE(ring->wind && ring->aux && ring->surf && ring->n_hours >= 1, "empty met ring");
    const GlideWindow w{ring->wind, ring->aux, ring->surf, ring->dtype, 0};
    return launch_step(tmpl, nullptr, 0, nullptr, s, g, &w, fg, e, stream, slots, wave_slot, ring);
}

extern "C" int glide_step_table(const GlideStepParams* tmpl, const GlideStepParams* table, int32_t n_rows,
                                int32_t* row, const GlideParticles* s, const GlideGrid* g, const GlideWindow* w,
                                const GlideFootprintGrid* fg, const GlideEmit* e, void* stream) {
    GLIDE_REQUIRE(table != nullptr && row != nullptr, "glide_step_table: NULL table/row");
    GLIDE_REQUIRE(n_rows >= 1, "glide_step_table: n_rows must be >= 1");
    return launch_step(tmpl, table, n_rows, row, s, g, w, fg, e, stream);
}

__global__ void row_set_kernel(int32_t* row, int32_t value) {
    for (int k = threadIdx.x; k < GLIDE_ROW_WORDS; k += blockDim.x) row[k] = k == 0 ? value : 0;
}

extern "C" int glide_row_set(int32_t* row, int32_t value, void* stream) {
    GLIDE_REQUIRE(row != nullptr, "NULL row");
    hipLaunchKernelGGL(row_set_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), row, value);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_footprint_keys(int64_t n, int64_t pos_stride, const float* lon, const float* lat,
                                    const float* z, const uint8_t* active, const float* weight,
                                    const int64_t* t_idx, const int64_t* rel_idx, double dt,
                                    const GlideFootprintGrid* fg, uint32_t* keys, float* vals,
                                    void* stream) {
    GLIDE_REQUIRE(fg, "NULL footprint grid");
    GLIDE_REQUIRE(n >= 0 && pos_stride >= 1, "bad n / pos_stride");
    GLIDE_REQUIRE(fg->n_r >= 1 && fg->n_t >= 1 && fg->n_z >= 1 && fg->n_y >= 1 && fg->n_x >= 1,
                  "All footprint dimensions must be > 0");
    const uint64_t cells = static_cast<uint64_t>(fg->n_r) * fg->n_t * fg->n_z * fg->n_y * fg->n_x;
    GLIDE_REQUIRE(cells < 0xffffffffull, "footprint has too many cells for 32-bit keys");
    GLIDE_REQUIRE(fg->n_z <= 64, "at most 64 footprint z bins");
    if (n == 0) return GLIDE_OK;
    GLIDE_REQUIRE(lon && lat && z && active && weight && t_idx && rel_idx && keys && vals && fg->z_edges,
                  "NULL array");
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(footprint_key_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), n, pos_stride, lon, lat, z, active, weight,
                       t_idx, rel_idx, static_cast<float>(dt), make_fp(*fg), keys, vals);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
