// Hanna (1982) / FLEXPART boundary-layer turbulence, free-troposphere closure
// and the engine primitives, in float, for the fused step kernel.
// Reference: src/lpdm/turbulence/hanna.py:47-426 and src/lpdm/gpu_engine.py:26-42.
#pragma once

#include "common.h"

namespace glide {

constexpr float kKarman = 0.4f;
constexpr float kGravity = 9.80665f;
constexpr float kRDry = 287.05f;
constexpr float kCp = 1005.0f;
constexpr float kOmega = 7.2921e-5f;
constexpr float kSigmaMin = 1e-3f;
constexpr float kTLMin = 1.0f;
constexpr float kUstarMin = 1e-3f;
constexpr float kBlhMin = 50.0f;
constexpr float kZMin = 1.0f;
constexpr float kMPerDegLat = 110540.0f;
constexpr float kMPerDegLonEq = 111320.0f;
constexpr float kDeg2Rad = 0.017453292519943295f;

// Fast float math on the hardware instructions (about 1 ulp; the per-step
// position bar is 1e-5 relative). The bit-exact index math does not use these.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^y for x >= 0, y > 0 (exp2(y log2 x); 0^y = 0)
__device__ __forceinline__ float fpow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float fcbrt(float x) { return fpow(x, 0.3333333333333333f); }
// cos / sin of an angle in degrees (the hardware takes revolutions)
__device__ __forceinline__ float fcos_deg(float d) { return __builtin_amdgcn_cosf(d * (1.0f / 360.0f)); }
__device__ __forceinline__ float fsin_deg(float d) { return __builtin_amdgcn_sinf(d * (1.0f / 360.0f)); }

// 111320 * max(|cos(lat)|, 0.05) (gpu_engine.py:467, hanna.py:884).
__device__ __forceinline__ float m_per_deg_lon(float lat_deg) {
    return kMPerDegLonEq * fmaxf(fabsf(fcos_deg(lat_deg)), 0.05f);
}

struct Surface {
    float blh, ustar, wstar, hol, L, fcor;
};

// obukhov_length + convective_velocity (hanna.py:148-189).
__device__ __forceinline__ void stability(float blh, float sp, float ustar, float shf, float tsf,
                                          float& L, float& wstar) {
    const float rho = fdiv(sp, kRDry * tsf);
    const float denom = kKarman * kGravity * shf;
    const float us = fmaxf(ustar, kUstarMin);
    const float num = -(us * us * us) * tsf * rho * kCp;
    L = (fabsf(denom) > 1e-10f) ? fdiv(num, denom) : __builtin_inff();
    const float cube = fdiv(kGravity * fmaxf(blh, kBlhMin) * shf, tsf * rho * kCp);
    wstar = (cube > 0.0f) ? fcbrt(cube) : 0.0f;
}

struct SigTL {
    float su, sv, sw, tu, tv, tw;
};

// in_bl_sigma_TL (hanna.py:192-337): the regime selected by h/L.
__device__ __forceinline__ SigTL in_bl(float z, const Surface& s) {
    SigTL r;
    const float blh = s.blh, us = s.ustar;
    if (s.hol > 1.0f) {  // stable, Hanna (1982) Eqs 7.19-7.24
        const float zoh = fminf(fmaxf(fdiv(z, blh), 0.0f), 1.0f);
        const float om = fmaxf(1.0f - zoh, 0.0f);
        r.su = 2.0f * us * om;
        r.sw = 1.3f * us * om;
        r.sv = r.sw;
        const float zf = fmaxf(zoh, 1e-6f);
        const float sq = fsqrt(zf);
        r.tu = fdiv(0.15f * blh, fmaxf(r.su, kSigmaMin)) * sq;
        r.tv = fdiv(0.07f * blh, fmaxf(r.sv, kSigmaMin)) * sq;
        r.tw = fdiv(0.10f * blh, fmaxf(r.sw, kSigmaMin)) * fpow(zf, 0.8f);
    } else if (s.hol < -1.0f) {  // unstable
        const float zoh = fminf(fmaxf(fdiv(z, blh), 0.0f), 1.0f);
        const float hc = fminf(fmaxf(s.hol, -1000.0f), 0.0f);
        r.su = fcbrt(fmaxf(12.0f - 0.5f * hc, 1e-3f)) * us;
        r.sv = r.su;
        const float sw2 = 1.2f * s.wstar * s.wstar * fmaxf(1.0f - 0.9f * zoh, 0.0f) *
                              fpow(zoh, 2.0f / 3.0f) +
                          (1.8f - 1.4f * zoh) * us * us;
        r.sw = fsqrt(fmaxf(sw2, kSigmaMin * kSigmaMin));
        r.tu = fdiv(0.15f * blh, fmaxf(r.su, kSigmaMin));
        r.tv = r.tu;
        const float azl = fmaxf(zoh * (-hc), 0.0f);
        const float swc = fmaxf(r.sw, kSigmaMin);
        if (azl < 1.0f) {
            r.tw = fdiv(0.1f * z, swc * fmaxf(0.55f - 0.38f * azl, 1e-2f));
        } else if (zoh < 0.1f) {
            r.tw = fdiv(0.59f * z, swc);
        } else {
            r.tw = fdiv(0.15f * blh, swc) * (1.0f - fexp(-5.0f * zoh));
        }
    } else {  // neutral, Hanna (1982) Eqs 7.25-7.27 with |f|
        const float usf = fmaxf(us, kUstarMin);
        const float fa = fmaxf(fabsf(s.fcor), 1e-5f);
        r.su = 2.0f * us * fexp(fdiv(-3.0f * fa * z, usf));
        r.sw = 1.3f * us * fexp(fdiv(-2.0f * fa * z, usf));
        r.sv = r.sw;
        const float den = 1.0f + fdiv(15.0f * fa * z, usf);
        r.tw = fdiv(fdiv(0.5f * z, fmaxf(r.sw, kSigmaMin)), fmaxf(den, 1.0f));
        r.tu = r.tw;
        r.tv = r.tw;
    }
    return r;
}

// sigma_w only (the finite-difference probes need nothing else).
__device__ __forceinline__ float in_bl_sw(float z, const Surface& s) {
    const float us = s.ustar;
    if (s.hol > 1.0f) {
        const float zoh = fminf(fmaxf(fdiv(z, s.blh), 0.0f), 1.0f);
        return 1.3f * us * fmaxf(1.0f - zoh, 0.0f);
    } else if (s.hol < -1.0f) {
        const float zoh = fminf(fmaxf(fdiv(z, s.blh), 0.0f), 1.0f);
        const float sw2 = 1.2f * s.wstar * s.wstar * fmaxf(1.0f - 0.9f * zoh, 0.0f) *
                              fpow(zoh, 2.0f / 3.0f) +
                          (1.8f - 1.4f * zoh) * us * us;
        return fsqrt(fmaxf(sw2, kSigmaMin * kSigmaMin));
    }
    const float usf = fmaxf(us, kUstarMin);
    const float fa = fmaxf(fabsf(s.fcor), 1e-5f);
    return 1.3f * us * fexp(fdiv(-2.0f * fa * z, usf));
}

// surface_layer_sigma_w (hanna.py:405-426).
__device__ __forceinline__ float surface_layer_sw(float z, float ustar, float L) {
    const bool finite = fabsf(L) > 1e-6f;
    const float zol = finite ? fminf(fmaxf(fdiv(z, L), -50.0f), 50.0f) : 0.0f;
    if (zol < 0.0f) return 1.3f * ustar * fcbrt(fmaxf(1.0f - 3.0f * zol, 1e-6f));
    return 1.3f * ustar;
}

// OU update (gpu_engine.py:26-42): a = exp(-dt/T_L); var = max(s2 (1 - a^2), 0).
__device__ __forceinline__ float ou(float w, float tl, float s2, float dt, float eta, float drift_dt) {
    const float a = fexp(-fdiv(dt, tl));
    const float var = fmaxf(s2 * (1.0f - a * a), 0.0f);
    return a * w + drift_dt + fsqrt(var) * eta;
}

// The same update split in two: the decay a and amplitude sqrt(var), which
// depend only on (T_L, sigma^2, dt), and their application to one draw.
struct OuCoef {
    float a, amp;
};
__device__ __forceinline__ OuCoef ou_coef(float tl, float s2, float dt) {
    const float a = fexp(-fdiv(dt, tl));
    return OuCoef{a, fsqrt(fmaxf(s2 * (1.0f - a * a), 0.0f))};
}
__device__ __forceinline__ float ou_apply(float w, const OuCoef& c, float eta, float drift_dt) {
    return c.a * w + drift_dt + c.amp * eta;
}

}  // namespace glide
