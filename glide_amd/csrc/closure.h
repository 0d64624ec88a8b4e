// Free-troposphere gradient-Richardson closure and the meander stencil std,
// in float, shared by the window build and the Hanna evaluation entry point.
// Reference: src/lpdm/turbulence/hanna.py:340-402 (free functions) and
// hanna.py:1786-1830 (_meander_sigma_fields / _windowed_std).
#pragma once

#include "common.h"

namespace glide {

constexpr float kKappaRdCp = 287.05f / 1005.0f;  // R_d / c_p

// theta = T (p0 / p)^kappa, p0 = 1000 hPa (hanna.py:340-345).
__device__ __forceinline__ float potential_temperature(float t, float p_pa) {
    return t * powf(100000.0f / p_pa, kKappaRdCp);
}

// N^2 = g / theta * dtheta/dz, theta clamped >= 1 K (hanna.py:348-353).
__device__ __forceinline__ float brunt_vaisala_squared(float theta, float dtheta_dz) {
    return (9.80665f / fmaxf(theta, 1.0f)) * dtheta_dz;
}

// Ri = N^2 / max(S^2, 1e-8) (hanna.py:356-361).
__device__ __forceinline__ float gradient_richardson(float n2, float shear_sq) {
    return n2 / fmaxf(shear_sq, 1e-8f);
}

// K = clamp(l^2 |S| F(Ri), 0.1, 50), l = kappa z / (1 + kappa z / 100) with z
// clamped >= 1 m, F = sqrt(1 - 16 Ri) for Ri < 0, (1 - Ri / 0.25)^2 clipped at
// zero above (hanna.py:364-385).
__device__ __forceinline__ float free_trop_diffusivity(float z, float shear_mag, float ri) {
    const float hz = fmaxf(z, 1.0f);
    const float ml = (0.4f * hz) / (1.0f + 0.4f * hz / 100.0f);
    float fri;
    if (ri < 0.0f) {
        fri = sqrtf(fmaxf(1.0f - 16.0f * ri, 0.0f));
    } else {
        const float q = fmaxf(1.0f - ri / 0.25f, 0.0f);
        fri = q * q;
    }
    return fminf(fmaxf(ml * ml * shear_mag * fri, 0.1f), 50.0f);
}

// T_Lw = clamp(0.5 / N, 1, 1000) (100 s when N^2 <= 1e-6); sigma_w =
// sqrt(max(K / T_Lw, 1e-6)) (hanna.py:388-402).
__device__ __forceinline__ void free_trop_sigma_tl(float kz, float n2, float& sw, float& tl) {
    tl = (n2 > 1e-6f) ? fminf(fmaxf(0.5f / sqrtf(fmaxf(n2, 1e-6f)), 1.0f), 1000.0f) : 100.0f;
    sw = sqrtf(fmaxf(kz / tl, 1e-6f));
}

// Two-pass population std of 0.5 * (s + e) over the edge-clipped stencil.
__device__ inline float local_std(const float* __restrict__ s, const float* __restrict__ e, int y, int x, int ny,
                           int nx, int r) {
    if (r == 1) {
        // the shipped 3 x 3 stencil: the midpoints stay in registers between
        // the passes (the same values, added in the same order)
        float v[9];
        bool in[9];
        int cnt = 0;
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int yy = y + q / 3 - 1, xx = x + q % 3 - 1;
            in[q] = yy >= 0 && yy < ny && xx >= 0 && xx < nx;
            v[q] = 0.0f;
            if (in[q]) {
                const int64_t c = static_cast<int64_t>(yy) * nx + xx;
                v[q] = 0.5f * (s[c] + e[c]);
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
    float sum = 0.0f;
    int cnt = 0;
    for (int dy = -r; dy <= r; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= ny) continue;
        for (int dx = -r; dx <= r; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= nx) continue;
            const int64_t c = static_cast<int64_t>(yy) * nx + xx;
            sum += 0.5f * (s[c] + e[c]);
            ++cnt;
        }
    }
    const float mean = sum / static_cast<float>(cnt);
    float ss = 0.0f;
    for (int dy = -r; dy <= r; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= ny) continue;
        for (int dx = -r; dx <= r; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= nx) continue;
            const int64_t c = static_cast<int64_t>(yy) * nx + xx;
            const float d = 0.5f * (s[c] + e[c]) - mean;
            ss += d * d;
        }
    }
    return sqrtf(ss / static_cast<float>(cnt));
}

}  // namespace glide
