// Per-hour unit conversions of raw source fields (the met reader's host-side
// conversions, met_reader.py:1052-1078, 1268-1300, 1353-1385), in place on the
// device before the resample, with the reader's arithmetic:
//   height AGL = (geopotential - surface geopotential) / g        float32
//   w = -(R_d T / (g p)) omega, p the level pressure broadcast   float64, to float32
//   shf (upward W/m^2) = -(accumulated J/m^2 / accumulation s)    float32
//                      or -(instantaneous downward W/m^2)
#include "common.h"

namespace glide {

constexpr double kUnitsRdry = 287.05;
constexpr double kUnitsGravity = 9.80665;

struct UnitsArgs {
    float* src3;
    int32_t zp, w_ch, t_ch;
    int64_t ncol;
    const double* p_level_pa;
    const float* geop;
    const float* geop_sfc;
    float* height;
    float* shf;
    int32_t shf_units;
    float accum_s;
};

__global__ void source_units_kernel(const UnitsArgs a) {
#pragma clang fp contract(off)
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (c >= a.ncol) return;
    const int64_t i = static_cast<int64_t>(k) * a.ncol + c;
    if (a.geop != nullptr) a.height[i] = (a.geop[i] - a.geop_sfc[c]) / static_cast<float>(kUnitsGravity);
    if (a.w_ch >= 0) {
        const double t = static_cast<double>(a.src3[static_cast<int64_t>(a.t_ch) * a.zp * a.ncol + i]);
        const double om = static_cast<double>(a.src3[static_cast<int64_t>(a.w_ch) * a.zp * a.ncol + i]);
        const double p = a.p_level_pa[k];
        a.src3[static_cast<int64_t>(a.w_ch) * a.zp * a.ncol + i] =
            static_cast<float>(-(kUnitsRdry * t / (kUnitsGravity * p)) * om);
    }
    if (k == 0 && a.shf_units != GLIDE_SHF_UPWARD_W_M2) {
        const float v = a.shf[c];
        a.shf[c] = a.shf_units == GLIDE_SHF_ACCUMULATED_J_M2 ? -(v / a.accum_s) : -v;
    }
}

}  // namespace glide

using namespace glide;

extern "C" int glide_source_units(float* src3, int32_t c3, int32_t zp, int32_t ny, int32_t nx, int32_t w_ch,
                                  int32_t t_ch, const double* p_level_pa, const float* geopotential,
                                  const float* geopotential_sfc, float* height_agl, float* shf, int32_t shf_units,
                                  double accumulation_s, void* stream) {
    GLIDE_REQUIRE(zp >= 1 && ny >= 1 && nx >= 1 && c3 >= 1, "bad source dimensions");
    GLIDE_REQUIRE(w_ch < c3 && t_ch < c3, "bad w/t channel index");
    GLIDE_REQUIRE(w_ch < 0 || (src3 && t_ch >= 0 && p_level_pa), "omega -> w needs src3, the t channel and p_level_pa");
    GLIDE_REQUIRE(geopotential == nullptr || (geopotential_sfc && height_agl),
                  "geopotential -> height needs the surface geopotential and a height output");
    GLIDE_REQUIRE(shf_units == GLIDE_SHF_UPWARD_W_M2 || shf_units == GLIDE_SHF_DOWNWARD_W_M2 ||
                      shf_units == GLIDE_SHF_ACCUMULATED_J_M2,
                  "unknown sensible heat flux units %d", shf_units);
    GLIDE_REQUIRE(shf_units == GLIDE_SHF_UPWARD_W_M2 || shf, "shf conversion needs the shf field");
    GLIDE_REQUIRE(shf_units != GLIDE_SHF_ACCUMULATED_J_M2 || accumulation_s > 0, "accumulation_s must be > 0");
    if (w_ch < 0 && geopotential == nullptr && shf_units == GLIDE_SHF_UPWARD_W_M2) return GLIDE_OK;
    UnitsArgs a{src3, zp, w_ch, t_ch, static_cast<int64_t>(ny) * nx, p_level_pa, geopotential, geopotential_sfc,
                height_agl, shf, shf_units, static_cast<float>(accumulation_s)};
    const int64_t blocks = (a.ncol + 255) / 256;
    GLIDE_REQUIRE(blocks < (1ll << 31), "too many columns");
    hipLaunchKernelGGL(source_units_kernel, dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(zp)), dim3(256),
                       0, static_cast<hipStream_t>(stream), a);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
