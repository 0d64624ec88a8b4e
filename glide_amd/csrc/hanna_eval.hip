// Hanna (1982) / FLEXPART free functions evaluated element-wise on the device
// with the step kernel's own float code (physics.h) and the window build's
// closure (closure.h): the reference's free-function API
// (turbulence/hanna.py:142-189, 192-337, 340-426, 1786-1830) over device
// arrays, so its unit tests run against the code the step executes.
#include "closure.h"
#include "physics.h"

namespace glide {

struct HannaEvalArgs {
    int32_t op;
    int64_t n;
    const float* in[6];
    float* out[6];
};

__global__ void hanna_eval_kernel(const HannaEvalArgs a) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    auto in = [&](int k) { return a.in[k][i]; };
    switch (a.op) {
        case GLIDE_HANNA_CORIOLIS:
            a.out[0][i] = 2.0f * kOmega * fsin_deg(in(0));
            break;
        case GLIDE_HANNA_AIR_DENSITY:  // the density inside stability()
            a.out[0][i] = fdiv(in(0), kRDry * in(1));
            break;
        case GLIDE_HANNA_OBUKHOV:
        case GLIDE_HANNA_CONVECTIVE_VELOCITY: {
            // in: (ustar, shf, t, sp) / (blh, shf, t, sp), as the reference orders them
            float L, wstar;
            if (a.op == GLIDE_HANNA_OBUKHOV) {
                stability(kBlhMin, in(3), in(0), in(1), in(2), L, wstar);
                a.out[0][i] = L;
            } else {
                stability(in(0), in(3), kUstarMin, in(1), in(2), L, wstar);
                a.out[0][i] = wstar;
            }
            break;
        }
        case GLIDE_HANNA_IN_BL: {
            Surface s;
            s.blh = in(1);
            s.ustar = in(2);
            s.wstar = in(3);
            s.hol = in(4);
            s.L = 0.0f;
            s.fcor = 2.0f * kOmega * fsin_deg(in(5));
            const SigTL t = in_bl(in(0), s);
            a.out[0][i] = t.su;
            a.out[1][i] = t.sv;
            a.out[2][i] = t.sw;
            a.out[3][i] = t.tu;
            a.out[4][i] = t.tv;
            a.out[5][i] = t.tw;
            break;
        }
        case GLIDE_HANNA_SURFACE_LAYER_SIGMA_W:
            a.out[0][i] = surface_layer_sw(in(0), in(1), in(2));
            break;
        case GLIDE_HANNA_POTENTIAL_TEMPERATURE:
            a.out[0][i] = potential_temperature(in(0), in(1));
            break;
        case GLIDE_HANNA_BRUNT_VAISALA:
            a.out[0][i] = brunt_vaisala_squared(in(0), in(1));
            break;
        case GLIDE_HANNA_RICHARDSON:
            a.out[0][i] = gradient_richardson(in(0), in(1));
            break;
        case GLIDE_HANNA_FT_DIFFUSIVITY:
            a.out[0][i] = free_trop_diffusivity(in(0), in(1), in(2));
            break;
        case GLIDE_HANNA_FT_SIGMA_TL: {
            float sw, tl;
            free_trop_sigma_tl(in(0), in(1), sw, tl);
            a.out[0][i] = sw;
            a.out[1][i] = tl;
            break;
        }
        default:
            break;
    }
}

__global__ void windowed_std_kernel(const float* field, int32_t nz, int32_t ny, int32_t nx, int32_t r,
                                    float* out) {
    const int64_t ncol = static_cast<int64_t>(ny) * nx;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nz) return;
    const int64_t k = i / ncol, c = i % ncol;
    const float* f = field + k * ncol;
    out[i] = local_std(f, f, static_cast<int>(c / nx), static_cast<int>(c % nx), ny, nx, r);
}

}  // namespace glide

using namespace glide;

extern "C" int glide_hanna_eval(int32_t op, int64_t n, const float* const* in, float* const* out, void* stream) {
    static const int kIn[] = {0, 1, 2, 4, 4, 6, 3, 2, 2, 2, 3, 2};
    static const int kOut[] = {0, 1, 1, 1, 1, 6, 1, 1, 1, 1, 1, 2};
    GLIDE_REQUIRE(op >= GLIDE_HANNA_CORIOLIS && op <= GLIDE_HANNA_FT_SIGMA_TL, "unknown Hanna function %d", op);
    GLIDE_REQUIRE(n >= 0, "n must be >= 0");
    GLIDE_REQUIRE(in != nullptr && out != nullptr, "NULL input/output pointer arrays");
    HannaEvalArgs a{};
    a.op = op;
    a.n = n;
    for (int k = 0; k < kIn[op]; ++k) {
        GLIDE_REQUIRE(in[k] != nullptr, "input %d of Hanna function %d is NULL", k, op);
        a.in[k] = in[k];
    }
    for (int k = 0; k < kOut[op]; ++k) {
        GLIDE_REQUIRE(out[k] != nullptr, "output %d of Hanna function %d is NULL", k, op);
        a.out[k] = out[k];
    }
    if (n == 0) return GLIDE_OK;
    hipLaunchKernelGGL(hanna_eval_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), a);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_windowed_std(const float* field, int32_t nz, int32_t ny, int32_t nx, int32_t radius, float* out,
                                  void* stream) {
    GLIDE_REQUIRE(field != nullptr && out != nullptr, "NULL field/out");
    GLIDE_REQUIRE(nz >= 1 && ny >= 1 && nx >= 1, "bad field dimensions");
    GLIDE_REQUIRE(radius >= 1, "meander_stencil_radius must be >= 1");
    const int64_t total = static_cast<int64_t>(nz) * ny * nx;
    hipLaunchKernelGGL(windowed_std_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), field, nz, ny, nx, radius, out);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
