// Fixed-point footprint (GLIDE_F_FIXED) to float.
//
// The accumulator may hold several copies of the cells (GlideEmit.fixed_copies,
// spreading the atomics of crowded cells); a cell's sum is over its copies.
//
// The step adds every particle's residence v (float weight * dt) of release r
// to an int64 cell as rint(v * 2^e_r) (wave_fixed_add). Here each cell becomes
// float(acc) * 2^-e_r: the int64 -> float conversion rounds once, to nearest,
// so a cell is the correctly rounded sum of its contributions whenever each was
// held exactly. The release totals (the reference's residence per release, the
// same contributions the footprint keeps) are the exact integer sums of the
// release's cells, in float64. HBM-bound: 8 bytes read and 4 written per cell.
#include "common.h"

namespace glide {

constexpr int kFinThreads = 256;
constexpr int kFinItems = 8;  // cells per thread per block-sized pass, before the grid stride

__global__ void fixed_zero_kernel(long long* w, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) w[j] = 0;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block (bx, r): cells bx*256 + t (+ grid stride) of release r's slab.
__global__ __launch_bounds__(kFinThreads) void fixed_finalize_kernel(const long long* fixed, int copies,
                                                                      uint64_t n_cells, uint32_t cpr,
                                                                      const int32_t* exp, float* slab,
                                                                      long long* rel) {
    __shared__ long long part[kFinThreads / 64];
    const uint32_t r = blockIdx.y;
    const int e = exp[r];
    const uint64_t base = static_cast<uint64_t>(r) * cpr;
    long long sum = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kFinThreads;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * kFinThreads + threadIdx.x; c < cpr; c += stride) {
        long long v = 0;
        for (int k = 0; k < copies; ++k) v += fixed[static_cast<uint64_t>(k) * n_cells + base + c];
        sum += v;
        slab[base + c] = ldexpf(static_cast<float>(v), -e);
    }
    if (rel == nullptr) return;
    sum = wave_sum_i64(sum);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kFinThreads / 64; ++w) t += part[w];
        if (t != 0) atomicAdd(reinterpret_cast<unsigned long long*>(rel + r), static_cast<unsigned long long>(t));
    }
}

__global__ void fixed_totals_kernel(const long long* rel, const int32_t* exp, int n_r, double* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_r) out[r] = ldexp(static_cast<double>(rel[r]), -exp[r]);
}

}  // namespace glide

using namespace glide;

extern "C" size_t glide_fixed_workspace_bytes(int32_t n_r) {
    return static_cast<size_t>(n_r > 0 ? n_r : 0) * sizeof(long long);
}

extern "C" int glide_fixed_finalize(const long long* fixed, int32_t copies, int64_t n_cells, uint32_t cells_per_release,
                                    const int32_t* fixed_exp, float* slab, double* release_total, void* workspace,
                                    size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(fixed && fixed_exp && slab, "glide_fixed_finalize: NULL fixed/fixed_exp/slab");
    GLIDE_REQUIRE(n_cells > 0 && n_cells < 0xffffffffll, "bad n_cells");
    GLIDE_REQUIRE(copies >= 0 && copies <= 1024, "copies must be in [0, 1024]");
    if (copies == 0) copies = 1;
    GLIDE_REQUIRE(cells_per_release > 0 && n_cells % cells_per_release == 0, "cells_per_release must divide n_cells");
    const int64_t n_r = n_cells / cells_per_release;
    GLIDE_REQUIRE(n_r <= 65535, "at most 65535 releases per footprint");
    hipStream_t st = static_cast<hipStream_t>(stream);
    long long* rel = nullptr;
    if (release_total != nullptr) {
        GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= glide_fixed_workspace_bytes(static_cast<int32_t>(n_r)),
                      "workspace too small: need %zu bytes", glide_fixed_workspace_bytes(static_cast<int32_t>(n_r)));
        rel = static_cast<long long*>(workspace);
        hipLaunchKernelGGL(fixed_zero_kernel, dim3(static_cast<unsigned>((n_r + 255) / 256)), dim3(256), 0, st, rel,
                           static_cast<int>(n_r));
    }
    const uint64_t per_block = static_cast<uint64_t>(kFinThreads) * kFinItems;
    uint64_t bx = (cells_per_release + per_block - 1) / per_block;
    if (bx > 4096) bx = 4096;
    hipLaunchKernelGGL(fixed_finalize_kernel, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(n_r)),
                       dim3(kFinThreads), 0, st, fixed, static_cast<int>(copies), static_cast<uint64_t>(n_cells),
                       cells_per_release, fixed_exp, slab, rel);
    if (release_total != nullptr) {
        hipLaunchKernelGGL(fixed_totals_kernel, dim3(static_cast<unsigned>((n_r + 255) / 256)), dim3(256), 0, st, rel,
                           fixed_exp, static_cast<int>(n_r), release_total);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
