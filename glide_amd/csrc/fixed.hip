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

// Where release r's cells lie: flat (r * cpr, cpr cells, out at the same
// index) or per-release slabs (slab_off[r], slab_nt[r] * cpb cells, out from
// slab_off[r] - slab_off[r0]: the range's slabs back to back).
struct FinLayout {
    uint64_t acc_cells;  // stride between the accumulator's copies
    uint32_t cpr;        // flat: cells per release
    const int64_t* slab_off;
    const int32_t* slab_nt;
    uint32_t cpb;        // slabs: cells per time bin
    int32_t r0;          // first release of the range
};

// Block (bx, j): cells bx*256 + t (+ grid stride) of release r0 + j.
__global__ __launch_bounds__(kFinThreads) void fixed_finalize_kernel(const long long* fixed, int copies, FinLayout L,
                                                                      const int32_t* exp, float* slab,
                                                                      long long* rel) {
    __shared__ long long part[kFinThreads / 64];
    const int32_t r = L.r0 + static_cast<int32_t>(blockIdx.y);
    const int e = exp[r];
    uint64_t base, cnt, obase;
    if (L.slab_off != nullptr) {
        base = static_cast<uint64_t>(L.slab_off[r]);
        cnt = static_cast<uint64_t>(L.slab_nt[r]) * L.cpb;
        obase = base - static_cast<uint64_t>(L.slab_off[L.r0]);
    } else {
        base = static_cast<uint64_t>(r) * L.cpr;
        cnt = L.cpr;
        obase = base;
    }
    long long sum = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kFinThreads;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * kFinThreads + threadIdx.x; c < cnt; c += stride) {
        // the copies' loads go out eight at a time (integer sums: any order is exact)
        const long long* q = fixed + base + c;
        long long v = 0;
        int k = 0;
        for (; k + 8 <= copies; k += 8) {
            long long t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = q[static_cast<uint64_t>(k + u) * L.acc_cells];
#pragma unroll
            for (int u = 0; u < 8; ++u) v += t[u];
        }
        for (; k < copies; ++k) v += q[static_cast<uint64_t>(k) * L.acc_cells];
        sum += v;
        slab[obase + c] = ldexpf(static_cast<float>(v), -e);
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

__global__ void fixed_totals_kernel(const long long* rel, const int32_t* exp, int r0, int r1, double* out) {
    const int r = r0 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < r1) out[r] = ldexp(static_cast<double>(rel[r]), -exp[r]);
}

__global__ void fixed_zero_range_kernel(long long* w, int r0, int r1) {
    const int r = r0 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < r1) w[r] = 0;
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
    const FinLayout lay{static_cast<uint64_t>(n_cells), cells_per_release, nullptr, nullptr, 0u, 0};
    hipLaunchKernelGGL(fixed_finalize_kernel, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(n_r)),
                       dim3(kFinThreads), 0, st, fixed, static_cast<int>(copies), lay, fixed_exp, slab, rel);
    if (release_total != nullptr) {
        hipLaunchKernelGGL(fixed_totals_kernel, dim3(static_cast<unsigned>((n_r + 255) / 256)), dim3(256), 0, st, rel,
                           fixed_exp, 0, static_cast<int>(n_r), release_total);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_fixed_finalize_slabs(const long long* fixed, int32_t copies, int64_t slab_cells,
                                          const int64_t* slab_off, const int32_t* slab_nt, uint32_t cells_per_bin,
                                          int32_t max_nt, const int32_t* fixed_exp, int32_t r0, int32_t r1, float* out,
                                          double* release_total, void* workspace, size_t workspace_bytes,
                                          void* stream) {
    GLIDE_REQUIRE(fixed && slab_off && slab_nt && fixed_exp && out, "glide_fixed_finalize_slabs: NULL argument");
    GLIDE_REQUIRE(slab_cells > 0, "bad slab_cells");
    GLIDE_REQUIRE(copies >= 0 && copies <= 1024, "copies must be in [0, 1024]");
    if (copies == 0) copies = 1;
    GLIDE_REQUIRE(cells_per_bin > 0 && max_nt > 0, "bad slab geometry");
    GLIDE_REQUIRE(r0 >= 0 && r1 >= r0 && r1 - r0 <= 65535, "bad release range [%d, %d)", r0, r1);
    if (r1 == r0) return GLIDE_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    long long* rel = nullptr;
    const unsigned nb = static_cast<unsigned>((r1 - r0 + 255) / 256);
    if (release_total != nullptr) {
        GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= glide_fixed_workspace_bytes(r1),
                      "workspace too small: need %zu bytes", glide_fixed_workspace_bytes(r1));
        rel = static_cast<long long*>(workspace);
        hipLaunchKernelGGL(fixed_zero_range_kernel, dim3(nb), dim3(256), 0, st, rel, r0, r1);
    }
    const uint64_t per_block = static_cast<uint64_t>(kFinThreads) * kFinItems;
    uint64_t bx = (static_cast<uint64_t>(max_nt) * cells_per_bin + per_block - 1) / per_block;
    if (bx > 4096) bx = 4096;
    const FinLayout lay{static_cast<uint64_t>(slab_cells), 0u, slab_off, slab_nt, cells_per_bin, r0};
    hipLaunchKernelGGL(fixed_finalize_kernel, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(r1 - r0)),
                       dim3(kFinThreads), 0, st, fixed, static_cast<int>(copies), lay, fixed_exp, out, rel);
    if (release_total != nullptr) {
        hipLaunchKernelGGL(fixed_totals_kernel, dim3(nb), dim3(256), 0, st, rel, fixed_exp, r0, r1, release_total);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
