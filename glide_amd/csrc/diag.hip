// Per-step trajectory diagnostics: the mean particle position of the batch
// (reference main.py:1565, particles[:, :3].mean over every particle of the
// batch) and the grid-mean wind of a resampled met hour (main.py:892-937,
// computed once per hour and interpolated in time on the host).
//
// Both are fixed-order reductions in float64: each workgroup sums a fixed
// contiguous chunk (per-thread strided sums, a butterfly inside the wave, the
// waves in order), a second single-workgroup kernel adds the partials in
// workgroup order. The grid size depends on n only, so a result is the same on
// every run. HBM-bound: 4 bytes per element per channel, read once.
#include "common.h"

namespace glide {

constexpr int kMeanThreads = 256;
constexpr int kMeanMaxBlocks = 512;
constexpr int kMeanCh = 4;

struct MeanArgs {
    const float* x[kMeanCh];
    int64_t n;
    int nch;
    double* partial;  // [blocks][kMeanCh]
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kMeanThreads) void mean_partial_kernel(const MeanArgs a) {
    __shared__ double red[kMeanThreads / 64][kMeanCh];
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t hi = lo + per < a.n ? lo + per : a.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < a.nch; ++c) {
        double s = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += kMeanThreads) s += static_cast<double>(a.x[c][i]);
        s = wave_sum(s);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < a.nch) {
        double s = 0.0;
        for (int w = 0; w < kMeanThreads / 64; ++w) s += red[w][threadIdx.x];
        a.partial[static_cast<int64_t>(blockIdx.x) * kMeanCh + threadIdx.x] = s;
    }
}

// One workgroup: out[r * nch + c] = (sum of partials in block order) / n, with
// r = row[0] - 1 when row is given (the step the step kernel has just
// finished; skipped when outside [0, n_rows)), else r = 0.
__global__ __launch_bounds__(kMeanThreads) void mean_final_kernel(const double* partial, int blocks, int nch,
                                                                  int64_t n, const int32_t* row, int32_t n_rows,
                                                                  double* out) {
    __shared__ double red[kMeanThreads / 64][kMeanCh];
    int r = 0;
    if (row != nullptr) {
        r = row[0] - 1;
        if (r < 0 || r >= n_rows) return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < nch; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < blocks; b += kMeanThreads) s += partial[static_cast<int64_t>(b) * kMeanCh + c];
        s = wave_sum(s);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < nch) {
        double s = 0.0;
        for (int w = 0; w < kMeanThreads / 64; ++w) s += red[w][threadIdx.x];
        out[static_cast<int64_t>(r) * nch + threadIdx.x] = s / static_cast<double>(n);
    }
}

static int launch_means(MeanArgs a, const int32_t* row, int32_t n_rows, double* out, void* ws, size_t ws_bytes,
                        hipStream_t st) {
    GLIDE_REQUIRE(out != nullptr && ws != nullptr, "NULL output/workspace");
    GLIDE_REQUIRE(ws_bytes >= glide_means_workspace_bytes(), "means workspace too small (%zu < %zu bytes)", ws_bytes,
                  glide_means_workspace_bytes());
    GLIDE_REQUIRE(a.n > 0, "means of an empty set");
    GLIDE_REQUIRE(row == nullptr || n_rows > 0, "row-indexed means need n_rows > 0");
    for (int c = 0; c < a.nch; ++c) GLIDE_REQUIRE(a.x[c] != nullptr, "NULL input channel %d", c);
    int64_t blocks = (a.n + 4 * kMeanThreads - 1) / (4 * kMeanThreads);
    if (blocks > kMeanMaxBlocks) blocks = kMeanMaxBlocks;
    a.partial = static_cast<double*>(ws);
    hipLaunchKernelGGL(mean_partial_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kMeanThreads), 0, st, a);
    GLIDE_HIP(hipGetLastError());
    hipLaunchKernelGGL(mean_final_kernel, dim3(1), dim3(kMeanThreads), 0, st, static_cast<const double*>(ws),
                       static_cast<int>(blocks), a.nch, a.n, row, n_rows, out);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

}  // namespace glide

using namespace glide;

extern "C" size_t glide_means_workspace_bytes(void) {
    return static_cast<size_t>(kMeanMaxBlocks) * kMeanCh * sizeof(double);
}

extern "C" int glide_particle_means(const GlideParticles* s, const int32_t* row, int32_t n_rows, double* out,
                                    void* workspace, size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(s != nullptr, "NULL particles");
    GLIDE_REQUIRE(s->pos_stride == 1, "particle means need pos_stride 1");
    GLIDE_REQUIRE(s->cell == nullptr, "particle means take absolute positions (cell is not NULL)");
    MeanArgs a{};
    a.x[0] = s->lon;
    a.x[1] = s->lat;
    a.x[2] = s->z;
    a.n = s->n;
    a.nch = 3;
    return launch_means(a, row, n_rows, out, workspace, workspace_bytes, static_cast<hipStream_t>(stream));
}

extern "C" int glide_field_means(const float* fields, int64_t n_per_channel, int32_t n_channels, double* out,
                                 void* workspace, size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(fields != nullptr, "NULL fields");
    GLIDE_REQUIRE(n_channels >= 1 && n_channels <= kMeanCh, "1 to %d channels", kMeanCh);
    MeanArgs a{};
    for (int c = 0; c < n_channels; ++c) a.x[c] = fields + static_cast<int64_t>(c) * n_per_channel;
    a.n = n_per_channel;
    a.nch = n_channels;
    return launch_means(a, nullptr, 0, out, workspace, workspace_bytes, static_cast<hipStream_t>(stream));
}
