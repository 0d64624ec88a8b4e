// Deterministic footprint accumulation (replaces scatter_add_ float atomics,
// reference footprint_gridder.py:197).
//
//  1. Stable LSD radix sort of (key, value) pairs by key, 8-bit digits over the
//     bits the key range needs. Per pass: per-tile digit histograms, one
//     exclusive scan in digit-major order, stable per-tile scatter (wavefront
//     ballot ranking, waves in order, rounds in order).
//  2. Segmented reduction of the sorted values: sequential per thread, a
//     fixed-shape segmented scan per tile, and a per-tile chain for segments
//     that cross tiles. Every cell's sum is formed in one fixed order, so the
//     result does not depend on scheduling.
//  3. One read-modify-write per touched cell: slab[key] += sum, optionally with
//     a Neumaier compensation term kept in comp[key].
#include "common.h"

namespace glide {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kRadix = 256;
constexpr int kRedThreads = 256;
constexpr int kRedItems = 8;
constexpr int kRedTile = kRedThreads * kRedItems;  // 2048
constexpr int kScanThreads = 1024;

struct Workspace {
    uint32_t* keys_a;
    float* vals_a;
    uint32_t* keys_b;
    float* vals_b;
    uint32_t* hist;  // [kRadix][n_sort_tiles]
    float* head_part;  // [n_red_tiles] sum before the tile's first head (whole tile if none)
    float* tail_part;  // [n_red_tiles] open segment at the tile end (from its last head)
    uint32_t* first_key;  // [n_red_tiles] key of the tile's first item
    uint32_t* flags;  // bit0 any head, bit1 open segment continues, bit2 first item is a head
};

static inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

static size_t layout(int64_t n, Workspace* w, char* base) {
    const int64_t ts = div_up(n, kSortTile);
    const int64_t tr = div_up(n, kRedTile);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes > 0 ? bytes : 1);
        return p;
    };
    char* ka = take(sizeof(uint32_t) * n);
    char* va = take(sizeof(float) * n);
    char* kb = take(sizeof(uint32_t) * n);
    char* vb = take(sizeof(float) * n);
    char* hi = take(sizeof(uint32_t) * kRadix * ts);
    char* hp = take(sizeof(float) * tr);
    char* tp = take(sizeof(float) * tr);
    char* fk = take(sizeof(uint32_t) * tr);
    char* fl = take(sizeof(uint32_t) * tr);
    if (w) {
        w->keys_a = reinterpret_cast<uint32_t*>(ka);
        w->vals_a = reinterpret_cast<float*>(va);
        w->keys_b = reinterpret_cast<uint32_t*>(kb);
        w->vals_b = reinterpret_cast<float*>(vb);
        w->hist = reinterpret_cast<uint32_t*>(hi);
        w->head_part = reinterpret_cast<float*>(hp);
        w->tail_part = reinterpret_cast<float*>(tp);
        w->first_key = reinterpret_cast<uint32_t*>(fk);
        w->flags = reinterpret_cast<uint32_t*>(fl);
    }
    return off;
}

// --------------------------------------------------------------------------
// Radix sort kernels
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                                  int shift, uint32_t mask,
                                                                  uint32_t* __restrict__ hist,
                                                                  int64_t n_tiles) {
    __shared__ uint32_t cnt[kRadix];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = base + r * kSortThreads + threadIdx.x;
        if (j < n) atomicAdd(&cnt[(keys[j] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * n_tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of m counts in place, one workgroup.
__global__ __launch_bounds__(kScanThreads) void exclusive_scan_kernel(uint32_t* __restrict__ a, int64_t m) {
    __shared__ uint32_t part[kScanThreads];
    const int64_t per = (m + kScanThreads - 1) / kScanThreads;
    const int64_t lo = threadIdx.x * per;
    const int64_t hi = lo + per < m ? lo + per : m;
    uint32_t s = 0;
    for (int64_t j = lo; j < hi; ++j) s += a[j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t v = a[j];
        a[j] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const float* __restrict__ vin, uint32_t* __restrict__ kout,
    float* __restrict__ vout, int64_t n, int shift, int bits, const uint32_t* __restrict__ offs,
    int64_t n_tiles) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const uint32_t mask = (1u << bits) - 1u;
    base[t] = offs[static_cast<int64_t>(t) * n_tiles + blockIdx.x];
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = tile0 + r * kSortThreads + t;
        const bool valid = j < n;
        uint32_t key = 0;
        float val = 0.0f;
        uint32_t d = 0;
        if (valid) {
            key = kin[j];
            val = vin[j];
            d = (key >> shift) & mask;
        }
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lt);
        const int total = __popcll(peers);
        if (valid && rank == 0) wcnt[wave][d] = static_cast<uint32_t>(total);
        __syncthreads();
        {
            uint32_t run = base[t];
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
            base[t] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = wcnt[wave][d] + static_cast<uint32_t>(rank);
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Segmented reduction
// --------------------------------------------------------------------------
__device__ __forceinline__ void add_cell(float* slab, float* comp, uint32_t key, float x) {
#pragma clang fp contract(off)
    if (comp == nullptr) {
        slab[key] += x;
        return;
    }
    // Neumaier: keep the low-order part lost by slab + x in comp.
    const float s = slab[key];
    const float t = s + x;
    const float c = (fabsf(s) >= fabsf(x)) ? (s - t) + x : (x - t) + s;
    slab[key] = t;
    comp[key] += c;
}

// One tile of sorted items. Emits every segment that starts and ends inside
// the tile; records the tile's leading continuation sum (head_part), its open
// trailing sum (tail_part) and flags for segments that cross tile boundaries.
__global__ __launch_bounds__(kRedThreads) void seg_reduce_tiles_kernel(
    const uint32_t* __restrict__ keys, const float* __restrict__ vals, int64_t n, uint32_t n_cells,
    float* __restrict__ slab, float* __restrict__ comp, Workspace w) {
#pragma clang fp contract(off)
    __shared__ uint32_t sk[kRedTile];
    __shared__ float sv[kRedTile];
    __shared__ float scan_v[kRedThreads];
    __shared__ uint32_t scan_f[kRedThreads];
    __shared__ uint32_t edge_key[2];  // key before the tile, key after the tile
    const int t = threadIdx.x;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRedTile;
    const int cnt = static_cast<int>((n - tile0) < kRedTile ? (n - tile0) : kRedTile);
    for (int q = 0; q < kRedItems; ++q) {
        const int idx = q * kRedThreads + t;
        if (idx < cnt) {
            sk[idx] = keys[tile0 + idx];
            sv[idx] = vals[tile0 + idx];
        }
    }
    if (t == 0) {
        edge_key[0] = tile0 > 0 ? keys[tile0 - 1] : 0xffffffffu;
        edge_key[1] = (tile0 + cnt < n) ? keys[tile0 + cnt] : 0xffffffffu;
    }
    __syncthreads();
    const bool first_tile = tile0 == 0;
    const bool last_tile = tile0 + cnt >= n;
    const bool item0_head = first_tile || sk[0] != edge_key[0];

    // Sequential pass over this thread's consecutive items.
    float acc = 0.0f;
    float first_part = 0.0f;
    bool has_head = false;
    uint32_t cur_key = 0;
    const int i0 = t * kRedItems;
    for (int q = 0; q < kRedItems; ++q) {
        const int idx = i0 + q;
        if (idx >= cnt) break;
        const uint32_t k = sk[idx];
        const bool head = (idx == 0) ? item0_head : (k != sk[idx - 1]);
        if (head) {
            if (has_head) {
                if (cur_key < n_cells) add_cell(slab, comp, cur_key, acc);
            } else {
                first_part = acc;
            }
            has_head = true;
            acc = 0.0f;
            cur_key = k;
        }
        acc += sv[idx];
    }
    if (!has_head) first_part = acc;

    // Inclusive segmented scan over threads (fixed shape).
    scan_v[t] = acc;
    scan_f[t] = has_head ? 1u : 0u;
    __syncthreads();
    for (int off = 1; off < kRedThreads; off <<= 1) {
        float v = 0.0f;
        uint32_t f = 0;
        if (t >= off) {
            v = scan_v[t - off];
            f = scan_f[t - off];
        }
        __syncthreads();
        if (t >= off) {
            if (!scan_f[t]) scan_v[t] = v + scan_v[t];
            scan_f[t] |= f;
        }
        __syncthreads();
    }
    if (has_head) {
        // The segment that closes just before this thread's first head.
        const float s_prev = t > 0 ? scan_v[t - 1] : 0.0f;
        const bool head_before = t > 0 && scan_f[t - 1] != 0;
        const float tot = s_prev + first_part;
        if (head_before) {
            const uint32_t k = sk[i0 - 1];
            if (k < n_cells) add_cell(slab, comp, k, tot);
        } else if (!item0_head) {
            w.head_part[blockIdx.x] = tot;  // continuation of an earlier tile's segment
        }
    }
    if (t == kRedThreads - 1) {
        const uint32_t lk = sk[cnt - 1];
        const bool next_head = last_tile || edge_key[1] != lk;
        const bool any_head = scan_f[t] != 0;
        const float open = scan_v[t];
        uint32_t fl = (any_head ? 1u : 0u) | (next_head ? 0u : 2u) | (item0_head ? 4u : 0u);
        if (!any_head) {
            w.head_part[blockIdx.x] = open;  // the whole tile continues one segment
        } else if (next_head) {
            if (lk < n_cells) add_cell(slab, comp, lk, open);
        } else {
            w.tail_part[blockIdx.x] = open;
        }
        w.first_key[blockIdx.x] = sk[0];
        w.flags[blockIdx.x] = fl;
    }
}

// Segments crossing tiles. With P_j the open-segment sum at the end of tile j
// (P_j = tail_j if tile j has a head, else P_{j-1} + head_j), the segment that
// closes in tile c has total P_{c-1} + head_c. One workgroup scans the tiles in
// fixed chunks, so the order of additions is fixed.
__global__ __launch_bounds__(kScanThreads) void seg_fixup_kernel(int64_t n_tiles, uint32_t n_cells,
                                                                 float* __restrict__ slab,
                                                                 float* __restrict__ comp, Workspace w) {
#pragma clang fp contract(off)
    __shared__ float agg_v[kScanThreads];
    __shared__ uint32_t agg_f[kScanThreads];
    const int t = threadIdx.x;
    const int64_t per = (n_tiles + kScanThreads - 1) / kScanThreads;
    const int64_t lo = t * per;
    const int64_t hi = lo + per < n_tiles ? lo + per : n_tiles;
    // Chunk aggregate: (any tile with a head, open sum at the chunk end).
    float p = 0.0f;
    uint32_t f = 0;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t fl = w.flags[j];
        if (fl & 1u) {
            p = (fl & 2u) ? w.tail_part[j] : 0.0f;
            f = 1u;
        } else {
            p = p + w.head_part[j];
            if (!(fl & 2u)) {
                p = 0.0f;
                f = 1u;
            }
        }
    }
    agg_v[t] = p;
    agg_f[t] = f;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        float v = 0.0f;
        uint32_t g = 0;
        if (t >= off) {
            v = agg_v[t - off];
            g = agg_f[t - off];
        }
        __syncthreads();
        if (t >= off) {
            if (!agg_f[t]) agg_v[t] = v + agg_v[t];
            agg_f[t] |= g;
        }
        __syncthreads();
    }
    float run = t > 0 ? agg_v[t - 1] : 0.0f;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t fl = w.flags[j];
        if (fl & 1u) {
            if (!(fl & 4u)) {
                const uint32_t k = w.first_key[j];
                if (k < n_cells) add_cell(slab, comp, k, run + w.head_part[j]);
            }
            run = (fl & 2u) ? w.tail_part[j] : 0.0f;
        } else {
            run = run + w.head_part[j];
            if (!(fl & 2u)) {
                // A pass-through tile whose segment ends exactly at its end.
                const uint32_t k = w.first_key[j];
                if (k < n_cells) add_cell(slab, comp, k, run);
                run = 0.0f;
            }
        }
    }
}

// Per-release residence totals in float64 over the sorted stream: release r
// owns keys [r * cpr, (r + 1) * cpr); one block per release sums its range in a
// fixed strided order and tree.
__global__ __launch_bounds__(256) void release_total_kernel(const uint32_t* __restrict__ keys,
                                                            const float* __restrict__ vals, int64_t n,
                                                            uint32_t n_cells, uint32_t cpr,
                                                            double* __restrict__ total) {
    __shared__ double part[256];
    const uint32_t r = blockIdx.x;
    const uint64_t klo = static_cast<uint64_t>(r) * cpr;
    uint64_t khi = klo + cpr;
    if (khi > n_cells) khi = n_cells;
    auto lower = [&](uint64_t key) {
        int64_t a = 0, b = n;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (static_cast<uint64_t>(keys[m]) < key) a = m + 1; else b = m;
        }
        return a;
    };
    const int64_t lo = lower(klo);
    const int64_t hi = lower(khi);
    double s = 0.0;
    for (int64_t j = lo + threadIdx.x; j < hi; j += blockDim.x) s += static_cast<double>(vals[j]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[r] += part[0];
}

}  // namespace glide

using namespace glide;

extern "C" size_t glide_footprint_workspace_bytes(int64_t n) {
    if (n < 0) n = 0;
    return layout(n, nullptr, nullptr);
}

extern "C" int glide_footprint_accumulate(const uint32_t* keys, const float* vals, int64_t n,
                                          uint32_t n_cells, uint32_t cells_per_release, float* slab,
                                          float* comp, double* release_total, void* workspace,
                                          size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return GLIDE_OK;
    GLIDE_REQUIRE(keys && vals && slab, "NULL keys/vals/slab");
    GLIDE_REQUIRE(n_cells > 0 && n_cells < 0xffffffffu, "bad n_cells");
    GLIDE_REQUIRE(n < (1ll << 32) - 1, "too many keys (positions must fit 32 bits)");
    const size_t need = layout(n, nullptr, nullptr);
    GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= need,
                  "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    if (release_total != nullptr) {
        GLIDE_REQUIRE(cells_per_release > 0 && n_cells % cells_per_release == 0,
                      "cells_per_release must divide n_cells");
    }
    Workspace w{};
    layout(n, &w, static_cast<char*>(workspace));
    hipStream_t st = static_cast<hipStream_t>(stream);

    // Key bits: enough to represent the sentinel n_cells.
    int key_bits = 1;
    while (key_bits < 32 && (static_cast<uint64_t>(n_cells) >> key_bits) != 0) ++key_bits;
    const int64_t ts = div_up(n, kSortTile);
    const int npass = (key_bits + 7) / 8;
    const uint32_t* kin = keys;
    const float* vin = vals;
    for (int ps = 0; ps < npass; ++ps) {
        const int shift = ps * 8;
        const int bits = (key_bits - shift) < 8 ? (key_bits - shift) : 8;
        const uint32_t mask = (1u << bits) - 1u;
        uint32_t* kout = (ps & 1) ? w.keys_b : w.keys_a;
        float* vout = (ps & 1) ? w.vals_b : w.vals_a;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin, n,
                           shift, mask, w.hist, ts);
        hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(kScanThreads), 0, st, w.hist,
                           static_cast<int64_t>(kRadix) * ts);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin,
                           vin, kout, vout, n, shift, bits, w.hist, ts);
        kin = kout;
        vin = vout;
    }
    GLIDE_HIP(hipGetLastError());
    const int64_t tr = div_up(n, kRedTile);
    hipLaunchKernelGGL(seg_reduce_tiles_kernel, dim3(static_cast<unsigned>(tr)), dim3(kRedThreads), 0, st, kin,
                       vin, n, n_cells, slab, comp, w);
    hipLaunchKernelGGL(seg_fixup_kernel, dim3(1), dim3(kScanThreads), 0, st, tr, n_cells, slab, comp, w);
    if (release_total != nullptr) {
        const uint32_t nr = n_cells / cells_per_release;
        hipLaunchKernelGGL(release_total_kernel, dim3(nr), dim3(256), 0, st, kin, vin, n, n_cells,
                           cells_per_release, release_total);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
