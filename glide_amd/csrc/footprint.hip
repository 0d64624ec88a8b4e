// Deterministic footprint accumulation (replaces scatter_add_ float atomics,
// reference footprint_gridder.py:197).
//
//  1. Stable radix sort of the (cell key, value) pairs by key; its first pass
//     drops the pairs whose key is not a cell (key >= n_cells), keeping order.
//  2. Segmented reduction of the sorted values: sequential per thread, a
//     fixed-shape segmented scan per tile, and one workgroup that scans the
//     tiles for segments crossing tile boundaries. Every cell's sum is formed in
//     one fixed order, so the result does not depend on scheduling.
//  3. One read-modify-write per touched cell: slab[key] += sum, optionally with
//     a Neumaier compensation term kept in comp[key].
//  4. Optionally, in the same launches, the reduction keyed by release
//     (key / cells_per_release) accumulates each release's residence total in
//     float64.
//
// Item counts after the first sort pass live in device memory; kernels are
// launched for the upper bound and tiles past the count exit.
#include "sort.h"

#include <mutex>

namespace glide {

constexpr int kRedThreads = 256;
constexpr int kRedItems = 8;
constexpr int kRedTile = kRedThreads * kRedItems;  // 2048
constexpr int kScanThreads = 1024;

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct Bump {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += align256(sizeof(T) * (count > 0 ? count : 1));
        return p;
    }
};

// --------------------------------------------------------------------------
// Segmented reduction, by cell (float sums into the slab) or by release
// (float64 sums into the release totals).
// --------------------------------------------------------------------------
struct RedWork {
    double* head_part;    // [n_red_tiles] sum before the tile's first head (whole tile if none)
    double* tail_part;    // [n_red_tiles] open segment at the tile end (from its last head)
    uint32_t* first_key;  // [n_red_tiles] segment key of the tile's first item
    uint32_t* flags;      // bit0 any head, bit1 open segment continues, bit2 first item is a head
};

struct CellTarget {
    using Acc = float;
    uint32_t n_cells;  // key space: flat cells, or n_r * cells_per_bin for release-binned keys
    float* slab;
    float* comp;
    const int32_t* t_bin;  // release-binned keys: slab cell (r * n_t + t_bin[r]) * cpb + k % cpb
    uint32_t cpb, n_t;
    __device__ __forceinline__ uint32_t seg(uint32_t key) const { return key; }
    __device__ __forceinline__ bool valid(uint32_t s) const { return s < n_cells; }
    __device__ __forceinline__ void emit(uint32_t k, float x) const {
#pragma clang fp contract(off)
        uint32_t key = k;
        if (t_bin != nullptr) {
            const uint32_t r = k / cpb;
            const int32_t tb = t_bin[r];
            if (tb < 0 || static_cast<uint32_t>(tb) >= n_t) return;  // never written this step: no such keys
            key = (r * n_t + static_cast<uint32_t>(tb)) * cpb + (k - r * cpb);
        }
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
};

struct ReleaseTarget {
    using Acc = double;
    uint32_t n_cells, cpr, n_rel;
    double* total;
    __device__ __forceinline__ uint32_t seg(uint32_t key) const { return key < n_cells ? key / cpr : n_rel; }
    __device__ __forceinline__ bool valid(uint32_t s) const { return s < n_rel; }
    __device__ __forceinline__ void emit(uint32_t r, double x) const { total[r] += x; }
};

// Segmented reduction of one tile (keys and values staged in LDS) for target
// T. edge_raw holds the raw keys just before and just after the tile.
template <typename T>
__device__ __forceinline__ void reduce_tile(const T& tg, const RedWork& w, const uint32_t* sk, const float* sv,
                                            int cnt, bool first_tile, bool last_tile, const uint32_t* edge_raw,
                                            typename T::Acc* scan_v, uint32_t* scan_f) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    const int t = threadIdx.x;
    const uint32_t e0 = tg.seg(edge_raw[0]);
    const uint32_t e1 = tg.seg(edge_raw[1]);
    const bool item0_head = first_tile || tg.seg(sk[0]) != e0;

    // Sequential pass over this thread's consecutive items.
    Acc acc = 0;
    Acc first_part = 0;
    bool has_head = false;
    uint32_t cur_key = 0;
    const int i0 = t * kRedItems;
    uint32_t prev = (i0 > 0 && i0 < cnt) ? tg.seg(sk[i0 - 1]) : 0u;
    for (int q = 0; q < kRedItems; ++q) {
        const int idx = i0 + q;
        if (idx >= cnt) break;
        const uint32_t k = tg.seg(sk[idx]);
        const bool head = (idx == 0) ? item0_head : (k != prev);
        prev = k;
        if (head) {
            if (has_head) {
                if (tg.valid(cur_key)) tg.emit(cur_key, acc);
            } else {
                first_part = acc;
            }
            has_head = true;
            acc = 0;
            cur_key = k;
        }
        acc += static_cast<Acc>(sv[idx]);
    }
    if (!has_head) first_part = acc;

    // Inclusive segmented scan over threads (fixed shape).
    scan_v[t] = acc;
    scan_f[t] = has_head ? 1u : 0u;
    __syncthreads();
    for (int off = 1; off < kRedThreads; off <<= 1) {
        Acc v = 0;
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
        const Acc s_prev = t > 0 ? scan_v[t - 1] : Acc(0);
        const bool head_before = t > 0 && scan_f[t - 1] != 0;
        const Acc tot = s_prev + first_part;
        if (head_before) {
            const uint32_t k = tg.seg(sk[i0 - 1]);
            if (tg.valid(k)) tg.emit(k, tot);
        } else if (!item0_head) {
            w.head_part[blockIdx.x] = static_cast<double>(tot);  // continuation of an earlier tile's segment
        }
    }
    if (t == kRedThreads - 1) {
        const uint32_t lk = tg.seg(sk[cnt - 1]);
        const bool next_head = last_tile || e1 != lk;
        const bool any_head = scan_f[t] != 0;
        const Acc open = scan_v[t];
        const uint32_t fl = (any_head ? 1u : 0u) | (next_head ? 0u : 2u) | (item0_head ? 4u : 0u);
        if (!any_head) {
            w.head_part[blockIdx.x] = static_cast<double>(open);  // the whole tile continues one segment
        } else if (next_head) {
            if (tg.valid(lk)) tg.emit(lk, open);
        } else {
            w.tail_part[blockIdx.x] = static_cast<double>(open);
        }
        w.first_key[blockIdx.x] = tg.seg(sk[0]);
        w.flags[blockIdx.x] = fl;
    }
}

// One launch reduces each tile by cell and, with kRel, by release.
template <bool kRel>
__global__ __launch_bounds__(kRedThreads) void seg_reduce_tiles_kernel(const uint32_t* __restrict__ keys,
                                                                       const float* __restrict__ vals,
                                                                       const uint32_t* __restrict__ n_dev,
                                                                       CellTarget ct, ReleaseTarget rt, RedWork wc,
                                                                       RedWork wr) {
    __shared__ uint32_t sk[kRedTile];
    __shared__ float sv[kRedTile];
    __shared__ float scan32[kRedThreads];
    __shared__ double scan64[kRedThreads];
    __shared__ uint32_t scan_f[kRedThreads];
    __shared__ uint32_t edge_raw[2];
    const int64_t n = *n_dev;
    const int t = threadIdx.x;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRedTile;
    if (tile0 >= n) return;  // uniform across the block
    const int cnt = static_cast<int>((n - tile0) < kRedTile ? (n - tile0) : kRedTile);
    for (int q = 0; q < kRedItems; ++q) {
        const int idx = q * kRedThreads + t;
        if (idx < cnt) {
            sk[idx] = keys[tile0 + idx];
            sv[idx] = vals[tile0 + idx];
        }
    }
    if (t == 0) {
        edge_raw[0] = tile0 > 0 ? keys[tile0 - 1] : 0u;
        edge_raw[1] = (tile0 + cnt < n) ? keys[tile0 + cnt] : 0u;
    }
    __syncthreads();
    const bool first_tile = tile0 == 0;
    const bool last_tile = tile0 + cnt >= n;
    reduce_tile(ct, wc, sk, sv, cnt, first_tile, last_tile, edge_raw, scan32, scan_f);
    if (kRel) {
        __syncthreads();
        reduce_tile(rt, wr, sk, sv, cnt, first_tile, last_tile, edge_raw, scan64, scan_f);
    }
}

// Segments crossing tiles. With P_j the open-segment sum at the end of tile j
// (P_j = tail_j if tile j has a head, else P_{j-1} + head_j), the segment that
// closes in tile c has total P_{c-1} + head_c. One workgroup scans the tiles in
// fixed chunks, so the order of additions is fixed. Partial sums are held in
// the target's accumulator type (float sums are exact in the double slots).
template <typename T>
__device__ __forceinline__ void fixup_tiles(const T& tg, const RedWork& w, int64_t n_tiles, typename T::Acc* agg_v,
                                            uint32_t* agg_f) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    const int t = threadIdx.x;
    const int64_t per = (n_tiles + kScanThreads - 1) / kScanThreads;
    const int64_t lo = t * per;
    const int64_t hi = lo + per < n_tiles ? lo + per : n_tiles;
    auto head = [&](int64_t j) { return static_cast<Acc>(w.head_part[j]); };
    auto tail = [&](int64_t j) { return static_cast<Acc>(w.tail_part[j]); };
    Acc p = 0;
    uint32_t f = 0;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t fl = w.flags[j];
        if (fl & 1u) {
            p = (fl & 2u) ? tail(j) : Acc(0);
            f = 1u;
        } else {
            p = p + head(j);
            if (!(fl & 2u)) {
                p = 0;
                f = 1u;
            }
        }
    }
    agg_v[t] = p;
    agg_f[t] = f;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        Acc v = 0;
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
    Acc run = t > 0 ? agg_v[t - 1] : Acc(0);
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t fl = w.flags[j];
        if (fl & 1u) {
            if (!(fl & 4u)) {
                const uint32_t k = w.first_key[j];
                if (tg.valid(k)) tg.emit(k, run + head(j));
            }
            run = (fl & 2u) ? tail(j) : Acc(0);
        } else {
            run = run + head(j);
            if (!(fl & 2u)) {
                // A pass-through tile whose segment ends exactly at its end.
                const uint32_t k = w.first_key[j];
                if (tg.valid(k)) tg.emit(k, run);
                run = 0;
            }
        }
    }
}

// Workgroup 0 fixes the cell segments, workgroup 1 (if launched) the releases.
__global__ __launch_bounds__(kScanThreads) void seg_fixup_kernel(const uint32_t* __restrict__ n_dev, CellTarget ct,
                                                                 ReleaseTarget rt, RedWork wc, RedWork wr) {
    __shared__ float agg32[kScanThreads];
    __shared__ double agg64[kScanThreads];
    __shared__ uint32_t agg_f[kScanThreads];
    const int64_t n_tiles = div_up(static_cast<int64_t>(*n_dev), kRedTile);
    if (blockIdx.x == 0) {
        fixup_tiles(ct, wc, n_tiles, agg32, agg_f);
    } else {
        fixup_tiles(rt, wr, n_tiles, agg64, agg_f);
    }
}

static size_t red_layout(int64_t n, RedWork* rw, Bump& b) {
    const int64_t tr = div_up(n, kRedTile);
    RedWork r{};
    r.head_part = b.take<double>(tr);
    r.tail_part = b.take<double>(tr);
    r.first_key = b.take<uint32_t>(tr);
    r.flags = b.take<uint32_t>(tr);
    if (rw) *rw = r;
    return b.off;
}

static size_t footprint_layout(int64_t n, SortWork* sw, RedWork* wc, RedWork* wr, char* base) {
    const size_t s = sort_layout(n, sw, base);
    Bump b{base ? base + s : nullptr};
    red_layout(n, wc, b);
    red_layout(n, wr, b);
    return s + b.off;
}

}  // namespace glide

using namespace glide;

extern "C" size_t glide_footprint_workspace_bytes(int64_t n) {
    if (n < 0) n = 0;
    return footprint_layout(n, nullptr, nullptr, nullptr, nullptr);
}

static int accumulate(const uint32_t* keys, const float* vals, int64_t n, uint32_t n_cells,
                      uint32_t cells_per_release, const int32_t* t_bin, uint32_t n_t, float* slab, float* comp,
                      double* release_total, void* workspace, size_t workspace_bytes, void* stream);

extern "C" int glide_footprint_accumulate(const uint32_t* keys, const float* vals, int64_t n,
                                          uint32_t n_cells, uint32_t cells_per_release, float* slab,
                                          float* comp, double* release_total, void* workspace,
                                          size_t workspace_bytes, void* stream) {
    return accumulate(keys, vals, n, n_cells, cells_per_release, nullptr, 1u, slab, comp, release_total, workspace,
                      workspace_bytes, stream);
}

extern "C" int glide_footprint_accumulate_binned(const uint32_t* keys, const float* vals, int64_t n, uint32_t n_r,
                                                 uint32_t n_t, uint32_t cells_per_bin, const int32_t* t_bin,
                                                 float* slab, float* comp, double* release_total, void* workspace,
                                                 size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(n_r >= 1 && n_t >= 1 && cells_per_bin >= 1, "bad binned footprint dimensions");
    GLIDE_REQUIRE(static_cast<uint64_t>(n_r) * n_t * cells_per_bin < 0xffffffffull,
                  "footprint has too many cells for 32-bit indices");
    GLIDE_REQUIRE(t_bin != nullptr, "NULL t_bin");
    return accumulate(keys, vals, n, n_r * cells_per_bin, cells_per_bin, t_bin, n_t, slab, comp, release_total,
                      workspace, workspace_bytes, stream);
}

static int accumulate(const uint32_t* keys, const float* vals, int64_t n, uint32_t n_cells,
                      uint32_t cells_per_release, const int32_t* t_bin, uint32_t n_t, float* slab, float* comp,
                      double* release_total, void* workspace, size_t workspace_bytes, void* stream) {
    GLIDE_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return GLIDE_OK;
    GLIDE_REQUIRE(keys && vals && slab, "NULL keys/vals/slab");
    GLIDE_REQUIRE(n_cells > 0 && n_cells < 0xffffffffu, "bad n_cells");
    GLIDE_REQUIRE(n < (1ll << 32) - 1, "too many keys (positions must fit 32 bits)");
    const size_t need = footprint_layout(n, nullptr, nullptr, nullptr, nullptr);
    GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= need,
                  "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    if (release_total != nullptr) {
        GLIDE_REQUIRE(cells_per_release > 0 && n_cells % cells_per_release == 0,
                      "cells_per_release must divide n_cells");
    }
    SortWork sw{};
    RedWork wc{}, wr{};
    footprint_layout(n, &sw, &wc, &wr, static_cast<char*>(workspace));
    hipStream_t st = static_cast<hipStream_t>(stream);
    // Key bits: enough for any key below n_cells (the others are dropped by the first pass).
    int key_bits = 1;
    while (key_bits < 32 && (static_cast<uint64_t>(n_cells - 1) >> key_bits) != 0) ++key_bits;
    const uint32_t* sk = nullptr;
    const uint32_t* sv = nullptr;
    radix_sort_pairs(keys, reinterpret_cast<const uint32_t*>(vals), n, key_bits, n_cells, sw, st, &sk, &sv);
    const float* svf = reinterpret_cast<const float*>(sv);
    const CellTarget ct{n_cells, slab, comp, t_bin, cells_per_release, n_t};
    const bool rel = release_total != nullptr;
    const uint32_t cpr = rel ? cells_per_release : 1u;
    const ReleaseTarget rt{n_cells, cpr, n_cells / cpr, release_total};
    const unsigned tr = static_cast<unsigned>(div_up(n, kRedTile));
    if (rel) {
        hipLaunchKernelGGL(seg_reduce_tiles_kernel<true>, dim3(tr), dim3(kRedThreads), 0, st, sk, svf, sw.count, ct,
                           rt, wc, wr);
    } else {
        hipLaunchKernelGGL(seg_reduce_tiles_kernel<false>, dim3(tr), dim3(kRedThreads), 0, st, sk, svf, sw.count, ct,
                           rt, wc, wr);
    }
    hipLaunchKernelGGL(seg_fixup_kernel, dim3(rel ? 2u : 1u), dim3(kScanThreads), 0, st, sw.count, ct, rt, wc, wr);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

// ---------------------------------------------------------------------------
// glide_footprint_accumulate_noalloc: the same accumulation with its
// workspace in a file-scope device array, for callers that cannot allocate.
// Each GPU holds its own copy of the array (at its own address) until the
// process ends; every call on a GPU uses that GPU's copy, so calls that
// overlap on one GPU race. Nothing in it must start at zero: every pass writes
// what the next one reads (the same contract as a caller's fresh workspace).
// ---------------------------------------------------------------------------
namespace glide {

constexpr size_t kNoallocScratchBytes = 19u << 20;  // footprint_layout(GLIDE_NOALLOC_MAX_PAIRS) = 18.03 MiB

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kNoallocMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoallocMaxDevices] = {};

// Device address of the current device's scratch array: resolved once per
// device (with that device current) and kept. The array's name is a host-side
// stand-in in host code, never a device address.
static int noalloc_scratch(void** out) {
    int dev = 0;
    GLIDE_HIP(hipGetDevice(&dev));
    GLIDE_REQUIRE(dev >= 0 && dev < kNoallocMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) return check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                              "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GLIDE_OK;
}

}  // namespace glide

extern "C" int glide_footprint_accumulate_noalloc(const uint32_t* keys, const float* vals, int64_t n,
                                                  uint32_t n_cells, uint32_t cells_per_release, float* slab,
                                                  float* comp, double* release_total, void* stream) {
    GLIDE_REQUIRE(n >= 0, "n must be >= 0");
    GLIDE_REQUIRE(n <= GLIDE_NOALLOC_MAX_PAIRS,
                  "glide_footprint_accumulate_noalloc: %lld pairs exceed the static workspace's maximum of %d",
                  static_cast<long long>(n), GLIDE_NOALLOC_MAX_PAIRS);
    const size_t need = footprint_layout(n, nullptr, nullptr, nullptr, nullptr);
    GLIDE_REQUIRE(need <= kNoallocScratchBytes, "static workspace too small: need %zu bytes, have %zu", need,
                  kNoallocScratchBytes);
    if (n == 0) return GLIDE_OK;
    void* ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc != GLIDE_OK) return rc;
    return accumulate(keys, vals, n, n_cells, cells_per_release, nullptr, 1u, slab, comp, release_total, ws,
                      kNoallocScratchBytes, stream);
}
