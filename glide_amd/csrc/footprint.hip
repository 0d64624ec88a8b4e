// Deterministic footprint accumulation (replaces scatter_add_ float atomics,
// reference footprint_gridder.py:197).
//
//  1. Stable LSD radix sort of (key, value) pairs by key, 8-bit digits over the
//     bits the key range needs. Per pass: per-tile digit histograms, a scan of
//     each digit's row of tile counts (one workgroup per digit), and a stable
//     per-tile scatter (wavefront ballot ranking; waves, rounds and tiles in
//     order) that adds the digit bases it scans itself.
//  2. Segmented reduction of the sorted values: sequential per thread, a
//     fixed-shape segmented scan per tile, and one workgroup that scans the
//     tiles for segments crossing tile boundaries. Every cell's sum is formed in
//     one fixed order, so the result does not depend on scheduling.
//  3. One read-modify-write per touched cell: slab[key] += sum, optionally with
//     a Neumaier compensation term kept in comp[key].
//  4. Optionally the same reduction keyed by release (key / cells_per_release)
//     accumulates each release's residence total in float64.
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
    uint32_t* hist;       // [kRadix][n_sort_tiles], scanned per row in place
    uint32_t* row_total;  // [kRadix]
    double* head_part;    // [n_red_tiles] sum before the tile's first head (whole tile if none)
    double* tail_part;    // [n_red_tiles] open segment at the tile end (from its last head)
    uint32_t* first_key;  // [n_red_tiles] segment key of the tile's first item
    uint32_t* flags;      // bit0 any head, bit1 open segment continues, bit2 first item is a head
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
    char* rt = take(sizeof(uint32_t) * kRadix);
    char* hp = take(sizeof(double) * tr);
    char* tp = take(sizeof(double) * tr);
    char* fk = take(sizeof(uint32_t) * tr);
    char* fl = take(sizeof(uint32_t) * tr);
    if (w) {
        w->keys_a = reinterpret_cast<uint32_t*>(ka);
        w->vals_a = reinterpret_cast<float*>(va);
        w->keys_b = reinterpret_cast<uint32_t*>(kb);
        w->vals_b = reinterpret_cast<float*>(vb);
        w->hist = reinterpret_cast<uint32_t*>(hi);
        w->row_total = reinterpret_cast<uint32_t*>(rt);
        w->head_part = reinterpret_cast<double*>(hp);
        w->tail_part = reinterpret_cast<double*>(tp);
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

// Exclusive scan of one digit's row of tile counts (in place) and its total.
__global__ __launch_bounds__(256) void radix_row_scan_kernel(uint32_t* __restrict__ hist, int64_t n_tiles,
                                                             uint32_t* __restrict__ row_total) {
    __shared__ uint32_t part[256];
    uint32_t* row = hist + static_cast<int64_t>(blockIdx.x) * n_tiles;
    const int t = threadIdx.x;
    // contiguous chunk per thread, coalesced in strided passes
    const int64_t per = (n_tiles + 255) / 256;
    const int64_t lo = t * per;
    const int64_t hi = lo + per < n_tiles ? lo + per : n_tiles;
    uint32_t s = 0;
    for (int64_t j = lo; j < hi; ++j) s += row[j];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t v = row[j];
        row[j] = run;
        run += v;
    }
    if (t == 255) row_total[blockIdx.x] = part[255];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const float* __restrict__ vin, uint32_t* __restrict__ kout,
    float* __restrict__ vout, int64_t n, int shift, int bits, const uint32_t* __restrict__ offs,
    const uint32_t* __restrict__ row_total, int64_t n_tiles) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const uint32_t mask = (1u << bits) - 1u;
    // digit base = sum of the totals of all smaller digits (block scan)
    base[t] = row_total[t];
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t v = t >= off ? base[t - off] : 0u;
        __syncthreads();
        base[t] += v;
        __syncthreads();
    }
    const uint32_t digit_base = base[t] - row_total[t];
    __syncthreads();
    base[t] = digit_base + offs[static_cast<int64_t>(t) * n_tiles + blockIdx.x];
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
// Segmented reduction, by cell (float sums into the slab) or by release
// (float64 sums into the release totals).
// --------------------------------------------------------------------------
struct CellTarget {
    using Acc = float;
    uint32_t n_cells;
    float* slab;
    float* comp;
    __device__ __forceinline__ uint32_t seg(uint32_t key) const { return key; }
    __device__ __forceinline__ bool valid(uint32_t s) const { return s < n_cells; }
    __device__ __forceinline__ void emit(uint32_t key, float x) const {
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
};

struct ReleaseTarget {
    using Acc = double;
    uint32_t n_cells, cpr, n_rel;
    double* total;
    __device__ __forceinline__ uint32_t seg(uint32_t key) const { return key < n_cells ? key / cpr : n_rel; }
    __device__ __forceinline__ bool valid(uint32_t s) const { return s < n_rel; }
    __device__ __forceinline__ void emit(uint32_t r, double x) const { total[r] += x; }
};

template <typename T>
__global__ __launch_bounds__(kRedThreads) void seg_reduce_tiles_kernel(const uint32_t* __restrict__ keys,
                                                                       const float* __restrict__ vals, int64_t n,
                                                                       T tg, Workspace w) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    __shared__ uint32_t sk[kRedTile];
    __shared__ float sv[kRedTile];
    __shared__ Acc scan_v[kRedThreads];
    __shared__ uint32_t scan_f[kRedThreads];
    __shared__ uint32_t edge_key[2];  // segment key before the tile, after the tile
    const int t = threadIdx.x;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRedTile;
    const int cnt = static_cast<int>((n - tile0) < kRedTile ? (n - tile0) : kRedTile);
    for (int q = 0; q < kRedItems; ++q) {
        const int idx = q * kRedThreads + t;
        if (idx < cnt) {
            sk[idx] = tg.seg(keys[tile0 + idx]);
            sv[idx] = vals[tile0 + idx];
        }
    }
    if (t == 0) {
        edge_key[0] = tile0 > 0 ? tg.seg(keys[tile0 - 1]) : 0xffffffffu;
        edge_key[1] = (tile0 + cnt < n) ? tg.seg(keys[tile0 + cnt]) : 0xffffffffu;
    }
    __syncthreads();
    const bool first_tile = tile0 == 0;
    const bool last_tile = tile0 + cnt >= n;
    const bool item0_head = first_tile || sk[0] != edge_key[0];

    // Sequential pass over this thread's consecutive items.
    Acc acc = 0;
    Acc first_part = 0;
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
            const uint32_t k = sk[i0 - 1];
            if (tg.valid(k)) tg.emit(k, tot);
        } else if (!item0_head) {
            w.head_part[blockIdx.x] = static_cast<double>(tot);  // continuation of an earlier tile's segment
        }
    }
    if (t == kRedThreads - 1) {
        const uint32_t lk = sk[cnt - 1];
        const bool next_head = last_tile || edge_key[1] != lk;
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
        w.first_key[blockIdx.x] = sk[0];
        w.flags[blockIdx.x] = fl;
    }
}

// Segments crossing tiles. With P_j the open-segment sum at the end of tile j
// (P_j = tail_j if tile j has a head, else P_{j-1} + head_j), the segment that
// closes in tile c has total P_{c-1} + head_c. One workgroup scans the tiles in
// fixed chunks, so the order of additions is fixed. Partial sums are held in
// the target's accumulator type (float sums are exact in the double slots).
template <typename T>
__global__ __launch_bounds__(kScanThreads) void seg_fixup_kernel(int64_t n_tiles, T tg, Workspace w) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    __shared__ Acc agg_v[kScanThreads];
    __shared__ uint32_t agg_f[kScanThreads];
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

template <typename T>
static void segmented_reduce(const uint32_t* keys, const float* vals, int64_t n, const T& tg, const Workspace& w,
                             hipStream_t st) {
    const int64_t tr = div_up(n, kRedTile);
    hipLaunchKernelGGL(seg_reduce_tiles_kernel<T>, dim3(static_cast<unsigned>(tr)), dim3(kRedThreads), 0, st, keys,
                       vals, n, tg, w);
    hipLaunchKernelGGL(seg_fixup_kernel<T>, dim3(1), dim3(kScanThreads), 0, st, tr, tg, w);
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
        hipLaunchKernelGGL(radix_row_scan_kernel, dim3(kRadix), dim3(256), 0, st, w.hist, ts, w.row_total);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin,
                           vin, kout, vout, n, shift, bits, w.hist, w.row_total, ts);
        kin = kout;
        vin = vout;
    }
    segmented_reduce(kin, vin, n, CellTarget{n_cells, slab, comp}, w, st);
    if (release_total != nullptr) {
        segmented_reduce(kin, vin, n, ReleaseTarget{n_cells, cells_per_release, n_cells / cells_per_release,
                                                    release_total}, w, st);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
