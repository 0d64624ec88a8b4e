// Deterministic footprint accumulation (replaces scatter_add_ float atomics,
// reference footprint_gridder.py:197), and the radix sort it is built on.
//
//  0. Order-preserving compaction of the valid (key < n_cells) pairs, with the
//     count kept in device memory: dropped particles never reach the sort.
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
//
// Item counts after compaction live in device memory; kernels are launched for
// the upper bound and tiles past the count exit (nothing waits on the host).
#include "sort.h"

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

size_t sort_layout(int64_t n, SortWork* w, char* base) {
    const int64_t ts = div_up(n, kSortTile);
    Bump b{base};
    SortWork s{};
    s.count = b.take<uint32_t>(1);
    s.tile_off = b.take<uint32_t>(ts);
    s.keys_a = b.take<uint32_t>(n);
    s.vals_a = b.take<uint32_t>(n);
    s.keys_b = b.take<uint32_t>(n);
    s.vals_b = b.take<uint32_t>(n);
    s.hist = b.take<uint32_t>(static_cast<size_t>(kRadix) * ts);
    s.row_total = b.take<uint32_t>(kRadix);
    if (w) *w = s;
    return b.off;
}

// --------------------------------------------------------------------------
// Compaction of valid pairs (order-preserving)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void compact_count_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                                     uint32_t limit, uint32_t* __restrict__ tile_cnt) {
    __shared__ uint32_t c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    uint32_t mine = 0;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = base + r * kSortThreads + threadIdx.x;
        if (j < n && keys[j] < limit) ++mine;
    }
    atomicAdd(&c, mine);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = c;
}

// Exclusive scan of m tile counts in place; writes the total to *count.
__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(uint32_t* __restrict__ a, int64_t m,
                                                                  uint32_t* __restrict__ count) {
    __shared__ uint32_t part[kScanThreads];
    const int t = threadIdx.x;
    const int64_t per = (m + kScanThreads - 1) / kScanThreads;
    const int64_t lo = t * per;
    const int64_t hi = lo + per < m ? lo + per : m;
    uint32_t s = 0;
    for (int64_t j = lo; j < hi; ++j) s += a[j];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t v = a[j];
        a[j] = run;
        run += v;
    }
    if (t == kScanThreads - 1) *count = part[t];
}

__global__ __launch_bounds__(kSortThreads) void compact_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, uint32_t limit,
    const uint32_t* __restrict__ tile_off, uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t run;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (t == 0) run = tile_off[blockIdx.x];
    __syncthreads();
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = tile0 + r * kSortThreads + t;
        uint32_t key = 0, val = 0;
        bool v = false;
        if (j < n) {
            key = kin[j];
            val = vin[j];
            v = key < limit;
        }
        const unsigned long long m = __ballot(v);
        if (lane == 0) wsum[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t before = run;
        for (int q = 0; q < wave; ++q) before += wsum[q];
        if (v) {
            const uint32_t pos = before + static_cast<uint32_t>(__popcll(m & lt));
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        if (t == 0) {
            uint32_t s = 0;
            for (int q = 0; q < kWaves; ++q) s += wsum[q];
            run += s;
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Radix sort kernels (counts from device memory)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys,
                                                                  const uint32_t* __restrict__ n_dev, int shift,
                                                                  uint32_t mask, uint32_t* __restrict__ hist,
                                                                  int64_t n_tiles) {
    __shared__ uint32_t cnt[kRadix];
    const int64_t n = *n_dev;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    if (base < n) {
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t j = base + r * kSortThreads + threadIdx.x;
            if (j < n) atomicAdd(&cnt[(keys[j] >> shift) & mask], 1u);
        }
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
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ n_dev, int shift, int bits,
    const uint32_t* __restrict__ offs, const uint32_t* __restrict__ row_total, int64_t n_tiles) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    const int64_t n = *n_dev;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
    if (tile0 >= n) return;  // uniform across the block
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
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = tile0 + r * kSortThreads + t;
        const bool valid = j < n;
        uint32_t key = 0, val = 0, d = 0;
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

__global__ void set_count_kernel(uint32_t* count, uint32_t n) { *count = n; }

// Sorts (keys, vals) 4-byte pairs by the low key_bits of the key. With
// compact_limit > 0 only pairs with key < compact_limit are kept (in order).
// Returns pointers to the sorted arrays (in the workspace) and leaves the item
// count in w.count.
void radix_sort_pairs(const uint32_t* keys, const uint32_t* vals, int64_t n, int key_bits, uint32_t compact_limit,
                      const SortWork& w, hipStream_t st, const uint32_t** out_keys, const uint32_t** out_vals) {
    const int64_t ts = div_up(n, kSortTile);
    const uint32_t* kin = keys;
    const uint32_t* vin = vals;
    int slot = 0;
    if (compact_limit > 0) {
        hipLaunchKernelGGL(compact_count_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, keys, n,
                           compact_limit, w.tile_off);
        hipLaunchKernelGGL(scan_tiles_kernel, dim3(1), dim3(kScanThreads), 0, st, w.tile_off, ts, w.count);
        hipLaunchKernelGGL(compact_scatter_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, keys,
                           vals, n, compact_limit, w.tile_off, w.keys_a, w.vals_a);
        kin = w.keys_a;
        vin = w.vals_a;
        slot = 1;
    } else {
        hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(1), 0, st, w.count, static_cast<uint32_t>(n));
    }
    const int npass = (key_bits + 7) / 8;
    for (int ps = 0; ps < npass; ++ps) {
        const int shift = ps * 8;
        const int bits = (key_bits - shift) < 8 ? (key_bits - shift) : 8;
        const uint32_t mask = (1u << bits) - 1u;
        uint32_t* kout = slot ? w.keys_b : w.keys_a;
        uint32_t* vout = slot ? w.vals_b : w.vals_a;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin, w.count,
                           shift, mask, w.hist, ts);
        hipLaunchKernelGGL(radix_row_scan_kernel, dim3(kRadix), dim3(256), 0, st, w.hist, ts, w.row_total);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin, vin,
                           kout, vout, w.count, shift, bits, w.hist, w.row_total, ts);
        kin = kout;
        vin = vout;
        slot ^= 1;
    }
    *out_keys = kin;
    *out_vals = vin;
}

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
                                                                       const float* __restrict__ vals,
                                                                       const uint32_t* __restrict__ n_dev, T tg,
                                                                       RedWork w) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    __shared__ uint32_t sk[kRedTile];
    __shared__ float sv[kRedTile];
    __shared__ Acc scan_v[kRedThreads];
    __shared__ uint32_t scan_f[kRedThreads];
    __shared__ uint32_t edge_key[2];  // segment key before the tile, after the tile
    const int64_t n = *n_dev;
    const int t = threadIdx.x;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRedTile;
    if (tile0 >= n) return;  // uniform across the block
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
__global__ __launch_bounds__(kScanThreads) void seg_fixup_kernel(const uint32_t* __restrict__ n_dev, T tg,
                                                                 RedWork w) {
#pragma clang fp contract(off)
    using Acc = typename T::Acc;
    __shared__ Acc agg_v[kScanThreads];
    __shared__ uint32_t agg_f[kScanThreads];
    const int64_t n_tiles = div_up(static_cast<int64_t>(*n_dev), kRedTile);
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
static void segmented_reduce(const uint32_t* keys, const float* vals, int64_t n_max, const uint32_t* n_dev,
                             const T& tg, const RedWork& w, hipStream_t st) {
    const int64_t tr = div_up(n_max, kRedTile);
    hipLaunchKernelGGL(seg_reduce_tiles_kernel<T>, dim3(static_cast<unsigned>(tr)), dim3(kRedThreads), 0, st, keys,
                       vals, n_dev, tg, w);
    hipLaunchKernelGGL(seg_fixup_kernel<T>, dim3(1), dim3(kScanThreads), 0, st, n_dev, tg, w);
}

static size_t footprint_layout(int64_t n, SortWork* sw, RedWork* rw, char* base) {
    const size_t s = sort_layout(n, sw, base);
    const int64_t tr = div_up(n, kRedTile);
    Bump b{base ? base + s : nullptr};
    RedWork r{};
    r.head_part = b.take<double>(tr);
    r.tail_part = b.take<double>(tr);
    r.first_key = b.take<uint32_t>(tr);
    r.flags = b.take<uint32_t>(tr);
    if (rw) *rw = r;
    return s + b.off;
}

}  // namespace glide

using namespace glide;

extern "C" size_t glide_footprint_workspace_bytes(int64_t n) {
    if (n < 0) n = 0;
    return footprint_layout(n, nullptr, nullptr, nullptr);
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
    const size_t need = footprint_layout(n, nullptr, nullptr, nullptr);
    GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= need,
                  "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    if (release_total != nullptr) {
        GLIDE_REQUIRE(cells_per_release > 0 && n_cells % cells_per_release == 0,
                      "cells_per_release must divide n_cells");
    }
    SortWork sw{};
    RedWork rw{};
    footprint_layout(n, &sw, &rw, static_cast<char*>(workspace));
    hipStream_t st = static_cast<hipStream_t>(stream);
    // Key bits: enough for any key below n_cells (dropped pairs are compacted out).
    int key_bits = 1;
    while (key_bits < 32 && (static_cast<uint64_t>(n_cells - 1) >> key_bits) != 0) ++key_bits;
    const uint32_t* sk = nullptr;
    const uint32_t* sv = nullptr;
    radix_sort_pairs(keys, reinterpret_cast<const uint32_t*>(vals), n, key_bits, n_cells, sw, st, &sk, &sv);
    const float* svf = reinterpret_cast<const float*>(sv);
    segmented_reduce(sk, svf, n, sw.count, CellTarget{n_cells, slab, comp}, rw, st);
    if (release_total != nullptr) {
        segmented_reduce(sk, svf, n, sw.count,
                         ReleaseTarget{n_cells, cells_per_release, n_cells / cells_per_release, release_total}, rw,
                         st);
    }
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" size_t glide_sort_workspace_bytes(int64_t n) {
    if (n < 0) n = 0;
    return sort_layout(n, nullptr, nullptr);
}

extern "C" int glide_sort_pairs(const uint32_t* keys, const uint32_t* vals, int64_t n, int32_t key_bits,
                                uint32_t* out_keys, uint32_t* out_vals, void* workspace, size_t workspace_bytes,
                                void* stream) {
    GLIDE_REQUIRE(n >= 0 && key_bits >= 1 && key_bits <= 32, "bad n / key_bits");
    if (n == 0) return GLIDE_OK;
    GLIDE_REQUIRE(keys && vals && out_keys && out_vals, "NULL array");
    GLIDE_REQUIRE(n < (1ll << 32) - 1, "too many keys (positions must fit 32 bits)");
    const size_t need = sort_layout(n, nullptr, nullptr);
    GLIDE_REQUIRE(workspace != nullptr && workspace_bytes >= need,
                  "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    SortWork sw{};
    sort_layout(n, &sw, static_cast<char*>(workspace));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t* sk = nullptr;
    const uint32_t* sv = nullptr;
    radix_sort_pairs(keys, vals, n, key_bits, 0u, sw, st, &sk, &sv);
    GLIDE_HIP(hipMemcpyAsync(out_keys, sk, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    GLIDE_HIP(hipMemcpyAsync(out_vals, sv, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
