// Stable LSD radix sort of (key, value) 4-byte pairs.
//
// Digits are 8 to 11 bits wide, balanced over the passes the key width needs
// (20-bit keys: two 10-bit passes). Each pass is three launches:
//   hist      per-tile digit counts (one LDS atomic per distinct digit of a
//             wavefront round; integer, so order-free),
//   row scan  one workgroup per digit scans that digit's row of tile counts,
//   scatter   each wavefront ranks its own 256 consecutive items against
//             wave-private LDS digit counters (ballot matching, no block
//             barriers), one barrier turns the counters into per-wave digit
//             offsets, and the items are written. Order is (wave, round, lane)
//             = input order, so the sort is stable.
// The first pass optionally keeps only keys < limit (order-preserving
// compaction) and writes the number of items kept to device memory; later
// passes read it and their tiles past it exit, so nothing waits on the host.
#include "sort.h"

namespace glide {

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

size_t sort_layout(int64_t n, SortWork* w, char* base) {
    const int64_t ts = div_up(n, kSortTile);
    size_t off = 0;
    auto take = [&](size_t count) {
        uint32_t* p = base ? reinterpret_cast<uint32_t*>(base + off) : nullptr;
        off += align256(sizeof(uint32_t) * (count > 0 ? count : 1));
        return p;
    };
    SortWork s{};
    s.count = take(1);
    s.keys_a = take(n);
    s.vals_a = take(n);
    s.keys_b = take(n);
    s.vals_b = take(n);
    s.hist = take(static_cast<size_t>(kMaxRadix) * ts);
    s.row_total = take(kMaxRadix);
    if (w) *w = s;
    return off;
}

__device__ __forceinline__ bool keep(uint32_t key, uint32_t limit) { return limit == 0u || key < limit; }

// Lanes of the wavefront whose digit equals this lane's (the ones with ok set).
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, bool ok, int bits) {
    unsigned long long peers = __ballot(ok);
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

template <int RB>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t n_host,
                                                                  const uint32_t* __restrict__ n_dev, uint32_t limit,
                                                                  int shift, int bits, uint32_t* __restrict__ hist,
                                                                  int64_t n_tiles) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t cnt[R];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t mask = (1u << bits) - 1u;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t n = n_dev ? static_cast<int64_t>(*n_dev) : n_host;
    for (int d = t; d < R; d += kSortThreads) cnt[d] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    if (base < n) {
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t j = base + r * kSortThreads + t;
            uint32_t k = 0;
            bool ok = false;
            if (j < n) {
                k = keys[j];
                ok = keep(k, limit);
            }
            const uint32_t d = (k >> shift) & mask;
            const unsigned long long peers = match_digit(d, ok, bits);
            if (ok && (peers & lt) == 0ull) atomicAdd(&cnt[d], static_cast<uint32_t>(__popcll(peers)));
        }
    }
    __syncthreads();
    for (int d = t; d < R; d += kSortThreads) hist[static_cast<int64_t>(d) * n_tiles + blockIdx.x] = cnt[d];
}

// Inclusive scan of one value per thread over a block of NW wavefronts.
template <int NW>
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* wave_tot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    if (lane == 63) wave_tot[wave] = v;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_tot[w];
    return v + before;
}

// Exclusive scan of one digit's row of tile counts (in place) and its total.
__global__ __launch_bounds__(256) void radix_row_scan_kernel(uint32_t* __restrict__ hist, int64_t n_tiles,
                                                             uint32_t* __restrict__ row_total) {
    __shared__ uint32_t wave_tot[4];
    uint32_t* row = hist + static_cast<int64_t>(blockIdx.x) * n_tiles;
    const int t = threadIdx.x;
    const int64_t per = (n_tiles + 255) / 256;
    const int64_t lo = t * per;
    const int64_t hi = lo + per < n_tiles ? lo + per : n_tiles;
    uint32_t s = 0;
    for (int64_t j = lo; j < hi; ++j) s += row[j];
    const uint32_t incl = block_inclusive_scan<4>(s, wave_tot);
    uint32_t run = incl - s;
    for (int64_t j = lo; j < hi; ++j) {
        const uint32_t v = row[j];
        row[j] = run;
        run += v;
    }
    if (t == 255) row_total[blockIdx.x] = incl;
}

template <int RB>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n_host, const uint32_t* __restrict__ n_dev, uint32_t limit, int shift,
    int bits, const uint32_t* __restrict__ offs, const uint32_t* __restrict__ row_total, int64_t n_tiles,
    uint32_t* __restrict__ count_out) {
    constexpr int R = 1 << RB;
    constexpr int DPT = (R + kSortThreads - 1) / kSortThreads;  // digits each thread scans
    __shared__ uint32_t wcnt[kSortWaves][R];
    __shared__ uint32_t dbase[R];
    __shared__ uint32_t wave_tot[kSortWaves];
    const int64_t n = n_dev ? static_cast<int64_t>(*n_dev) : n_host;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
    if (tile0 >= n) return;  // uniform across the block
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint32_t mask = (1u << bits) - 1u;

    // Digit bases: exclusive scan of the digit totals (thread t owns digits t*DPT .. t*DPT+DPT-1).
    uint32_t tot[DPT];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < DPT; ++i) {
        const int d = t * DPT + i;
        tot[i] = d < R ? row_total[d] : 0u;
        s += tot[i];
    }
    for (int w = 0; w < kSortWaves; ++w)
        for (int d = t; d < R; d += kSortThreads) wcnt[w][d] = 0;
    uint32_t run = block_inclusive_scan<kSortWaves>(s, wave_tot) - s;
#pragma unroll
    for (int i = 0; i < DPT; ++i) {
        const int d = t * DPT + i;
        if (d < R) dbase[d] = run;
        run += tot[i];
    }
    if (count_out != nullptr && blockIdx.x == 0 && t == kSortThreads - 1) *count_out = run;
    __syncthreads();

    // Phase 1: each wavefront ranks its chunk against its own digit counters.
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t chunk0 = tile0 + static_cast<int64_t>(wave) * kSortChunk;
    uint32_t key[kSortItems], val[kSortItems], loc[kSortItems];
    uint32_t okmask = 0;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t j = chunk0 + r * 64 + lane;
        uint32_t k = 0, v = 0;
        bool ok = false;
        if (j < n) {
            k = kin[j];
            v = vin[j];
            ok = keep(k, limit);
        }
        const uint32_t d = (k >> shift) & mask;
        const unsigned long long peers = match_digit(d, ok, bits);
        const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
        uint32_t prior = 0;
        if (ok) prior = wcnt[wave][d];
        // All lanes of the wave have read the counter before its leader updates it
        // (LDS operations of one wavefront execute in order).
        if (ok && rank == 0) wcnt[wave][d] = prior + static_cast<uint32_t>(__popcll(peers));
        key[r] = k;
        val[r] = v;
        loc[r] = prior + rank;
        okmask |= (ok ? 1u : 0u) << r;
    }
    __syncthreads();

    // Phase 2: per digit, the offsets of the waves' runs.
    for (int d = t; d < R; d += kSortThreads) {
        uint32_t pos = dbase[d] + offs[static_cast<int64_t>(d) * n_tiles + blockIdx.x];
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = pos;
            pos += c;
        }
    }
    __syncthreads();

    // Phase 3: write.
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if ((okmask >> r) & 1u) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t pos = wcnt[wave][d] + loc[r];
            kout[pos] = key[r];
            vout[pos] = val[r];
        }
    }
}

template <int RB>
static void launch_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, int64_t n_host,
                        const uint32_t* n_dev, uint32_t limit, int shift, int bits, const SortWork& w, int64_t ts,
                        uint32_t* count_out, hipStream_t st) {
    hipLaunchKernelGGL(radix_hist_kernel<RB>, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin, n_host,
                       n_dev, limit, shift, bits, w.hist, ts);
    hipLaunchKernelGGL(radix_row_scan_kernel, dim3(1u << RB), dim3(256), 0, st, w.hist, ts, w.row_total);
    hipLaunchKernelGGL(radix_scatter_kernel<RB>, dim3(static_cast<unsigned>(ts)), dim3(kSortThreads), 0, st, kin, vin,
                       kout, vout, n_host, n_dev, limit, shift, bits, w.hist, w.row_total, ts, count_out);
}

void radix_sort_pairs(const uint32_t* keys, const uint32_t* vals, int64_t n, int key_bits, uint32_t compact_limit,
                      const SortWork& w, hipStream_t st, const uint32_t** out_keys, const uint32_t** out_vals,
                      uint32_t* dst_keys, uint32_t* dst_vals) {
    const int64_t ts = div_up(n, kSortTile);
    if (key_bits < 1) key_bits = 1;
    const int npass = static_cast<int>(div_up(key_bits, kMaxDigitBits));
    const int per = static_cast<int>(div_up(key_bits, npass));
    const uint32_t* kin = keys;
    const uint32_t* vin = vals;
    int slot = 0;
    int shift = 0;
    for (int ps = 0; ps < npass; ++ps) {
        const int bits = (key_bits - shift) < per ? (key_bits - shift) : per;
        const int rb = bits < kMinDigitBits ? kMinDigitBits : bits;
        uint32_t* kout = slot ? w.keys_b : w.keys_a;
        uint32_t* vout = slot ? w.vals_b : w.vals_a;
        if (ps == npass - 1 && dst_keys != nullptr && dst_vals != nullptr) {
            kout = dst_keys;
            vout = dst_vals;
        }
        const bool first = ps == 0;
        const uint32_t* n_dev = first ? nullptr : w.count;
        const uint32_t limit = first ? compact_limit : 0u;
        uint32_t* count_out = first ? w.count : nullptr;
        switch (rb) {
            case 8: launch_pass<8>(kin, vin, kout, vout, n, n_dev, limit, shift, bits, w, ts, count_out, st); break;
            case 9: launch_pass<9>(kin, vin, kout, vout, n, n_dev, limit, shift, bits, w, ts, count_out, st); break;
            case 10: launch_pass<10>(kin, vin, kout, vout, n, n_dev, limit, shift, bits, w, ts, count_out, st); break;
            default: launch_pass<11>(kin, vin, kout, vout, n, n_dev, limit, shift, bits, w, ts, count_out, st); break;
        }
        kin = kout;
        vin = vout;
        slot ^= 1;
        shift += bits;
    }
    *out_keys = kin;
    *out_vals = vin;
}

}  // namespace glide

using namespace glide;

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
    GLIDE_REQUIRE(out_keys != keys && out_vals != vals && out_keys != vals && out_vals != keys,
                  "glide_sort_pairs cannot sort in place");
    const uint32_t* sk = nullptr;
    const uint32_t* sv = nullptr;
    radix_sort_pairs(keys, vals, n, key_bits, 0u, sw, st, &sk, &sv, out_keys, out_vals);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
