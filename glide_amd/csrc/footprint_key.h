// Footprint cell key and the escape predicate, bit-exact with the reference's
// float32 index math (footprint_gridder.py:156-196, main.py:690-711).
#pragma once

#include "common.h"

namespace glide {

struct FpDev {
    float lon_min, lat_min, span_x, span_y;  // float32 values, as torch sees them
    float fnx, fny;
    int n_r, n_t, n_z, n_y, n_x;
    uint32_t n_cells;  // flat cells n_r*n_t*n_z*n_y*n_x (the 32-bit key paths)
    uint32_t n_keys;   // key space; the dropped key (n_cells, or n_r*n_z*n_y*n_x when binned)
    uint64_t acc_cells;  // cells of the fixed-point accumulator (the flat cells, or the slabs' total)
    const int64_t* slab_off;  // per-release slab offsets (NULL: r * n_t * cells per bin)
    const int32_t* slab_nt;   // per-release time bins (NULL: n_t)
    const float* z_edges;
    float dom_lon_min, dom_lon_max, dom_lat_min, dom_lat_max, dom_alt_max;
    int32_t* t_bin;  // release-binned keys: per-release time bin of this step (NULL: flat keys)
};

inline FpDev make_fp(const GlideFootprintGrid& g) {
    FpDev d;
    d.lon_min = static_cast<float>(g.lon_min);
    d.lat_min = static_cast<float>(g.lat_min);
    const double sx = g.lon_max - g.lon_min;
    const double sy = g.lat_max - g.lat_min;
    d.span_x = static_cast<float>(sx > 1e-6 ? sx : 1e-6);
    d.span_y = static_cast<float>(sy > 1e-6 ? sy : 1e-6);
    d.fnx = static_cast<float>(g.n_x);
    d.fny = static_cast<float>(g.n_y);
    d.n_r = g.n_r;
    d.n_t = g.n_t;
    d.n_z = g.n_z;
    d.n_y = g.n_y;
    d.n_x = g.n_x;
    const uint64_t flat = static_cast<uint64_t>(g.n_r) * g.n_t * g.n_z * g.n_y * g.n_x;
    d.n_cells = static_cast<uint32_t>(flat < 0xffffffffull ? flat : 0xffffffffull);
    d.slab_off = g.slab_off;
    d.slab_nt = g.slab_nt;
    d.acc_cells = g.slab_off ? static_cast<uint64_t>(g.slab_cells) : flat;
    d.t_bin = g.t_bin;
    d.n_keys = g.t_bin ? static_cast<uint32_t>(static_cast<uint64_t>(g.n_r) * g.n_z * g.n_y * g.n_x) : d.n_cells;
    d.z_edges = g.z_edges;
    d.dom_lon_min = static_cast<float>(g.dom_lon_min);
    d.dom_lon_max = static_cast<float>(g.dom_lon_max);
    d.dom_lat_min = static_cast<float>(g.dom_lat_min);
    d.dom_lat_max = static_cast<float>(g.dom_lat_max);
    d.dom_alt_max = static_cast<float>(g.dom_alt_max);
    return d;
}

// floor(((c - c_min) / span) * n) as an index, or -1 when outside [0, n).
__device__ __forceinline__ int bin_index(float c, float c_min, float span, float fn, int n) {
#pragma clang fp contract(off)
    const float f = floorf(((c - c_min) / span) * fn);
    if (!(f >= 0.0f) || f >= static_cast<float>(n)) return -1;
    return static_cast<int>(f);
}

// Key of one particle: flat cell index, or n_keys when the particle is
// inactive or outside any dimension (the reference's w_eff = 0 case). With
// release-binned keys the time bin leaves the key and goes to t_bin[r]
// (every lane of a release stores the same value).
__device__ __forceinline__ uint32_t footprint_key(const FpDev& g, bool active, float lon, float lat,
                                                  float z, int64_t t_idx, int64_t r_idx) {
    if (!active) return g.n_keys;
    if (r_idx < 0 || r_idx >= g.n_r || t_idx < 0 || t_idx >= g.n_t) return g.n_keys;
    const int xi = bin_index(lon, g.lon_min, g.span_x, g.fnx, g.n_x);
    const int yi = bin_index(lat, g.lat_min, g.span_y, g.fny, g.n_y);
    if (xi < 0 || yi < 0) return g.n_keys;
    // bucketize(z, edges, right=False) - 1: bins (e_b, e_{b+1}].
    int cnt = 0;
    for (int k = 0; k <= g.n_z; ++k) cnt += (g.z_edges[k] < z) ? 1 : 0;
    const int zi = cnt - 1;
    if (zi < 0 || zi >= g.n_z) return g.n_keys;
    if (g.t_bin != nullptr) {
        g.t_bin[r_idx] = static_cast<int32_t>(t_idx);
        return static_cast<uint32_t>(((static_cast<uint64_t>(r_idx) * g.n_z + zi) * g.n_y + yi) * g.n_x + xi);
    }
    const uint64_t key =
        ((((static_cast<uint64_t>(r_idx) * g.n_t + t_idx) * g.n_z + zi) * g.n_y + yi) * g.n_x + xi);
    return static_cast<uint32_t>(key);
}

// Accumulator cell of one particle for the fixed-point footprint: release
// r's slab offset plus its flat in-slab cell ((t * n_z + z) * n_y + y) * n_x +
// x, 64-bit; kNoCell when the particle is outside any dimension or its time
// bin is past its release's slab (slab_nt). The bin arithmetic is
// footprint_key's, bit for bit.
constexpr uint64_t kNoCell = ~0ull;

__device__ __forceinline__ uint64_t fixed_cell(const FpDev& g, float lon, float lat, float z, int64_t t_idx,
                                               int64_t r_idx) {
    if (r_idx < 0 || r_idx >= g.n_r || t_idx < 0 || t_idx >= g.n_t) return kNoCell;
    if (g.slab_nt != nullptr && t_idx >= g.slab_nt[r_idx]) return kNoCell;
    const int xi = bin_index(lon, g.lon_min, g.span_x, g.fnx, g.n_x);
    const int yi = bin_index(lat, g.lat_min, g.span_y, g.fny, g.n_y);
    if (xi < 0 || yi < 0) return kNoCell;
    int cnt = 0;
    for (int k = 0; k <= g.n_z; ++k) cnt += (g.z_edges[k] < z) ? 1 : 0;
    const int zi = cnt - 1;
    if (zi < 0 || zi >= g.n_z) return kNoCell;
    const uint64_t local = ((static_cast<uint64_t>(t_idx) * g.n_z + zi) * g.n_y + yi) * g.n_x + xi;
    const uint64_t base = g.slab_off != nullptr ? static_cast<uint64_t>(g.slab_off[r_idx])
                                                : static_cast<uint64_t>(r_idx) * g.n_t * g.n_z * g.n_y * g.n_x;
    return base + local;
}

// Time-ago bin (main.py:1459-1465) from the pre-step cursor.
__device__ __forceinline__ int64_t time_bin(float we_off, float cursor_off, int n_t) {
#pragma clang fp contract(off)
    const float el = fmaxf(we_off - cursor_off, 0.0f);
    int64_t t = static_cast<int64_t>(floorf(el / 3600.0f));
    if (n_t == 1 && t > 0) t = 0;
    return t;
}

__device__ __forceinline__ bool within_domain(const FpDev& g, float lon, float lat, float z) {
    return lon >= g.dom_lon_min && lon <= g.dom_lon_max && lat >= g.dom_lat_min &&
           lat <= g.dom_lat_max && z <= g.dom_alt_max;
}

// Fixed-point residence of one particle: rint(v * 2^e) (v * 2^e is exact in
// double; the rounding only drops bits below 2^0).
__device__ __forceinline__ long long to_fixed(float v, int e) {
    return static_cast<long long>(llrint(ldexp(static_cast<double>(v), e)));
}

// 64-bit read of another lane's value through DPP (CTRL: the DPP control
// word), 0 where the source lane is outside the row or the row is masked off
// (ROWMASK). The wavefront must be fully active.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    const uint32_t lo = __builtin_amdgcn_update_dpp(0u, static_cast<uint32_t>(x), CTRL, ROWMASK, 0xf, false);
    const uint32_t hi = __builtin_amdgcn_update_dpp(0u, static_cast<uint32_t>(x >> 32), CTRL, ROWMASK, 0xf, false);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Inclusive segmented scan over the wavefront: lane l gets the sum of lanes
// [start, l] (start <= l, the first lane of its segment). Hillis-Steele steps
// within rows of 16 lanes (row_shr 1, 2, 4, 8), then the last lane of row 0
// into row 1 and of row 2 into row 3 (row_bcast:15), then lane 31 into rows 2
// and 3 (row_bcast:31); a step adds only when its source lane is in the same
// segment. Exact in integers.
__device__ __forceinline__ long long seg_scan_i64(long long v, int lane, int start) {
    unsigned long long u = static_cast<unsigned long long>(v), o;
    o = dpp_u64<kDppRowShr1, 0xf>(u);
    if (lane - 1 >= start) u += o;
    o = dpp_u64<kDppRowShr2, 0xf>(u);
    if (lane - 2 >= start) u += o;
    o = dpp_u64<kDppRowShr4, 0xf>(u);
    if (lane - 4 >= start) u += o;
    o = dpp_u64<kDppRowShr8, 0xf>(u);
    if (lane - 8 >= start) u += o;
    o = dpp_u64<kDppRowBcast15, 0xa>(u);
    if ((lane & ~15) - 1 >= start) u += o;
    o = dpp_u64<kDppRowBcast31, 0xc>(u);
    if (31 >= start) u += o;
    return static_cast<long long>(u);
}

// Sum of v over the wavefront (the same DPP steps without segments, read
// from lane 63).
__device__ __forceinline__ long long wave_sum_i64_dpp(long long v) {
    unsigned long long u = static_cast<unsigned long long>(v);
    u += dpp_u64<kDppRowShr1, 0xf>(u);
    u += dpp_u64<kDppRowShr2, 0xf>(u);
    u += dpp_u64<kDppRowShr4, 0xf>(u);
    u += dpp_u64<kDppRowShr8, 0xf>(u);
    u += dpp_u64<kDppRowBcast15, 0xa>(u);
    u += dpp_u64<kDppRowBcast31, 0xc>(u);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), 63);
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Adds each lane's fixed-point value to acc[key] (kNoCell dropped), exactly:
// integer addition is associative, so the sums do not depend on how the
// lanes are grouped or the order the atomics land in.
//
// A wavefront's particles were binned by met column at the last re-binning
// and have drifted since, so its keys are few but interleaved. First, up to
// kMatchRounds times: the lowest pending lane's key, the wavefront sum of the
// lanes holding that key, one 64-bit atomic from that lane. What is left
// (a wavefront over many cells) goes as runs of equal keys in lane order: an
// inclusive segmented scan, one atomic per run from its last lane.
// Every lane of the wavefront must call it (no lane may have returned).
constexpr int kMatchRounds = 8;
__device__ __forceinline__ void wave_fixed_add(uint64_t key, long long v, long long* acc) {
    bool pend = key != kNoCell && v != 0;
    unsigned long long m = __ballot(pend);
    if (m == 0ull) return;
    const int lane = static_cast<int>(__lane_id());
    for (int round = 0; round < kMatchRounds && m != 0ull; ++round) {
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        const uint32_t klo = __builtin_amdgcn_readlane(static_cast<uint32_t>(key), leader);
        const uint32_t khi = __builtin_amdgcn_readlane(static_cast<uint32_t>(key >> 32), leader);
        const uint64_t kk = (static_cast<uint64_t>(khi) << 32) | klo;
        const bool mine = pend && key == kk;
        const long long t = wave_sum_i64_dpp(mine ? v : 0);
        if (lane == leader)
            atomicAdd(reinterpret_cast<unsigned long long*>(acc + kk), static_cast<unsigned long long>(t));
        pend = pend && !mine;
        m = __ballot(pend);
    }
    if (m == 0ull) return;
    const uint64_t k = pend ? key : kNoCell;
    const uint64_t prev = __shfl_up(static_cast<unsigned long long>(k), 1u);
    const unsigned long long heads = __ballot(lane == 0 || prev != k);
    const int start = 63 - __clzll(static_cast<long long>(heads & ((2ull << lane) - 1ull)));
    const long long s = seg_scan_i64(pend ? v : 0, lane, start);
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull) != 0ull;
    if (tail && k != kNoCell && s != 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(acc + k), static_cast<unsigned long long>(s));
}

// This workgroup's copy of the fixed-point accumulator (GlideEmit.fixed_copies).
__device__ __forceinline__ long long* fixed_copy(const GlideEmit& e, uint64_t acc_cells) {
    const uint32_t copies = e.fixed_copies > 1 ? static_cast<uint32_t>(e.fixed_copies) : 1u;
    return e.fixed + static_cast<uint64_t>(blockIdx.x % copies) * acc_cells;
}

// Adds the number of lanes with flag set to *ctr: wave ballot, one integer atomic per wave.
__device__ __forceinline__ void count_flags(bool esc, unsigned long long* ctr) {
    const unsigned long long m = __ballot(esc);
    if (m == 0ull) return;
    const unsigned long long act = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(act)) - 1;
    if (static_cast<int>(__lane_id()) == leader) atomicAdd(ctr, static_cast<unsigned long long>(__popcll(m)));
}

}  // namespace glide
