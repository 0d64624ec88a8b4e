// Stable LSD radix sort of 4-byte (key, value) pairs, shared by the footprint
// accumulation and the particle reordering.
#pragma once

#include "common.h"

namespace glide {

constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 4;                         // per thread
constexpr int kSortChunk = 64 * kSortItems;           // items one wavefront ranks (256)
constexpr int kSortTile = kSortThreads * kSortItems;  // items per workgroup (4096)
constexpr int kMinDigitBits = 8;
constexpr int kMaxDigitBits = 11;
constexpr int kMaxRadix = 1 << kMaxDigitBits;

__host__ __device__ inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct SortWork {
    uint32_t* count;      // [1] items being sorted (device; written by the first pass)
    uint32_t* keys_a;
    uint32_t* vals_a;
    uint32_t* keys_b;
    uint32_t* vals_b;
    uint32_t* hist;       // [kMaxRadix][n_sort_tiles], scanned per row in place
    uint32_t* row_total;  // [kMaxRadix]
};

// Bytes of sort workspace for n pairs; fills *w when base is not NULL.
size_t sort_layout(int64_t n, SortWork* w, char* base);

// Sorts by the low key_bits of the key. compact_limit > 0 keeps only the pairs
// with key < compact_limit (order-preserving; the first pass does it). The
// item count is left in w.count. The last pass writes into (dst_keys,
// dst_vals) when both are given, else into the workspace; *out_keys /
// *out_vals point at the result.
void radix_sort_pairs(const uint32_t* keys, const uint32_t* vals, int64_t n, int key_bits, uint32_t compact_limit,
                      const SortWork& w, hipStream_t st, const uint32_t** out_keys, const uint32_t** out_vals,
                      uint32_t* dst_keys = nullptr, uint32_t* dst_vals = nullptr);

}  // namespace glide
