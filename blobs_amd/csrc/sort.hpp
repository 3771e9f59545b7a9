// sort.hpp -- device-wide primitives for the broadphase: an exclusive scan
// (reduce, scan of tile sums, scan) and a stable LSD radix sort of
// (key, value) u32 pairs with wave64 ballot ranking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bg {

struct SortTemp {
    uint32_t* hist = nullptr;      // [256 * max_blocks]
    uint32_t* tile_sums = nullptr;  // [max_scan_tiles]
    uint32_t max_items = 0;
    uint32_t max_blocks = 0;
    uint32_t max_scan_tiles = 0;
};

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kScanTile = kSortTile;

hipError_t sort_temp_reserve(SortTemp& t, uint32_t max_items);
void sort_temp_free(SortTemp& t);

// out[i] = sum(in[0..i)) for i < n.  in and out may alias.
// out[i] = in[0] + ... + in[i-1] for i < n; with `last`, only the items up to
// index *last (a device-side count, read by the kernels) are scanned and
// the outputs past it are left alone.
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, SortTemp& t, hipStream_t s,
                              const uint32_t* last = nullptr);

// Stable sort by the low `bits` bits of keys.  Sorted data ends in
// (*keys, *vals); the pointers are swapped with the tmp buffers when the
// number of 8-bit passes is odd.
hipError_t radix_sort_pairs(uint32_t** keys, uint32_t** vals, uint32_t** keys_tmp, uint32_t** vals_tmp, uint32_t n,
                            int bits, SortTemp& t, hipStream_t s);

}  // namespace bg
