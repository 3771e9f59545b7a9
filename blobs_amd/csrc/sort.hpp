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

// A cell table cut into segments (one per world of a batch): segment w owns
// cells [seg_base[w], seg_base[w+1]) and uses the first seg_used[w * stride]
// + 1 of them this rebuild (its grid's cells and the end sentinel); the rest
// hold zero counts and nobody reads their starts.  tile_seg[t] is the segment
// holding the first cell of scan tile t.
struct ScanSegments {
    const uint32_t* tile_seg;
    const uint32_t* seg_base;
    const uint32_t* seg_used;
    uint32_t seg_used_stride;
    uint32_t n_seg;
};

// exclusive_scan_u32 over such a table: a tile that no segment uses is
// neither read nor written (its starts keep stale values), and the tile sums
// are scanned by one block of their own (three launches: the tile count of a
// large batch makes each tile reducing its predecessors' sums quadratic).
hipError_t exclusive_scan_u32_segments(const uint32_t* in, uint32_t* out, uint32_t n, SortTemp& t, hipStream_t s,
                                       const ScanSegments& g);

// Stable sort by the low `bits` bits of keys.  Sorted data ends in
// (*keys, *vals); the pointers are swapped with the tmp buffers when the
// number of 8-bit passes is odd.
hipError_t radix_sort_pairs(uint32_t** keys, uint32_t** vals, uint32_t** keys_tmp, uint32_t** vals_tmp, uint32_t n,
                            int bits, SortTemp& t, hipStream_t s);

}  // namespace bg
