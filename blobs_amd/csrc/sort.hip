// sort.hip -- exclusive scan and stable LSD radix sort for gfx950.
//
// Scan: two launches with no inter-block waiting -- per-tile sums, then the
// per-tile scan, each block first reducing the sums of the tiles before it
// for its offset.  The input is read twice; nothing spins on another block's
// flag, so the cost is bandwidth and launch latency only.  The segmented form
// (a batch's per-world cell tables) skips tiles no world uses and scans the
// tile sums in a launch of their own.  Radix sort: 8-bit digits; per pass a digit
// histogram per 4096-item tile, the scan above over the digit-major matrix,
// and a scatter that ranks items with wave64 ballots (peer masks from 8
// ballots, popcount of lower lanes), orders the tile in LDS and writes
// digit runs out contiguously.
#include <stdlib.h>

#include "common.hpp"
#include "sort.hpp"

namespace bg {

namespace {

// Block-wide exclusive prefix of one value per thread; *total gets the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const int tid = threadIdx.x;
    const int lane = tid & (BG_WAVE - 1);
    const int wid = tid / BG_WAVE;
    const uint32_t incl = wave_incl_scan_u32(v);
    if (lane == BG_WAVE - 1) s_wave[wid] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / BG_WAVE; ++w) {
        off += (w < wid) ? s_wave[w] : 0u;
        tot += s_wave[w];
    }
    *total = tot;
    return off + incl - v;
}

// Items to scan: n, or fewer when the device holds the count (the last
// item's index at *last, e.g. a grid's cell count).
__device__ __forceinline__ uint32_t scan_len(uint32_t n, const uint32_t* last) {
    if (!last) return n;
    const uint32_t m = *last + 1u;
    return m < n ? m : n;
}

// Sum of the items of the tile starting at tile0 (coalesced: item k * 256 + tid),
// valid in thread 0.
__device__ __forceinline__ uint32_t tile_sum(const uint32_t* __restrict__ in, uint32_t n, uint32_t tile0,
                                             uint32_t* s_wave) {
    const uint32_t base = tile0 + threadIdx.x;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t i = base + k * kSortThreads;
        sum += (i < n) ? in[i] : 0u;
    }
    sum = wave_sum_u32(sum);
    if ((threadIdx.x & (BG_WAVE - 1)) == 0) s_wave[threadIdx.x / BG_WAVE] = sum;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / BG_WAVE; ++w) t += s_wave[w];
    return t;
}

// sums[tile] = sum of the tile's items.
__global__ __launch_bounds__(kSortThreads) void k_tile_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                            uint32_t* __restrict__ sums, const uint32_t* last) {
    __shared__ uint32_t s_wave[kSortThreads / BG_WAVE];
    n = scan_len(n, last);
    if (blockIdx.x * kScanTile >= n) {  // a tile past the items: nothing to read
        if (threadIdx.x == 0) sums[blockIdx.x] = 0u;
        return;
    }
    const uint32_t t = tile_sum(in, n, blockIdx.x * kScanTile, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// out = exclusive scan of in.  The tile goes through LDS (padded against
// bank conflicts) so global loads and stores are coalesced while thread t
// scans the 16 consecutive items [t*16, t*16+16).
__device__ __forceinline__ uint32_t scan_pad(uint32_t i) { return i + (i >> 5); }

// Exclusive scan of the tile starting at tile0, offset by tile_off.
__device__ __forceinline__ void scan_tile(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t tile0,
                                          uint32_t tile_off, uint32_t* s_wave, uint32_t* s_t) {
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = k * kSortThreads + threadIdx.x;
        s_t[scan_pad(idx)] = (tile0 + idx < n) ? in[tile0 + idx] : 0u;
    }
    __syncthreads();
    const uint32_t base = threadIdx.x * kSortItems;
    uint32_t v[kSortItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        v[k] = s_t[scan_pad(base + k)];
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = tile_off + block_excl_scan(sum, s_wave, &total);
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        s_t[scan_pad(base + k)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = k * kSortThreads + threadIdx.x;
        if (tile0 + idx < n) out[tile0 + idx] = s_t[scan_pad(idx)];
    }
}

__global__ __launch_bounds__(kSortThreads) void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n,
                                                             const uint32_t* __restrict__ sums, const uint32_t* last) {
    __shared__ uint32_t s_wave[kSortThreads / BG_WAVE];
    __shared__ uint32_t s_t[kScanTile + kScanTile / 32];
    n = scan_len(n, last);
    const uint32_t tile0 = blockIdx.x * kScanTile;
    if (tile0 >= n) return;  // the whole block leaves together (no barrier is skipped)
    // this tile's offset: the sum of the earlier tiles' sums, reduced by the block itself
    // (no separate pass over the tile sums)
    uint32_t pre = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kSortThreads) pre += sums[t];
    uint32_t tile_off;
    (void)block_excl_scan(pre, s_wave, &tile_off);
    __syncthreads();  // s_wave is reused below
    scan_tile(in, out, n, tile0, tile_off, s_wave, s_t);
}

// Whether some segment uses a cell of tile t: only the segment holding the
// tile's first cell can end before the tile; one starting inside it uses at
// least its first cell.
__device__ __forceinline__ bool tile_used(const ScanSegments& g, uint32_t t) {
    const uint32_t c0 = t * kScanTile, c1 = c0 + kScanTile;
    const uint32_t w = g.tile_seg[t];
    if (w >= g.n_seg) return false;
    if (g.seg_base[w] + g.seg_used[(size_t)w * g.seg_used_stride] + 1u > c0) return true;
    return w + 1u < g.n_seg && g.seg_base[w + 1u] < c1;
}

__global__ __launch_bounds__(kSortThreads) void k_tile_sums_seg(const uint32_t* __restrict__ in, uint32_t n,
                                                                uint32_t* __restrict__ sums, ScanSegments g) {
    __shared__ uint32_t s_wave[kSortThreads / BG_WAVE];
    if (!tile_used(g, blockIdx.x)) {
        if (threadIdx.x == 0) sums[blockIdx.x] = 0u;
        return;
    }
    const uint32_t t = tile_sum(in, n, blockIdx.x * kScanTile, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// sums[0..n) -> their exclusive scan, in place, by one block (kScanTile sums a pass).
__global__ __launch_bounds__(kSortThreads) void k_scan_sums(uint32_t* sums, uint32_t n) {
    __shared__ uint32_t s_wave[kSortThreads / BG_WAVE];
    uint32_t carry = 0;
    for (uint32_t p0 = 0; p0 < n; p0 += kScanTile) {
        const uint32_t base = p0 + threadIdx.x * kSortItems;
        uint32_t v[kSortItems];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            v[k] = base + k < n ? sums[base + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_excl_scan(sum, s_wave, &total);
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            if (base + k < n) sums[base + k] = run;
            run += v[k];
        }
        carry += total;
        __syncthreads();  // s_wave is reused by the next pass
    }
}

__global__ __launch_bounds__(kSortThreads) void k_scan_tiles_seg(const uint32_t* in, uint32_t* out, uint32_t n,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 ScanSegments g) {
    __shared__ uint32_t s_wave[kSortThreads / BG_WAVE];
    __shared__ uint32_t s_t[kScanTile + kScanTile / 32];
    if (!tile_used(g, blockIdx.x)) return;  // the whole block leaves together
    scan_tile(in, out, n, blockIdx.x * kScanTile, offsets[blockIdx.x], s_wave, s_t);
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* keys, uint32_t n, int shift,
                                                             uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        uint32_t i = base + k * kSortThreads + tid;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[tid * nblocks + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, uint32_t n,
                                                                int shift, const uint32_t* __restrict__ offsets,
                                                                uint32_t nblocks) {
    constexpr int kWaves = kSortThreads / BG_WAVE;
    __shared__ uint32_t wcount[kWaves][256];
    __shared__ uint32_t dbo[256];
    __shared__ uint32_t s_dwave[kWaves];
    __shared__ uint32_t stage_k[kSortTile];
    __shared__ uint32_t stage_v[kSortTile];
    const int tid = threadIdx.x;
    const int lane = tid & (BG_WAVE - 1);
    const int wid = tid / BG_WAVE;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
    __syncthreads();

    const uint32_t tile_base = blockIdx.x * kSortTile;
    const uint32_t sub_base = tile_base + wid * (kSortTile / kWaves);
    const uint64_t lt = lanemask_lt();
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = sub_base + r * BG_WAVE + lane;
        bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0u;
        val[r] = valid ? vals_in[i] : 0u;
        uint32_t d = (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bits = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bits : ~bits;
        }
        uint32_t below = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t prior = valid ? wcount[wid][d] : 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wcount[wid][d] = prior + cnt;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        rank[r] = prior + below;
    }
    __syncthreads();
    // per digit: exclusive prefix over waves, then over digits for the block
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        uint32_t c = wcount[w][tid];
        wcount[w][tid] = tot;
        tot += c;
    }
    uint32_t incl = wave_incl_scan_u32(tot);
    if (lane == BG_WAVE - 1) s_dwave[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        if (w < wid) woff += s_dwave[w];
    dbo[tid] = woff + incl - tot;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = sub_base + r * BG_WAVE + lane;
        if (i < n) {
            uint32_t d = (key[r] >> shift) & 255u;
            uint32_t p = dbo[d] + wcount[wid][d] + rank[r];
            stage_k[p] = key[r];
            stage_v[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t count = (n - tile_base) < (uint32_t)kSortTile ? (n - tile_base) : (uint32_t)kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        uint32_t p = k * kSortThreads + tid;
        if (p < count) {
            uint32_t kk = stage_k[p];
            uint32_t d = (kk >> shift) & 255u;
            uint32_t g = offsets[d * nblocks + blockIdx.x] + (p - dbo[d]);
            if (g < n) {  // always true for a consistent histogram/scan
                keys_out[g] = kk;
                vals_out[g] = stage_v[p];
            }
        }
    }
}

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

hipError_t sort_temp_reserve(SortTemp& t, uint32_t max_items) {
    if (max_items <= t.max_items && t.hist) return hipSuccess;
    sort_temp_free(t);
    uint32_t mi = max_items < 4096 ? 4096 : max_items;
    t.max_items = mi;
    t.max_blocks = div_up(mi, kSortTile);
    // the scan runs over 256 * max_blocks digit counters or over max_items
    uint32_t scan_items = 256u * t.max_blocks;
    if (scan_items < mi + 1) scan_items = mi + 1;
    t.max_scan_tiles = div_up(scan_items, kScanTile);
    hipError_t e = hipMalloc(&t.hist, sizeof(uint32_t) * 256u * t.max_blocks);
    if (e != hipSuccess) return e;
    return hipMalloc(&t.tile_sums, sizeof(uint32_t) * t.max_scan_tiles);
}

void sort_temp_free(SortTemp& t) {
    if (t.hist) (void)hipFree(t.hist);
    if (t.tile_sums) (void)hipFree(t.tile_sums);
    t = SortTemp{};
}

hipError_t exclusive_scan_u32_segments(const uint32_t* in, uint32_t* out, uint32_t n, SortTemp& t, hipStream_t s,
                                       const ScanSegments& g) {
    if (n == 0) return hipSuccess;
    const uint32_t ntiles = div_up(n, kScanTile);
    if (ntiles > t.max_scan_tiles) return hipErrorInvalidValue;
    k_tile_sums_seg<<<ntiles, kSortThreads, 0, s>>>(in, n, t.tile_sums, g);
    k_scan_sums<<<1, kSortThreads, 0, s>>>(t.tile_sums, ntiles);
    k_scan_tiles_seg<<<ntiles, kSortThreads, 0, s>>>(in, out, n, t.tile_sums, g);
    return hipGetLastError();
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, SortTemp& t, hipStream_t s,
                              const uint32_t* last) {
    if (n == 0) return hipSuccess;
    uint32_t ntiles = div_up(n, kScanTile);
    if (ntiles > t.max_scan_tiles) return hipErrorInvalidValue;
    k_tile_sums<<<ntiles, kSortThreads, 0, s>>>(in, n, t.tile_sums, last);
    k_scan_tiles<<<ntiles, kSortThreads, 0, s>>>(in, out, n, t.tile_sums, last);
    return hipGetLastError();
}

// ---- host-buffer test hooks (bgx_: internal, not part of the Physics ABI) --

extern "C" int bgx_radix_sort_u32(const uint32_t* keys, const uint32_t* vals, uint32_t n, int bits, uint32_t* keys_out,
                                  uint32_t* vals_out) {
    uint32_t *k = nullptr, *v = nullptr, *kt = nullptr, *vt = nullptr;
    SortTemp t;
    size_t bytes = sizeof(uint32_t) * (n ? n : 1);
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&k, bytes)) || (e = hipMalloc(&v, bytes)) || (e = hipMalloc(&kt, bytes)) ||
        (e = hipMalloc(&vt, bytes)) || (e = sort_temp_reserve(t, n)))
        goto done;
    if ((e = hipMemcpy(k, keys, sizeof(uint32_t) * n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(v, vals, sizeof(uint32_t) * n, hipMemcpyHostToDevice)))
        goto done;
    if ((e = radix_sort_pairs(&k, &v, &kt, &vt, n, bits, t, nullptr))) goto done;
    if ((e = hipMemcpy(keys_out, k, sizeof(uint32_t) * n, hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(vals_out, v, sizeof(uint32_t) * n, hipMemcpyDeviceToHost)))
        goto done;
done:
    (void)hipFree(k), (void)hipFree(v), (void)hipFree(kt), (void)hipFree(vt);
    sort_temp_free(t);
    return (int)e;
}

__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Copies `bytes` (a multiple of 16) between two device buffers with a plain
// grid-stride kernel, `repeats` times: the known-size reference that hardware
// byte counters are calibrated against.
extern "C" int bgx_copy_probe(uint64_t bytes, uint32_t repeats) {
    if (bytes == 0 || bytes % 16 != 0) return (int)hipErrorInvalidValue;
    uint4 *a = nullptr, *b = nullptr;
    hipError_t e = hipSuccess;
    const size_t n = bytes / 16;
    if ((e = hipMalloc(&a, bytes)) || (e = hipMalloc(&b, bytes))) goto done;
    if ((e = hipMemset(a, 1, bytes))) goto done;
    for (uint32_t r = 0; r < repeats; ++r) k_copy_probe<<<2048, 256>>>(a, b, n);
    if ((e = hipGetLastError())) goto done;
    e = hipDeviceSynchronize();
done:
    (void)hipFree(a), (void)hipFree(b);
    return (int)e;
}

extern "C" int bgx_exclusive_scan_u32(const uint32_t* in, uint32_t n, uint32_t* out) {
    uint32_t* d = nullptr;
    SortTemp t;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&d, sizeof(uint32_t) * (n ? n : 1))) || (e = sort_temp_reserve(t, n))) goto done;
    if ((e = hipMemcpy(d, in, sizeof(uint32_t) * n, hipMemcpyHostToDevice))) goto done;
    if ((e = exclusive_scan_u32(d, d, n, t, nullptr))) goto done;
    e = hipMemcpy(out, d, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
done:
    (void)hipFree(d);
    sort_temp_free(t);
    return (int)e;
}

// The segmented scan over a host table: segment w owns [seg_base[w], seg_base[w+1])
// and uses its first seg_used[w] + 1 cells; outputs elsewhere are left as `out` held them.
extern "C" int bgx_exclusive_scan_u32_segments(const uint32_t* in, uint32_t n, const uint32_t* seg_base,
                                               const uint32_t* seg_used, uint32_t n_seg, uint32_t* out) {
    if (n_seg == 0 || seg_base[n_seg] + 1u != n) return (int)hipErrorInvalidValue;
    const uint32_t ntiles = div_up(n, kScanTile);
    uint32_t *d_in = nullptr, *d_out = nullptr, *d_base = nullptr, *d_used = nullptr, *d_tile = nullptr;
    SortTemp t;
    hipError_t e = hipSuccess;
    uint32_t* tile = (uint32_t*)malloc(sizeof(uint32_t) * ntiles);
    ScanSegments g;
    if (!tile) return (int)hipErrorOutOfMemory;
    for (uint32_t k = 0, w = 0; k < ntiles; ++k) {
        while (w + 1u < n_seg && seg_base[w + 1] <= k * (uint32_t)kScanTile) ++w;
        tile[k] = w;
    }
    if ((e = hipMalloc(&d_in, sizeof(uint32_t) * n)) || (e = hipMalloc(&d_out, sizeof(uint32_t) * n)) ||
        (e = hipMalloc(&d_base, sizeof(uint32_t) * (n_seg + 1))) || (e = hipMalloc(&d_used, sizeof(uint32_t) * n_seg)) ||
        (e = hipMalloc(&d_tile, sizeof(uint32_t) * ntiles)) || (e = sort_temp_reserve(t, n)))
        goto done;
    if ((e = hipMemcpy(d_in, in, sizeof(uint32_t) * n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_out, out, sizeof(uint32_t) * n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_base, seg_base, sizeof(uint32_t) * (n_seg + 1), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_used, seg_used, sizeof(uint32_t) * n_seg, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_tile, tile, sizeof(uint32_t) * ntiles, hipMemcpyHostToDevice)))
        goto done;
    g.tile_seg = d_tile;
    g.seg_base = d_base;
    g.seg_used = d_used;
    g.seg_used_stride = 1;
    g.n_seg = n_seg;
    if ((e = exclusive_scan_u32_segments(d_in, d_out, n, t, nullptr, g))) goto done;
    if ((e = hipDeviceSynchronize())) goto done;
    e = hipMemcpy(out, d_out, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
done:
    free(tile);
    (void)hipFree(d_in), (void)hipFree(d_out), (void)hipFree(d_base), (void)hipFree(d_used), (void)hipFree(d_tile);
    sort_temp_free(t);
    return (int)e;
}

__global__ __launch_bounds__(256) void k_sqrt_lt_check(const float* d2, const float* r, uint32_t n, uint8_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (uint8_t)((sqrtf(d2[i]) < r[i] ? 1u : 0u) | (sqrt_lt(d2[i], r[i]) ? 2u : 0u));
}

// For each pair: bit 0 = (sqrtf(d2) < r), bit 1 = sqrt_lt(d2, r), both evaluated on the device.
extern "C" int bgx_sqrt_lt(const float* d2, const float* r, uint32_t n, uint8_t* out) {
    float *a = nullptr, *b = nullptr;
    uint8_t* o = nullptr;
    hipError_t e = hipSuccess;
    const size_t m = n ? n : 1;
    if ((e = hipMalloc(&a, 4 * m)) || (e = hipMalloc(&b, 4 * m)) || (e = hipMalloc(&o, m))) goto done;
    if ((e = hipMemcpy(a, d2, 4ull * n, hipMemcpyHostToDevice)) || (e = hipMemcpy(b, r, 4ull * n, hipMemcpyHostToDevice)))
        goto done;
    if (n) k_sqrt_lt_check<<<div_up(n, 256u), 256>>>(a, b, n, o);
    if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) goto done;
    e = hipMemcpy(out, o, n, hipMemcpyDeviceToHost);
done:
    (void)hipFree(a), (void)hipFree(b), (void)hipFree(o);
    return (int)e;
}

hipError_t radix_sort_pairs(uint32_t** keys, uint32_t** vals, uint32_t** keys_tmp, uint32_t** vals_tmp, uint32_t n,
                            int bits, SortTemp& t, hipStream_t s) {
    if (n <= 1 || bits <= 0) return hipSuccess;
    if (n > t.max_items) return hipErrorInvalidValue;
    uint32_t nblocks = div_up(n, kSortTile);
    for (int shift = 0; shift < bits; shift += 8) {
        k_radix_hist<<<nblocks, kSortThreads, 0, s>>>(*keys, n, shift, t.hist, nblocks);
        hipError_t e = exclusive_scan_u32(t.hist, t.hist, 256u * nblocks, t, s);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<nblocks, kSortThreads, 0, s>>>(*keys, *vals, *keys_tmp, *vals_tmp, n, shift, t.hist,
                                                         nblocks);
        uint32_t* tk = *keys;
        *keys = *keys_tmp;
        *keys_tmp = tk;
        uint32_t* tv = *vals;
        *vals = *vals_tmp;
        *vals_tmp = tv;
    }
    return hipGetLastError();
}

}  // namespace bg
