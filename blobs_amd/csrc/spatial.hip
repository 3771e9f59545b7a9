// spatial.hip -- SpatialHash::query (spatial.rs:155-195) for a batch of
// query points on the GPU.
//
// The reference keeps, per cell (floor(x / cell_size), floor(y / cell_size))
// cast to i32, a set of point ids (spatial.rs:62-87); query(position, r)
// visits the 3x3 cells around the query's cell in the order of
// get_neighbor_cells (:140-153) and returns every point of those cells with
// |p - q|^2 <= (r + p.radius)^2 (:166-183).  Points more than one cell away
// are never returned, whatever r is.
//
// Here the caller passes the cell memberships as records (position, radius,
// the cell it is filed under, id): normally one per point, filed under its
// own position's cell; SpatialHash::insert_with_id on an existing id leaves
// the old cell's entry in place, which is a second record of the same id.
// One call builds a hash table of the occupied cells (open addressing over a
// power-of-two table at least twice the record count), files the records
// into per-cell runs (counting sort by table slot, each run ordered by record
// index), then answers every query in two passes: count, exclusive scan into
// the caller's offsets, write.  Hits of a query come in neighbour-cell order
// and, inside a cell, in record order (the reference's HashSet order is
// unspecified).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/blobs_gpu.h"
#include "common.hpp"
#include "sort.hpp"
#include "step_common.hpp"

namespace bg {

void set_last_error(const std::string& msg);

namespace {

int fail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}


// Key of cell (cx, cy); the one cell whose key equals kEmpty is filed in the
// extra slot T of the table instead.
constexpr unsigned long long kEmpty = 0x8000000080000000ull;

__device__ __forceinline__ unsigned long long cell_key(int32_t cx, int32_t cy) {
    return ((unsigned long long)(uint32_t)cx << 32) | (uint32_t)cy;
}

__device__ __forceinline__ uint32_t key_hash(unsigned long long k, uint32_t mask) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}

// (f32 / cell_size).floor() as i32: Rust's saturating cast, NaN -> 0 (spatial.rs:55-60).
__device__ __forceinline__ int32_t cell_coord_sat(float x, float cs) {
    const float f = floorf(x / cs);
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)f;
}

struct SpatialScratch {
    unsigned long long* keys;  // [T] cell key of each slot (kEmpty: free)
    uint32_t* start;           // [T + 2] record count per slot, then (after the scan) run starts
    uint32_t* slot;            // [n] slot of each record
    uint32_t* rank;            // [n] rank of a record within its slot (atomic order)
    uint32_t* order;           // [n] record indices, grouped by slot
    uint32_t T;                // table size (power of two)
    SortTemp scan;
};

__global__ void k_sh_clear(SpatialScratch s) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.T + 2; i += gridDim.x * blockDim.x) {
        if (i < s.T) s.keys[i] = kEmpty;
        s.start[i] = 0u;
    }
}

__global__ void k_sh_insert(const bg_spatial_point* __restrict__ pts, uint32_t n, SpatialScratch s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = cell_key(pts[i].cell[0], pts[i].cell[1]);
    uint32_t h;
    if (key == kEmpty) {
        h = s.T;
    } else {
        const uint32_t mask = s.T - 1u;
        h = key_hash(key, mask);
        for (;;) {  // at most n of the T >= 2n slots are taken: a free one is always found
            const unsigned long long old = atomicCAS(&s.keys[h], kEmpty, key);
            if (old == kEmpty || old == key) break;
            h = (h + 1u) & mask;
        }
    }
    s.slot[i] = h;
    s.rank[i] = atomicAdd(&s.start[h], 1u);
}

__global__ void k_sh_scatter(uint32_t n, SpatialScratch s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s.order[s.start[s.slot[i]] + s.rank[i]] = i;
}

// Each run into record order (insertion sort: runs are a cell's points).
__global__ void k_sh_sort_runs(SpatialScratch s) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h > s.T) return;
    const uint32_t a = s.start[h], b = s.start[h + 1];
    for (uint32_t i = a + 1; i < b; ++i) {
        const uint32_t v = s.order[i];
        uint32_t j = i;
        while (j > a && s.order[j - 1] > v) {
            s.order[j] = s.order[j - 1];
            --j;
        }
        s.order[j] = v;
    }
}

// Run of cell (cx, cy): [*a, *b), empty when no record is filed there.
__device__ __forceinline__ void cell_run(const SpatialScratch& s, int32_t cx, int32_t cy, uint32_t* a, uint32_t* b) {
    const unsigned long long key = cell_key(cx, cy);
    uint32_t h;
    if (key == kEmpty) {
        h = s.T;
    } else {
        const uint32_t mask = s.T - 1u;
        h = key_hash(key, mask);
        for (;;) {  // a free slot ends the probe
            const unsigned long long k = s.keys[h];
            if (k == key) break;
            if (k == kEmpty) {
                *a = *b = 0u;
                return;
            }
            h = (h + 1u) & mask;
        }
    }
    *a = s.start[h];
    *b = s.start[h + 1];
}

// One query (spatial.rs:155-195): emit(record index) for each hit, in order.
template <class F>
__device__ __forceinline__ void run_query(const bg_spatial_point* __restrict__ pts, const SpatialScratch& s, float cs,
                                          const bg_spatial_query& q, F&& emit) {
    const int32_t cx = cell_coord_sat(q.position[0], cs), cy = cell_coord_sat(q.position[1], cs);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {  // get_neighbor_cells order (spatial.rs:140-153), wrapping i32
            uint32_t a, b;
            cell_run(s, (int32_t)((uint32_t)cx + (uint32_t)dx), (int32_t)((uint32_t)cy + (uint32_t)dy), &a, &b);
            for (uint32_t t = a; t < b; ++t) {
                const uint32_t i = s.order[t];
                const bg_spatial_point p = pts[i];
                const float ex = p.position[0] - q.position[0], ey = p.position[1] - q.position[1];
                const float dist = q.radius + p.radius;  // :168-171
                if (ex * ex + ey * ey <= dist * dist) emit(i, p);
            }
        }
}

__global__ void k_sh_count(const bg_spatial_point* __restrict__ pts, const bg_spatial_query* __restrict__ qs,
                           uint32_t nq, SpatialScratch s, float cs, uint32_t* counts) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nq) return;
    uint32_t c = 0;
    if (q < nq) run_query(pts, s, cs, qs[q], [&](uint32_t, const bg_spatial_point&) { ++c; });
    counts[q] = c;  // counts[nq] = 0: the scan's total lands there
}

__global__ void k_sh_fill(const bg_spatial_point* __restrict__ pts, const bg_spatial_query* __restrict__ qs,
                          uint32_t nq, SpatialScratch s, float cs, const uint32_t* __restrict__ offsets,
                          bg_spatial_hit* hits, uint32_t cap) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t o = offsets[q];
    run_query(pts, s, cs, qs[q], [&](uint32_t i, const bg_spatial_point& p) {
        if (o < cap) {
            bg_spatial_hit h;
            h.id = p.id;
            h.position[0] = p.position[0];
            h.position[1] = p.position[1];
            h.radius = p.radius;
            h.point = i;
            hits[o] = h;
        }
        ++o;
    });
}

constexpr uint32_t table_size(uint32_t n) {
    uint32_t t = 64;
    while (t < 2u * n) t <<= 1;
    return t;
}

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of scratch for n records and nq queries, and the views into it.
constexpr size_t scratch_bytes(uint32_t n, uint32_t nq) {
    const uint32_t T = table_size(n);
    const uint32_t m = T + 2 > nq + 1 ? T + 2 : nq + 1;
    const uint32_t tiles = (m + kScanTile - 1) / kScanTile;
    return align256(sizeof(unsigned long long) * T) + align256(sizeof(uint32_t) * (T + 2)) +
           3 * align256(sizeof(uint32_t) * (n > 1u ? n : 1u)) + align256(sizeof(uint32_t) * tiles);
}

// The no-allocation entry point's scratch: one array per GPU, held until the
// process ends, large enough for kNoAllocMaxPoints records and
// kNoAllocMaxQueries queries (36 MiB).  Host code never passes `scratch`
// itself to a kernel (in host code the name is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress returns per device.
constexpr uint32_t kNoAllocMaxPoints = 1u << 20, kNoAllocMaxQueries = 1u << 20;
constexpr size_t kScratchBytes = scratch_bytes(kNoAllocMaxPoints, kNoAllocMaxQueries);
static_assert(kScratchBytes == 37751296, "36 MiB of scratch");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_on_current_device(char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return BG_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return fail(BG_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = (char*)g_scratch_ptr[dev];
    return BG_OK;
}

SpatialScratch carve(char* base, uint32_t n, uint32_t nq) {
    SpatialScratch s;
    s.T = table_size(n);
    const uint32_t tiles = div_up(std::max(s.T + 2, nq + 1), kScanTile);
    char* p = base;
    s.keys = (unsigned long long*)p;
    p += align256(sizeof(unsigned long long) * s.T);
    s.start = (uint32_t*)p;
    p += align256(sizeof(uint32_t) * (s.T + 2));
    s.slot = (uint32_t*)p;
    p += align256(sizeof(uint32_t) * std::max(n, 1u));
    s.rank = (uint32_t*)p;
    p += align256(sizeof(uint32_t) * std::max(n, 1u));
    s.order = (uint32_t*)p;
    p += align256(sizeof(uint32_t) * std::max(n, 1u));
    s.scan.tile_sums = (uint32_t*)p;
    s.scan.max_scan_tiles = tiles;
    s.scan.hist = nullptr;
    s.scan.max_items = 0;
    s.scan.max_blocks = 0;
    return s;
}

// The whole query on `stream`, scratch views given; leaves offsets[nq] = total.
hipError_t spatial_query_on(const SpatialScratch& s, float cs, const bg_spatial_point* pts, uint32_t n,
                            const bg_spatial_query* qs, uint32_t nq, uint32_t* offsets, bg_spatial_hit* hits,
                            uint32_t cap, hipStream_t stream) {
    const uint32_t B = 256;
    k_sh_clear<<<std::min(div_up(s.T + 2, B), 4096u), B, 0, stream>>>(s);
    if (n > 0) k_sh_insert<<<div_up(n, B), B, 0, stream>>>(pts, n, s);
    hipError_t e = exclusive_scan_u32(s.start, s.start, s.T + 2, const_cast<SortTemp&>(s.scan), stream);
    if (e != hipSuccess) return e;
    if (n > 0) {
        k_sh_scatter<<<div_up(n, B), B, 0, stream>>>(n, s);
        k_sh_sort_runs<<<div_up(s.T + 1, B), B, 0, stream>>>(s);
    }
    k_sh_count<<<div_up(nq + 1, B), B, 0, stream>>>(pts, qs, nq, s, cs, offsets);
    e = exclusive_scan_u32(offsets, offsets, nq + 1, const_cast<SortTemp&>(s.scan), stream);
    if (e != hipSuccess) return e;
    if (nq > 0) k_sh_fill<<<div_up(nq, B), B, 0, stream>>>(pts, qs, nq, s, cs, offsets, hits, cap);
    return hipGetLastError();
}

int check_args(float cs, const bg_spatial_point* pts, uint32_t n, const bg_spatial_query* qs, uint32_t nq,
               const uint32_t* offsets, const bg_spatial_hit* hits, uint32_t cap) {
    if (!offsets || (n > 0 && !pts) || (nq > 0 && !qs) || (cap > 0 && !hits))
        return fail(BG_ERR_INVALID_ARGUMENT, "spatial query: offsets, and points / queries / hits when counted, are required");
    if (n > (1u << 30) || nq > (1u << 30)) return fail(BG_ERR_INVALID_ARGUMENT, "spatial query: more than 2^30 records");
    (void)cs;
    return BG_OK;
}

int finish(hipError_t e, const uint32_t* offsets, uint32_t nq, uint32_t cap, uint64_t* total_out) {
    if (e != hipSuccess) return fail(BG_ERR_HIP, std::string("spatial query: ") + hipGetErrorString(e));
    uint32_t total = 0;
    e = hipMemcpy(&total, offsets + nq, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(BG_ERR_HIP, std::string("spatial query: total: ") + hipGetErrorString(e));
    if (total_out) *total_out = total;
    return total > cap ? fail(BG_ERR_CAPACITY, "spatial query: more hits than hit_cap (offsets are complete)") : BG_OK;
}

}  // namespace

}  // namespace bg

extern "C" {

int bg_spatial_hash_query(float cell_size, const bg_spatial_point* points, uint32_t n_points,
                     const bg_spatial_query* queries, uint32_t n_queries, uint32_t* offsets, bg_spatial_hit* hits,
                     uint32_t hit_cap, uint64_t* total_hits) {
    using namespace bg;
    if (int st = check_args(cell_size, points, n_points, queries, n_queries, offsets, hits, hit_cap)) return st;
    char* mem = nullptr;
    const hipError_t ea = hipMalloc((void**)&mem, scratch_bytes(n_points, n_queries));
    if (ea != hipSuccess) return fail(BG_ERR_HIP, std::string("spatial query: scratch: ") + hipGetErrorString(ea));
    const SpatialScratch s = carve(mem, n_points, n_queries);
    hipError_t e = spatial_query_on(s, cell_size, points, n_points, queries, n_queries, offsets, hits, hit_cap,
                                    nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    const int st = finish(e, offsets, n_queries, hit_cap, total_hits);
    (void)hipFree(mem);
    return st;
}

int bg_spatial_hash_query_host(float cell_size, const bg_spatial_point* points, uint32_t n_points,
                               const bg_spatial_query* queries, uint32_t n_queries, uint32_t* offsets,
                               bg_spatial_hit* hits, uint32_t hit_cap, uint64_t* total_hits) {
    using namespace bg;
    if (int st = check_args(cell_size, points, n_points, queries, n_queries, offsets, hits, hit_cap)) return st;
    const size_t bp = sizeof(bg_spatial_point) * std::max(n_points, 1u);
    const size_t bq = sizeof(bg_spatial_query) * std::max(n_queries, 1u);
    const size_t bo = sizeof(uint32_t) * ((size_t)n_queries + 1);
    const size_t bh = sizeof(bg_spatial_hit) * std::max(hit_cap, 1u);
    char* dev = nullptr;
    hipError_t e = hipMalloc((void**)&dev, bp + bq + bo + bh + 4 * 256);
    if (e != hipSuccess) return fail(BG_ERR_HIP, std::string("spatial query: buffers: ") + hipGetErrorString(e));
    bg_spatial_point* dp = (bg_spatial_point*)dev;
    bg_spatial_query* dq = (bg_spatial_query*)(dev + align256(bp));
    uint32_t* doff = (uint32_t*)((char*)dq + align256(bq));
    bg_spatial_hit* dh = (bg_spatial_hit*)((char*)doff + align256(bo));
    int st = BG_OK;
    if ((n_points && (e = hipMemcpy(dp, points, sizeof(bg_spatial_point) * n_points, hipMemcpyHostToDevice))) ||
        (n_queries && (e = hipMemcpy(dq, queries, sizeof(bg_spatial_query) * n_queries, hipMemcpyHostToDevice)))) {
        st = fail(BG_ERR_HIP, std::string("spatial query: upload: ") + hipGetErrorString(e));
    } else {
        uint64_t total = 0;
        st = bg_spatial_hash_query(cell_size, dp, n_points, dq, n_queries, doff, dh, hit_cap, &total);
        if (st == BG_OK || st == BG_ERR_CAPACITY) {
            const size_t nh = std::min<uint64_t>(total, hit_cap);
            if ((e = hipMemcpy(offsets, doff, bo, hipMemcpyDeviceToHost)) ||
                (nh && (e = hipMemcpy(hits, dh, sizeof(bg_spatial_hit) * nh, hipMemcpyDeviceToHost))))
                st = fail(BG_ERR_HIP, std::string("spatial query: download: ") + hipGetErrorString(e));
            if (total_hits) *total_hits = total;
        }
    }
    (void)hipFree(dev);
    return st;
}

// ---- device-buffer test hooks (bgx_: internal, not part of the Physics ABI) ----
int bgx_device_alloc(uint64_t bytes, void** out) {
    return hipMalloc(out, bytes ? bytes : 1) == hipSuccess ? BG_OK : BG_ERR_HIP;
}
int bgx_device_free(void* p) { return hipFree(p) == hipSuccess ? BG_OK : BG_ERR_HIP; }
int bgx_copy_to_device(void* dst, const void* src, uint64_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? BG_OK : BG_ERR_HIP;
}
int bgx_copy_to_host(void* dst, const void* src, uint64_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? BG_OK : BG_ERR_HIP;
}

int bg_spatial_hash_query_noalloc(float cell_size, const bg_spatial_point* points, uint32_t n_points,
                                  const bg_spatial_query* queries, uint32_t n_queries, uint32_t* offsets,
                                  bg_spatial_hit* hits, uint32_t hit_cap, uint64_t* total_hits) {
    using namespace bg;
    if (int st = check_args(cell_size, points, n_points, queries, n_queries, offsets, hits, hit_cap)) return st;
    if (n_points > kNoAllocMaxPoints || n_queries > kNoAllocMaxQueries)
        return fail(BG_ERR_CAPACITY, "spatial query without allocation: at most 1,048,576 points and 1,048,576 queries");
    char* mem = nullptr;
    if (int st = scratch_on_current_device(&mem)) return st;
    // every table word the call reads is cleared by its first kernel (k_sh_clear)
    const SpatialScratch s = carve(mem, n_points, n_queries);
    hipError_t e = spatial_query_on(s, cell_size, points, n_points, queries, n_queries, offsets, hits, hit_cap,
                                    nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return finish(e, offsets, n_queries, hit_cap, total_hits);
}

}  // extern "C"
