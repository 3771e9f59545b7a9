// batch_kernels.hip -- the persistent batched step: each world of a batch
// of single-sphere worlds with 1,025 to 4,096 bodies advances every substep
// of a launch inside ONE workgroup of 1,024 threads, its snapshot and cell
// grid in LDS and its bodies' state in registers (four per thread), so HBM
// sees the world once per launch instead of once per substep.  Workgroups
// take worlds from a queue ordered largest first until it is empty.  The
// other worlds stay on the grid path (one launch per substep over all of
// their bodies), which skips the persistent worlds' slots; contacts never
// cross worlds, so the two schedules share nothing but the step's counters.
//
// The substep is the fused sphere path's (physics.rs:401-421 for one
// collider per body, no springs or joints): gravity, contacts against the
// substep's snapshot summed in ascending partner id (physics.rs:241-317;
// push-outs by the pair rule of the sphere path, flagged), Verlet with the
// world's own old_dt (Q1, physics.rs:326-358), collider transform
// (physics.rs:360-366).  The same f32 operations in the same order as the
// grid path's kernel, so a world evolves bit-identically on either.  The
// grid is rebuilt every substep (bounding box, counts with LDS atomics, a
// block scan, a scatter): it decides which candidates are tested, never a
// result.
#include "common.hpp"
#include "device_state.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

constexpr uint32_t kPer = kPersistPer;  // bodies per thread
constexpr int kMaxWaves = 16;          // waves of the largest workgroup
constexpr int kContactsP = 4;  // contacts summed from registers; more take a selection pass in id order

struct Grid {
    float ox, oy, inv_cell;
    int32_t nx, ny;
    uint32_t ncells;
};

struct Shared {
    uint32_t* red;  // [4][waves]
    Grid* grid;
};

// Exclusive scan of cnt[0, n) by the workgroup of NT threads (n <= CELLS + 1).
template <uint32_t NT, uint32_t CELLS>
__device__ void scan_counts(uint32_t* cnt, uint32_t n, const Shared& sh) {
    constexpr uint32_t kItems = (CELLS + 1 + NT - 1) / NT;
    const uint32_t base = threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        v[k] = (base + k < n) ? cnt[base + k] : 0u;
        sum += v[k];
    }
    const uint32_t incl = wave_incl_scan_u32(sum);
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    if (lane == BG_WAVE - 1) sh.red[wid] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wid; ++w) off += sh.red[w];
    uint32_t run = off + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        if (base + k < n) cnt[base + k] = run;
        run += v[k];
    }
    __syncthreads();
}

// Cell grid of the snapshot SN[0, nb): a counting sort of the local indices
// by cell into SRT, cell starts in CNT.  Cells of side 2 r_max (doubled
// until the bounding box fits CELLS cells).
template <uint32_t NT, uint32_t PER, uint32_t CELLS, class Idx>
__device__ void build_grid(const float4* SN, uint32_t* CNT, Idx* SRT, uint32_t nb, float rmax, const Shared& sh) {
    constexpr int kWavesPT = NT / BG_WAVE;
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    uint32_t b0 = 0xFFFFFFFFu, b1 = 0xFFFFFFFFu, b2 = 0u, b3 = 0u;
    for (uint32_t li = threadIdx.x; li < nb; li += NT) {
        const float4 s = SN[li];
        if (!finitef(s.x) || !finitef(s.y)) continue;
        b0 = min(b0, f2ord(s.x));
        b1 = min(b1, f2ord(s.y));
        b2 = max(b2, f2ord(s.x));
        b3 = max(b3, f2ord(s.y));
    }
    b0 = wave_min_u32(b0);
    b1 = wave_min_u32(b1);
    b2 = wave_max_u32(b2);
    b3 = wave_max_u32(b3);
    if (lane == 0) {
        sh.red[wid] = b0;
        sh.red[kWavesPT + wid] = b1;
        sh.red[2 * kWavesPT + wid] = b2;
        sh.red[3 * kWavesPT + wid] = b3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < kWavesPT; ++w) {
            b0 = min(b0, sh.red[w]);
            b1 = min(b1, sh.red[kWavesPT + w]);
            b2 = max(b2, sh.red[2 * kWavesPT + w]);
            b3 = max(b3, sh.red[3 * kWavesPT + w]);
        }
        float minx = 0.0f, miny = 0.0f, maxx = 0.0f, maxy = 0.0f;
        if (b0 <= b2 && b1 <= b3) {
            minx = ord2f(b0);
            miny = ord2f(b1);
            maxx = ord2f(b2);
            maxy = ord2f(b3);
        }
        float cell = 2.0f * rmax;
        if (!(cell > 0.0f) || !finitef(cell)) cell = 1.0f;
        const float ex = maxx - minx, ey = maxy - miny;
        int32_t nx = 1, ny = 1;
        for (;;) {
            const float fx = floorf(ex / cell) + 1.0f, fy = floorf(ey / cell) + 1.0f;
            if (fx <= 65536.0f && fy <= 65536.0f && (uint64_t)fx * (uint64_t)fy <= (uint64_t)CELLS) {
                nx = (int32_t)fx;
                ny = (int32_t)fy;
                break;
            }
            cell = cell * 2.0f;
        }
        sh.grid->ox = minx;
        sh.grid->oy = miny;
        sh.grid->inv_cell = 1.0f / cell;
        sh.grid->nx = nx;
        sh.grid->ny = ny;
        sh.grid->ncells = (uint32_t)nx * (uint32_t)ny;
    }
    __syncthreads();
    const Grid g = *sh.grid;
    for (uint32_t i = threadIdx.x; i <= g.ncells; i += NT) CNT[i] = 0u;
    __syncthreads();
    uint32_t key[PER], rank[PER];
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t li = threadIdx.x + j * NT;
        key[j] = 0u;
        rank[j] = 0u;
        if (li >= nb) continue;
        const float4 s = SN[li];
        const int32_t cx = cell_coord(s.x, g.ox, g.inv_cell, g.nx);
        const int32_t cy = cell_coord(s.y, g.oy, g.inv_cell, g.ny);
        key[j] = (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
        rank[j] = atomicAdd(&CNT[key[j]], 1u);
    }
    __syncthreads();
    scan_counts<NT, CELLS>(CNT, g.ncells + 1, sh);
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t li = threadIdx.x + j * NT;
        if (li < nb) SRT[CNT[key[j]] + rank[j]] = (Idx)li;
    }
    __syncthreads();
}

// Local indices of the grid rows overlapping the square of half-width Q
// around (qx, qy), row by row.
template <class Idx, class F>
__device__ __forceinline__ void grid_query(const Grid& g, const uint32_t* CNT, const Idx* SRT, float qx, float qy,
                                           float Q, F&& visit) {
    const float R = Q + (1e-5f * (fabsf(qx) + fabsf(qy) + Q) + 1e-6f);
    int32_t cx0 = 0, cx1 = g.nx - 1, cy0 = 0, cy1 = g.ny - 1;
    if (R < 1e30f && finitef(qx) && finitef(qy)) {
        cx0 = cell_coord(qx - R, g.ox, g.inv_cell, g.nx);
        cx1 = cell_coord(qx + R, g.ox, g.inv_cell, g.nx);
        cy0 = cell_coord(qy - R, g.oy, g.inv_cell, g.ny);
        cy1 = cell_coord(qy + R, g.oy, g.inv_cell, g.ny);
    }
    for (int32_t cy = cy0; cy <= cy1; ++cy) {
        const uint32_t row = (uint32_t)cy * (uint32_t)g.nx;
        const uint32_t t0 = CNT[row + (uint32_t)cx0], t1 = CNT[row + (uint32_t)cx1 + 1u];
        for (uint32_t t = t0; t < t1; ++t) visit((uint32_t)SRT[t]);
    }
}

// One body's state between substeps.
struct Body {
    float2 p, q;
    float rot, omega;
    float m, gm;
    uint32_t fl;
};

// One world, every substep of the launch, by a workgroup of NT threads:
// bodies in registers (PER per thread), snapshot and grid in LDS
// (nb <= NT * PER).
template <uint32_t NT, uint32_t PER>
__device__ void run_world(const PersistArgs& k, uint32_t w, float4* SN, uint32_t* CNT, uint16_t* SRT,
                          const Shared& sh) {
    constexpr uint32_t CELLS = 2 * NT * PER;
    constexpr int kWavesPT = NT / BG_WAVE;
    const uint32_t b0 = k.wstart[w], nb = k.wstart[w + 1] - b0;
    Body reg[PER];
    float2 vlast[PER];
    float wold = 0.0f;  // Q1: this world's old_dt, held by the thread of its first dynamic body
    bool has_first = false;
    float rmax = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t li = threadIdx.x + j * NT;
        vlast[j] = make_float2(0.0f, 0.0f);
        reg[j].fl = 0u;
        if (li >= nb) continue;
        const uint32_t g = b0 + li;
        const float4 s = k.src.snap[g];
        rmax = fmaxf(rmax, s.z);
        SN[li] = make_float4(s.x, s.y, s.z, __uint_as_float(__float_as_uint(s.w) & ~kSnapOutlier));
        const float4 pp = k.src.pp[g];
        const float2 rw = k.src.rw[g], mg = k.src.mg[g];
        reg[j].p = make_float2(pp.x, pp.y);
        reg[j].q = make_float2(pp.z, pp.w);
        reg[j].rot = rw.x;
        reg[j].omega = rw.y;
        reg[j].m = mg.x;
        reg[j].gm = mg.y;
        reg[j].fl = k.src.bflags[g];
        if (reg[j].fl & BF_FIRST_DYN) {
            has_first = true;
            wold = k.wold_dt[w];
        }
    }
    // the world's largest radius (grid cell size only)
    {
        uint32_t r = wave_max_u32(__float_as_uint(rmax));  // radii are >= 0: integer order is float order
        const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
        if (lane == 0) sh.red[wid] = r;
        __syncthreads();
        r = 0u;
        for (int i = 0; i < kWavesPT; ++i) r = max(r, sh.red[i]);
        rmax = __uint_as_float(r);
        __syncthreads();
    }
    uint32_t count = 0, pushes = 0, flags = 0;
    const float h = k.h;
    for (uint32_t sub = 0; sub < k.nsub; ++sub) {
        const bool fresh = k.fresh && sub == 0;
        if (k.collisions) build_grid<NT, PER, CELLS>(SN, CNT, SRT, nb, rmax, sh);
        const Grid gr = *sh.grid;
        float2 xn[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t li = threadIdx.x + j * NT;
            xn[j] = make_float2(0.0f, 0.0f);
            if (li >= nb) continue;
            const uint32_t g = b0 + li;
            Body b = reg[j];
            const float4 own = SN[li];
            const uint32_t id = __float_as_uint(own.w);
            BodyState st;
            st.p = b.p;
            st.po = b.q;
            if (k.collisions && finitef(own.x) && finitef(own.y)) {
                const float ra = own.z;
                // reference predicate (physics.rs:255-269) against the snapshot, A = higher id
                auto enumerate = [&](auto&& sink) {
                    grid_query(gr, CNT, SRT, own.x, own.y, ra + rmax, [&](uint32_t lj) {
                        if (lj == li) return;
                        const float4 sb = SN[lj];
                        const float dx = own.x - sb.x, dy = own.y - sb.y;
                        const float rs = ra + sb.z;
                        const float dot = dx * dx + dy * dy;
                        if (dot > (rs * rs) * 1.00000095367431640625f) return;
                        const float dist = sqrtf(dot);
                        if (!(dist < rs)) return;
                        const uint32_t idb = __float_as_uint(sb.w);
                        const bool isA = id > idb;
                        sink(lj, idb, isA, isA ? dx : -dx, isA ? dy : -dy, dist, rs);
                    });
                };
                // correction on this body from one pair (physics.rs:270-317); a push-out is
                // folded into the same entry by the pair rule
                auto correction = [&](bool first, auto&& put, uint32_t lj, uint32_t idb, bool isA, float ax,
                                      float ay, float dist, float rs) {
                    float px = 0.0f, py = 0.0f;
                    const bool pushed = dist < 1e-6f;
                    if (pushed) {
                        ax = ax + 0.02f;
                        dist = vlen(ax, ay);
                        px = isA ? 0.01f : -0.01f;
                        py = isA ? 0.0f : -0.0f;
                        if (first) {
                            ++pushes;
                            flags |= INFO_PUSH_OUT;
                        }
                    }
                    const float mpart = k.src.mg[b0 + lj].x;
                    const float mA = isA ? b.m : mpart, mB = isA ? mpart : b.m;
                    const float ratio = 1.0f - mA / (mA + mB);
                    const float nx = ax / dist, ny = ay / dist;
                    if (first && (nx != nx || ny != ny)) flags |= ERR_NAN_NORMAL;
                    const float delta = rs - dist;
                    float cx, cy;
                    if (isA) {
                        const float kk = ratio * delta;
                        cx = kk * nx;
                        cy = kk * ny;
                    } else {
                        const float kk = (1.0f - ratio) * delta;
                        cx = -(kk * nx);
                        cy = -(kk * ny);
                    }
                    if (pushed) {
                        cx = px + cx;
                        cy = py + cy;
                    }
                    put(idb, cx, cy);
                };
                ContactList<kContactsP, uint32_t> L;
                L.init();
                enumerate([&](uint32_t lj, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                    if (isA) ++count;
                    correction(true, [&](uint32_t kk, float x, float y) { L.add(kk, x, y); }, lj, idb, isA, ax, ay,
                               dist, rs);
                });
                if (L.n > 0) {
                    uint32_t lastk = L.apply(st.p);
                    uint32_t done = L.n < (uint32_t)kContactsP ? L.n : (uint32_t)kContactsP;
                    while (done < L.n) {  // more than kContactsP: next id by selection, same order
                        uint32_t best = kNil;
                        float bx = 0.0f, by = 0.0f;
                        enumerate([&](uint32_t lj, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                            correction(
                                false,
                                [&](uint32_t kk, float x, float y) {
                                    if (kk > lastk && kk < best) {
                                        best = kk;
                                        bx = x;
                                        by = y;
                                    }
                                },
                                lj, idb, isA, ax, ay, dist, rs);
                        });
                        if (best == kNil) break;
                        st.p.x = st.p.x + bx;
                        st.p.y = st.p.y + by;
                        lastk = best;
                        ++done;
                    }
                }
            }
            // update_objects (physics.rs:326-358)
            st.rot = b.rot;
            st.omega = b.omega;
            float2 vel;
            uint32_t fl_out = b.fl;
            if (b.fl & BF_STATIC) {
                st.po = st.p;
                vel = make_float2(0.0f, 0.0f);
            } else {
                float2 acc = fresh ? k.src.acc[g] : make_float2(0.0f, 0.0f);
                acc.x = acc.x + k.gx * b.gm;
                acc.y = acc.y + k.gy * b.gm;
                const bool vr = (b.fl & BF_HAS_VREQ) != 0u;
                const float2 vreq = vr ? k.src.vreq[g] : make_float2(0.0f, 0.0f);
                const float tq = fresh ? k.src.torque[g] : 0.0f;
                const float toi = (fresh || !(b.fl & BF_INERTIA_POS)) ? tq / k.src.inertia[g] : 0.0f;
                float ratio = k.ratio_rest;
                if (b.fl & BF_FIRST_DYN) {  // Q1: this world's old_dt
                    ratio = h / wold;
                    wold = h;
                }
                verlet_dynamic(st, acc, toi, vr, vreq, h, ratio);
                vel = st.vel;
                if (vr) fl_out = b.fl & ~BF_HAS_VREQ;
            }
            // collider transform from the position (physics.rs:360-366)
            xn[j] = body_xform_zero_off(st.rot, st.p);
            b.p = st.p;
            b.q = st.po;
            b.rot = st.rot;
            b.omega = st.omega;
            b.fl = fl_out;
            reg[j] = b;
            vlast[j] = vel;
        }
        __syncthreads();  // every query of this substep has read the snapshot
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint32_t li = threadIdx.x + j * NT;
            if (li < nb) {
                const float4 s = SN[li];
                SN[li] = make_float4(xn[j].x, xn[j].y, s.z, s.w);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t li = threadIdx.x + j * NT;
        if (li >= nb) continue;
        const uint32_t g = b0 + li;
        const Body& b = reg[j];
        // every record of the body goes to dst (which the grid path may have swapped in)
        k.dst.pp[g] = make_float4(b.p.x, b.p.y, b.q.x, b.q.y);
        k.dst.rw[g] = make_float2(b.rot, b.omega);
        k.dst.bflags[g] = b.fl;
        k.dst.vel[g] = k.nsub > 0 ? vlast[j] : k.src.vel[g];
        k.dst.mg[g] = make_float2(b.m, b.gm);
        k.dst.inertia[g] = k.src.inertia[g];
        k.dst.torque[g] = (k.fresh && k.nsub > 0 && !(b.fl & BF_STATIC)) ? 0.0f : k.src.torque[g];
        if (b.fl & BF_HAS_VREQ) k.dst.vreq[g] = k.src.vreq[g];
        k.dst.snap[g] = SN[li];
    }
    if (has_first && k.nsub > 0) k.wold_dt[w] = wold;
    // counters: contacts into the slot of the world's first body, summed on the host
    if (flags) atomicOr(&k.accum->error_flags, flags);
    const uint32_t c_sum = wave_sum_u32(count);
    const uint32_t p_sum = wave_sum_u32(pushes);
    if ((threadIdx.x & (BG_WAVE - 1)) == 0) {
        if (c_sum) atomicAdd(&k.block_contacts[b0 / 64u], (unsigned long long)c_sum);
        if (p_sum) atomicAdd(&k.accum->push_outs, p_sum);
    }
    __syncthreads();  // LDS is rewritten by the next world
}

// Worlds of one size class (order[0, n) of the class, largest first) by
// workgroups of NT threads, taken from the queue until it is empty (every
// workgroup reaches the exit).
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_batch_persistent(PersistArgs k, const uint32_t* order, uint32_t n,
                                                         uint32_t* counter) {
    __shared__ float4 s_snap[NT * kPer];
    __shared__ uint32_t s_cnt[2 * NT * kPer + 1];
    __shared__ uint16_t s_sorted[NT * kPer];
    __shared__ uint32_t s_red[4 * kMaxWaves];
    __shared__ Grid s_grid;
    __shared__ uint32_t s_item;
    Shared sh;
    sh.red = s_red;
    sh.grid = &s_grid;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= n) break;
        run_world<NT, kPer>(k, order[item], s_snap, s_cnt, s_sorted, sh);
    }
}

}  // namespace

hipError_t launch_batch_persistent(const PersistArgs& k, hipStream_t s) {
    // the counters are reset by a fill kernel (graph-safe), then each class runs
    hipError_t e = launch_fill_u32(k.counter, 0u, kPersistClasses, s);
    if (e != hipSuccess) return e;
    for (uint32_t c = 0; c < kPersistClasses; ++c) {
        const uint32_t n = k.class_start[c + 1] - k.class_start[c];
        if (n == 0 || k.groups[c] == 0) continue;
        const uint32_t* order = k.order + k.class_start[c];
        uint32_t* counter = k.counter + c;
        static_assert(kPersistClasses == 1, "one launch per class");
        k_batch_persistent<kPersistClassThreads[0]><<<k.groups[c], kPersistClassThreads[0], 0, s>>>(k, order, n,
                                                                                                      counter);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace bg
