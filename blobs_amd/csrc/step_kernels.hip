// step_kernels.hip -- gfx950 kernels of one `Physics::step` substep.
//
// Reference stage order per substep (physics.rs:401-421): gravity, springs,
// contacts, fixed joints, Verlet update (+ collider transforms), circle
// constraints.  Two device paths implement it:
//
//  * general path (any world): k_forces (gravity + springs, per body),
//    k_contacts (per body, over its colliders), k_joints (one thread per
//    independent joint component), k_integrate (Verlet, collider transforms,
//    constraints, per body);
//  * fused sphere path (one collider per body, no springs/joints): all stages
//    in one kernel per substep over bodies kept in broadphase-cell order.
//
// Contacts are order-independent because the reference tests the collider
// snapshot written by the previous Verlet (physics.rs:264, 360-366), so the
// pair set and every correction vector depend only on the snapshot; only the
// summation order on a body position matters, and each body thread sums its
// own corrections in the reference's (i, j<i) lexicographic order.
#include <float.h>

#include "common.hpp"
#include "device_state.hpp"
#include "sort.hpp"

namespace bg {

namespace {

constexpr int kThreads = 256;

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---- sorted register list of contact corrections -------------------------
// Keeps the K smallest keys in ascending order (unrolled insertion with
// compile-time indices, so it stays in VGPRs); `n` counts every insertion.
template <int K, typename Key>
struct ContactList {
    Key key[K];
    float cx[K], cy[K];
    uint32_t n;
    __device__ __forceinline__ void init() {
        n = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            key[i] = (Key)~(Key)0;
            cx[i] = 0.0f;
            cy[i] = 0.0f;
        }
    }
    __device__ __forceinline__ void add(Key k, float x, float y) {
        Key ck = k;
        float ccx = x, ccy = y;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            bool sw = ck < key[i];
            Key tk = key[i];
            float tx = cx[i], ty = cy[i];
            key[i] = sw ? ck : tk;
            cx[i] = sw ? ccx : tx;
            cy[i] = sw ? ccy : ty;
            ck = sw ? tk : ck;
            ccx = sw ? tx : ccx;
            ccy = sw ? ty : ccy;
        }
        ++n;
    }
    // Adds the stored corrections in key order; returns the last key applied.
    __device__ __forceinline__ Key apply(float2& p) const {
        Key last = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if ((uint32_t)i < n) {
                p.x = p.x + cx[i];
                p.y = p.y + cy[i];
                last = key[i];
            }
        }
        return last;
    }
};

// Visit every slot in the cells overlapping the square of half-width Q
// around (qx, qy).  Cell coordinates are monotone in x, so the visited set
// is a superset of all snapshots within distance Q of the query point.
template <class F>
__device__ __forceinline__ void scan_stencil(const GridParams& g, const uint32_t* __restrict__ cell_start, uint32_t n,
                                             float qx, float qy, float Q, F&& visit) {
    float R = Q + (1e-5f * (fabsf(qx) + fabsf(qy) + Q) + 1e-6f);
    int32_t cx0, cx1, cy0, cy1;
    if (R < 1e30f && finitef(qx) && finitef(qy)) {
        cx0 = cell_coord(qx - R, g.ox, g.inv_cell, g.nx);
        cx1 = cell_coord(qx + R, g.ox, g.inv_cell, g.nx);
        cy0 = cell_coord(qy - R, g.oy, g.inv_cell, g.ny);
        cy1 = cell_coord(qy + R, g.oy, g.inv_cell, g.ny);
    } else {
        cx0 = 0;
        cx1 = g.nx - 1;
        cy0 = 0;
        cy1 = g.ny - 1;
    }
    for (int32_t cy = cy0; cy <= cy1; ++cy) {
        uint32_t row = (uint32_t)cy * (uint32_t)g.nx;
        uint32_t t0 = cell_start[row + (uint32_t)cx0];
        uint32_t t1 = cell_start[row + (uint32_t)cx1 + 1u];
        t1 = t1 < n ? t1 : n;  // cell_start is a monotone table in [0, n]
        for (uint32_t t = t0; t < t1; ++t) visit(t);
    }
}

__device__ __forceinline__ float read_dmax(const StepAccum* acc) {
    float d = __uint_as_float(__hip_atomic_load(&acc->dmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    return d;
}

// Block-level reduction of the per-thread step records, one atomic each.
struct LocalAccum {
    float dmax;
    float minx, miny, maxx, maxy;
    uint32_t contacts;
    uint32_t flags;
    uint32_t push_outs;
    __device__ __forceinline__ void init() {
        dmax = 0.0f;
        minx = miny = FLT_MAX;
        maxx = maxy = -FLT_MAX;
        contacts = 0;
        flags = 0;
        push_outs = 0;
    }
    __device__ __forceinline__ void point(float x, float y, float x0, float y0, float dx, float dy) {
        if (finitef(x) && finitef(y)) {
            minx = fminf(minx, x);
            maxx = fmaxf(maxx, x);
            miny = fminf(miny, y);
            maxy = fmaxf(maxy, y);
            float ex = (x - x0) - dx, ey = (y - y0) - dy;
            float d = sqrtf(ex * ex + ey * ey);
            if (finitef(d)) dmax = fmaxf(dmax, d);
        }
    }
};

// Only blocks that extend the running extremum issue an atomic (a relaxed
// read first): thousands of blocks hitting one word serialize otherwise.
__device__ __forceinline__ void atomic_max_if_greater(uint32_t* p, uint32_t v) {
    if (v > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, v);
}
__device__ __forceinline__ void atomic_min_if_less(uint32_t* p, uint32_t v) {
    if (v < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(p, v);
}

// Contacts go to a per-block slot (summed on the host when queried), never
// to one shared counter.
__device__ void flush_accum(const LocalAccum& l, StepAccum* g, unsigned long long* block_contacts, uint32_t parity,
                            bool with_points) {
    __shared__ uint32_t s_red[8][kThreads / BG_WAVE];
    __shared__ unsigned long long s_cnt[kThreads / BG_WAVE];
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    uint32_t v[8];
    v[0] = with_points ? __float_as_uint(l.dmax) : 0u;
    v[1] = f2ord(l.minx);
    v[2] = f2ord(l.miny);
    v[3] = f2ord(l.maxx);
    v[4] = f2ord(l.maxy);
    v[5] = l.flags;
    v[6] = l.push_outs;
    v[0] = wave_max_u32(v[0]);
    v[1] = wave_min_u32(v[1]);
    v[2] = wave_min_u32(v[2]);
    v[3] = wave_max_u32(v[3]);
    v[4] = wave_max_u32(v[4]);
    for (int o = 32; o > 0; o >>= 1) v[5] |= __shfl_xor(v[5], o, BG_WAVE);
    v[6] = wave_sum_u32(v[6]);
    uint64_t c = wave_sum_u64((uint64_t)l.contacts);
    if (lane == 0) {
        for (int k = 0; k < 7; ++k) s_red[k][wid] = v[k];
        s_cnt[wid] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / BG_WAVE; ++w) {
            s_red[0][0] = s_red[0][w] > s_red[0][0] ? s_red[0][w] : s_red[0][0];
            s_red[1][0] = s_red[1][w] < s_red[1][0] ? s_red[1][w] : s_red[1][0];
            s_red[2][0] = s_red[2][w] < s_red[2][0] ? s_red[2][w] : s_red[2][0];
            s_red[3][0] = s_red[3][w] > s_red[3][0] ? s_red[3][w] : s_red[3][0];
            s_red[4][0] = s_red[4][w] > s_red[4][0] ? s_red[4][w] : s_red[4][0];
            s_red[5][0] |= s_red[5][w];
            s_red[6][0] += s_red[6][w];
            s_cnt[0] += s_cnt[w];
        }
        if (with_points) {
            atomic_max_if_greater(&g->dmax, s_red[0][0]);
            atomic_min_if_less(&g->bbox[parity][0], s_red[1][0]);
            atomic_min_if_less(&g->bbox[parity][1], s_red[2][0]);
            atomic_max_if_greater(&g->bbox[parity][2], s_red[3][0]);
            atomic_max_if_greater(&g->bbox[parity][3], s_red[4][0]);
        }
        if (s_red[5][0]) atomicOr(&g->error_flags, s_red[5][0]);
        if (s_red[6][0]) atomicAdd(&g->push_outs, s_red[6][0]);
        if (s_cnt[0]) block_contacts[blockIdx.x] += s_cnt[0];
    }
}

// ---- Verlet update of one body, physics.rs:326-358 -----------------------
struct BodyState {
    float2 p, po, vel;
    float rot, omega;
};

__device__ __forceinline__ void verlet_dynamic(BodyState& s, float2 acc, float torque_over_inertia, bool has_vreq,
                                               float2 vreq, float h, float ratio) {
    if (has_vreq) {  // :334-336
        s.po.x = s.p.x - vreq.x * h;
        s.po.y = s.p.y - vreq.y * h;
    }
    float dx = (s.p.x - s.po.x) * ratio;  // :338
    float dy = (s.p.y - s.po.y) * ratio;
    s.po = s.p;                                    // :343
    s.p.x = s.p.x + (dx + (acc.x * h) * h);        // :344
    s.p.y = s.p.y + (dy + (acc.y * h) * h);
    s.omega = s.omega + torque_over_inertia * h;   // :346
    s.rot = s.rot + s.omega * h;                   // :349
    s.vel.x = dx / h;                              // :357
    s.vel.y = dy / h;
}

// Circle constraints, physics.rs:377-395 (centre clamp, static bodies too).
__device__ __forceinline__ void apply_constraints(float2& p, const float2* __restrict__ con_pos,
                                                  const float* __restrict__ con_r, uint32_t nk) {
    for (uint32_t k = 0; k < nk; ++k) {
        float2 c = con_pos[k];
        float r = con_r[k];
        float tx = p.x - c.x, ty = p.y - c.y;
        float dist = vlen(tx, ty);
        if (dist > r) {
            float nx = tx / dist, ny = ty / dist;
            p.x = c.x + nx * r;
            p.y = c.y + ny * r;
        }
    }
}

// ============================================================ broadphase

__device__ __forceinline__ GridParams grid_from_bbox(const uint32_t* bb, float cell, uint32_t cells_cap) {
    GridParams g;
    float minx = ord2f(bb[0]), miny = ord2f(bb[1]), maxx = ord2f(bb[2]), maxy = ord2f(bb[3]);
    if (!(minx <= maxx) || !(miny <= maxy) || !finitef(minx) || !finitef(maxx) || !finitef(miny) ||
        !finitef(maxy)) {
        minx = miny = 0.0f;
        maxx = maxy = 0.0f;
    }
    if (!(cell > 0.0f) || !finitef(cell)) cell = 1.0f;
    float ex = maxx - minx, ey = maxy - miny;
    int32_t nx, ny;
    for (;;) {
        float fx = floorf(ex / cell) + 1.0f, fy = floorf(ey / cell) + 1.0f;
        if (fx * fy <= (float)cells_cap && fx <= 65536.0f && fy <= 65536.0f) {
            nx = (int32_t)fx;
            ny = (int32_t)fy;
            break;
        }
        cell = cell * 2.0f;
    }
    g.ox = minx;
    g.oy = miny;
    g.cell = cell;
    g.inv_cell = 1.0f / cell;
    g.nx = nx;
    g.ny = ny;
    g.ncells = (uint32_t)nx * (uint32_t)ny;
    g.pad = 0;
    return g;
}

__global__ __launch_bounds__(kThreads) void k_rebuild_keys(const float2* __restrict__ abs, uint32_t nc,
                                                           StepAccum* accum, uint32_t parity, float cell,
                                                           uint32_t cells_cap, GridParams* grid_out,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                           float2* __restrict__ x0) {
    const GridParams g = grid_from_bbox(accum->bbox[parity], cell, cells_cap);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *grid_out = g;
        accum->bbox[parity ^ 1][0] = 0xFFFFFFFFu;
        accum->bbox[parity ^ 1][1] = 0xFFFFFFFFu;
        accum->bbox[parity ^ 1][2] = 0u;
        accum->bbox[parity ^ 1][3] = 0u;
        accum->dmax = 0u;
    }
    if (i < nc) {
        float2 x = abs[i];
        int32_t cx = cell_coord(x.x, g.ox, g.inv_cell, g.nx);
        int32_t cy = cell_coord(x.y, g.oy, g.inv_cell, g.ny);
        keys[i] = (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
        vals[i] = i;
        if (x0) x0[i] = x;
    }
}

// cell_start[c] = first sorted slot with key >= c, for c in [0, ncells]:
// one thread per cell, lower_bound over the sorted keys (balanced however
// the bodies are spread over the grid).
__global__ __launch_bounds__(kThreads) void k_cell_start(const uint32_t* __restrict__ keys, uint32_t n,
                                                         const GridParams* __restrict__ grid,
                                                         uint32_t* __restrict__ cell_start) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ncells = grid->ncells;
    if (c > ncells) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    cell_start[c] = lo;
}

// ============================================================ general path

// Gravity + springs -> acc (physics.rs:404-408, springs.rs:25-47).
__global__ __launch_bounds__(kThreads) void k_forces(DeviceWorld d, SubstepArgs a) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nb) return;
    uint32_t fl = d.b.bflags[b];
    if (fl & BF_STATIC) return;
    float2 acc = a.fresh ? d.b.acc[b] : make_float2(0.0f, 0.0f);
    float gm = d.b.gmod[b];
    acc.x = acc.x + a.gx * gm;
    acc.y = acc.y + a.gy * gm;
    const float m = d.b.mass[b];
    for (uint32_t k = d.bs_start[b]; k < d.bs_start[b + 1]; ++k) {
        uint32_t s = d.bs_list[k];
        int32_t ia = d.sp_a[s], ib = d.sp_b[s];
        float2 pa = d.b.pos[ia], pb = d.b.pos[ib];
        float2 va = d.b.vel[ia], vb = d.b.vel[ib];
        float dx = pb.x - pa.x, dy = pb.y - pa.y;
        float dist = vlen(dx, dy);
        float ux = dx / dist, uy = dy / dist;
        float rvx = va.x - vb.x, rvy = va.y - vb.y;
        float dot = rvx * ux + rvy * uy;
        float cd = d.sp_c[s] * dot;
        float dfx = cd * ux, dfy = cd * uy;
        float mag = d.sp_k[s] * (dist - d.sp_rest[s]);
        float fx = ux * (mag - dfx), fy = uy * (mag - dfy);
        if ((int32_t)b == ia) {
            acc.x = acc.x + fx / m;
            acc.y = acc.y + fy / m;
        } else {
            acc.x = acc.x + (-fx) / m;
            acc.y = acc.y + (-fy) / m;
        }
    }
    d.b.acc[b] = acc;
}

// Contacts (physics.rs:241-317), one thread per body over its colliders.
template <bool GROUPS, bool SENSORS>
__global__ __launch_bounds__(kThreads) void k_contacts(DeviceWorld d, SubstepArgs a) {
    LocalAccum la;
    la.init();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < d.nb) {
        const GridParams g = *d.grid;
        const float D = read_dmax(d.accum);
        float2 p = d.b.pos[b];
        const float mown = d.b.mass[b];
        ContactList<16, uint64_t> L;
        L.init();
        const uint32_t k0 = d.bc_start[b], k1 = d.bc_start[b + 1];
        // one enumeration of this body's contact corrections, in visit order
        auto enumerate = [&](auto&& emit) {
            for (uint32_t kk = k0; kk < k1; ++kk) {
                const uint32_t c = d.bc_list[kk];
                const float2 xa = d.c.abs[c];
                if (!finitef(xa.x) || !finitef(xa.y)) continue;
                const float ra = d.c.radius[c];
                uint2 ga = make_uint2(0, 0);
                if (GROUPS) ga = d.c.groups[c];
                const bool sen_a = SENSORS ? (d.c.cflags[c] & 1u) : false;
                const float Q = ra + a.rmax + D;
                scan_stencil(g, d.cell_start, d.nc, xa.x - a.drift_prev_x, xa.y - a.drift_prev_y, Q, [&](uint32_t t) {
                    const uint32_t j = d.vals[t];
                    if (j == c || j >= d.nc) return;
                    const int32_t pj = d.c.parent[j];
                    if (pj < 0 || pj == (int32_t)b) return;
                    if (GROUPS) {
                        uint2 gb = d.c.groups[j];
                        if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) return;
                    }
                    const float2 xb = d.c.abs[j];
                    const bool isA = c > j;
                    float ax = isA ? xa.x - xb.x : xb.x - xa.x;
                    float ay = isA ? xa.y - xb.y : xb.y - xa.y;
                    const float rs = ra + d.c.radius[j];
                    float dot = ax * ax + ay * ay;
                    if (dot > (rs * rs) * 1.00000095367431640625f) return;
                    float dist = sqrtf(dot);
                    if (!(dist < rs)) return;
                    const uint64_t key = isA ? (((uint64_t)c << 32) | j) : (((uint64_t)j << 32) | c);
                    emit(key, isA, j, pj, ax, ay, dist, rs, sen_a);
                });
            }
        };
        uint32_t count = 0, pushes = 0, flags = 0;
        // Correction vector of one contact for this body's side.  A push-out
        // (physics.rs:272-286) is folded into the same entry: it is order
        // dependent in the reference and reproduced here for the pair only.
        auto correction = [&](bool first, auto&& sink, uint64_t key, bool isA, uint32_t j, int32_t pj, float ax,
                              float ay, float dist, float rs, bool sen_a) {
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
            const bool sen_b = SENSORS ? (d.c.cflags[j] & 1u) : false;
            if (sen_a || sen_b) {
                if (pushed) sink(key, px, py);
                return;
            }
            const float mpart = d.b.mass[pj];
            const float mA = isA ? mown : mpart, mB = isA ? mpart : mown;
            const float ratio = 1.0f - mA / (mA + mB);
            const float nx = ax / dist, ny = ay / dist;
            if (first && (nx != nx || ny != ny)) flags |= ERR_NAN_NORMAL;
            const float delta = rs - dist;
            float cx, cy;
            if (isA) {
                const float k = ratio * delta;
                cx = k * nx;
                cy = k * ny;
            } else {
                const float k = (1.0f - ratio) * delta;
                cx = -(k * nx);
                cy = -(k * ny);
            }
            if (pushed) {
                cx = px + cx;
                cy = py + cy;
            }
            sink(key, cx, cy);
        };
        enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float ax, float ay, float dist, float rs,
                      bool sen_a) {
            if (isA) ++count;
            correction(true, [&](uint64_t k, float x, float y) { L.add(k, x, y); }, key, isA, j, pj, ax, ay, dist,
                       rs, sen_a);
        });
        if (L.n > 0) {
            uint64_t last = L.apply(p);
            uint32_t done = L.n < 16 ? L.n : 16;
            // more corrections than the register list holds: apply the rest
            // by repeated selection of the next key (same order)
            while (done < L.n) {
                uint64_t best = ~0ull;
                float bx = 0.0f, by = 0.0f;
                enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float ax, float ay, float dist,
                              float rs, bool sen_a) {
                    correction(
                        false,
                        [&](uint64_t k, float x, float y) {
                            if (k > last && k < best) {
                                best = k;
                                bx = x;
                                by = y;
                            }
                        },
                        key, isA, j, pj, ax, ay, dist, rs, sen_a);
                });
                if (best == ~0ull) break;
                p.x = p.x + bx;
                p.y = p.y + by;
                last = best;
                ++done;
            }
            d.b.pos[b] = p;
        }
        la.contacts = count;
        la.push_outs = pushes;
        la.flags = flags;
    }
    flush_accum(la, d.accum, d.block_contacts, 0, false);
}

// Fixed joints (physics.rs:424-477): Gauss-Seidel in arena order inside each
// component of the joint graph; components touch disjoint bodies, so running
// them concurrently reproduces the sequential sweep.
__global__ __launch_bounds__(64) void k_joints(DeviceWorld d, SubstepArgs a, uint32_t iterations) {
    uint32_t comp = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t flags = 0;
    if (comp < d.n_jcomp) {
        const uint32_t k0 = d.jc_start[comp], k1 = d.jc_start[comp + 1];
        for (uint32_t it = 0; it < iterations; ++it) {
            for (uint32_t kk = k0; kk < k1; ++kk) {
                const uint32_t j = d.jc_list[kk];
                const int32_t ia = d.j_a[j], ib = d.j_b[j];
                float2 pa = d.b.pos[ia], pb = d.b.pos[ib];
                const float2 an_a = d.j_anchor_a[j], an_b = d.j_anchor_b[j];
                float wax = pa.x + an_a.x, way = pa.y + an_a.y;
                float wbx = pb.x + an_b.x, wby = pb.y + an_b.y;
                float dx = wbx - wax, dy = wby - way;
                float dist = vlen(dx, dy);
                if (dist < 1e-6f) continue;
                float off_by = dist - d.j_dist[j];
                float cx = (off_by * dx) / dist, cy = (off_by * dy) / dist;
                float ma = d.b.mass[ia], mb = d.b.mass[ib];
                if (!(ma > 0.0f) || !(mb > 0.0f)) flags |= ERR_JOINT_MASS;
                float inv_a = 1.0f / ma, inv_b = 1.0f / mb;
                float ims = inv_a + inv_b;
                bool sa = d.b.bflags[ia] & BF_STATIC, sb = d.b.bflags[ib] & BF_STATIC;
                if (sa) {
                    pb.x = pb.x - ims * cx;
                    pb.y = pb.y - ims * cy;
                } else if (sb) {
                    pa.x = pa.x + ims * cx;
                    pa.y = pa.y + ims * cy;
                } else {
                    float ratio = inv_a / ims;
                    float rb = 1.0f - ratio;
                    pa.x = pa.x + ratio * cx;
                    pa.y = pa.y + ratio * cy;
                    pb.x = pb.x - rb * cx;
                    pb.y = pb.y - rb * cy;
                }
                d.b.pos[ia] = pa;
                d.b.pos[ib] = pb;
                float angle_a = atan2_f32(dy, dx);
                float angle_b = -atan2_f32(dy, -dx);
                float diff = (angle_b - angle_a) - d.j_target[j];
                float corr = diff * 0.5f;
                float ra = d.b.rot[ia] + corr * a.h;
                float rb2 = d.b.rot[ib] - corr * a.h;
                d.b.rot[ia] = ra;
                d.b.rot[ib] = rb2;
                if (!finitef(ra) || !finitef(rb2)) flags |= ERR_JOINT_ROT;
            }
        }
    }
    if (flags) atomicOr(&d.accum->error_flags, flags);
}

// Verlet + collider transforms + constraints (physics.rs:323-395).
__global__ __launch_bounds__(kThreads) void k_integrate(DeviceWorld d, SubstepArgs a) {
    LocalAccum la;
    la.init();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < d.nb) {
        const uint32_t fl = d.b.bflags[b];
        BodyState s;
        s.p = d.b.pos[b];
        s.rot = d.b.rot[b];
        s.omega = d.b.omega[b];
        if (fl & BF_STATIC) {  // :327-332
            d.b.pos_old[b] = s.p;
            d.b.vel[b] = make_float2(0.0f, 0.0f);
        } else {
            float2 acc;
            if (a.has_springs) {
                acc = d.b.acc[b];
            } else {
                acc = a.fresh ? d.b.acc[b] : make_float2(0.0f, 0.0f);
                float gm = d.b.gmod[b];
                acc.x = acc.x + a.gx * gm;
                acc.y = acc.y + a.gy * gm;
            }
            s.po = d.b.pos_old[b];
            const bool vr = fl & BF_HAS_VREQ;
            float2 vreq = vr ? d.b.vreq[b] : make_float2(0.0f, 0.0f);
            float tq = a.fresh ? d.b.torque[b] : 0.0f;
            float toi = (a.fresh || !(fl & BF_INERTIA_POS)) ? tq / d.b.inertia[b] : 0.0f;
            float ratio = a.ratio_rest;
            if ((int32_t)b == a.first_dyn) {
                ratio = a.h / d.accum->old_dt;
                d.accum->old_dt = a.h;
            }
            const float rot0 = s.rot, om0 = s.omega;
            verlet_dynamic(s, acc, toi, vr, vreq, a.h, ratio);
            d.b.pos_old[b] = s.po;
            d.b.vel[b] = s.vel;
            if (__float_as_uint(s.rot) != __float_as_uint(rot0)) d.b.rot[b] = s.rot;
            if (__float_as_uint(s.omega) != __float_as_uint(om0)) d.b.omega[b] = s.omega;
            if (vr) d.b.bflags[b] = fl & ~BF_HAS_VREQ;
            if (a.fresh) d.b.torque[b] = 0.0f;
        }
        // collider transforms from the pre-constraint position (:360-366)
        for (uint32_t kk = d.bc_start[b]; kk < d.bc_start[b + 1]; ++kk) {
            uint32_t c = d.bc_list[kk];
            float2 x = body_xform(s.rot, s.p, d.c.off[c]);
            d.c.abs[c] = x;
            float2 x0 = d.x0[c];
            la.point(x.x, x.y, x0.x, x0.y, a.drift_x, a.drift_y);
        }
        apply_constraints(s.p, d.con_pos, d.con_r, d.nk);
        d.b.pos[b] = s.p;
    }
    flush_accum(la, d.accum, d.block_contacts, a.parity, true);
}

// ============================================================ sphere path

template <bool GROUPS, bool SENSORS, bool ZERO_OFF>
__global__ __launch_bounds__(kThreads) void k_sphere_substep(DeviceWorld d, SubstepArgs a,
                                                             const float2* __restrict__ abs_in,
                                                             float2* __restrict__ abs_out) {
    LocalAccum la;
    la.init();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < d.nb) {
        const uint32_t fl = d.b.bflags[s];
        const uint32_t id = d.b.aid[s];
        BodyState st;
        st.p = d.b.pos[s];
        const float mown = d.b.mass[s];
        uint32_t count = 0, pushes = 0, flags = 0;
        if (a.collisions) {
            const float2 xa = abs_in[s];
            if (finitef(xa.x) && finitef(xa.y)) {
                const GridParams g = *d.grid;
                const float D = read_dmax(d.accum);
                const float ra = d.c.radius[s];
                uint2 ga = make_uint2(0, 0);
                if (GROUPS) ga = d.c.groups[s];
                const bool sen_a = SENSORS ? (fl & BF_SENSOR) : false;
                ContactList<8, uint32_t> L;
                L.init();
                auto enumerate = [&](auto&& sink) {
                    scan_stencil(g, d.cell_start, d.nb, xa.x - a.drift_prev_x, xa.y - a.drift_prev_y, ra + a.rmax + D,
                                 [&](uint32_t t) {
                                     if (t == s) return;
                                     if (GROUPS) {
                                         uint2 gb = d.c.groups[t];
                                         if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) return;
                                     }
                                     const float2 xb = abs_in[t];
                                     const uint32_t idb = d.b.aid[t];
                                     const bool isA = id > idb;
                                     float ax = isA ? xa.x - xb.x : xb.x - xa.x;
                                     float ay = isA ? xa.y - xb.y : xb.y - xa.y;
                                     const float rs = ra + d.c.radius[t];
                                     float dot = ax * ax + ay * ay;
                                     if (dot > (rs * rs) * 1.00000095367431640625f) return;
                                     float dist = sqrtf(dot);
                                     if (!(dist < rs)) return;
                                     sink(t, idb, isA, ax, ay, dist, rs);
                                 });
                };
                auto correction = [&](bool first, auto&& put, uint32_t t, uint32_t idb, bool isA, float ax,
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
                    const bool sen_b = SENSORS ? (d.b.bflags[t] & BF_SENSOR) : false;
                    if (sen_a || sen_b) {
                        if (pushed) put(idb, px, py);
                        return;
                    }
                    const float mpart = d.b.mass[t];
                    const float mA = isA ? mown : mpart, mB = isA ? mpart : mown;
                    const float ratio = 1.0f - mA / (mA + mB);
                    const float nx = ax / dist, ny = ay / dist;
                    if (first && (nx != nx || ny != ny)) flags |= ERR_NAN_NORMAL;
                    const float delta = rs - dist;
                    float cx, cy;
                    if (isA) {
                        const float k = ratio * delta;
                        cx = k * nx;
                        cy = k * ny;
                    } else {
                        const float k = (1.0f - ratio) * delta;
                        cx = -(k * nx);
                        cy = -(k * ny);
                    }
                    if (pushed) {
                        cx = px + cx;
                        cy = py + cy;
                    }
                    put(idb, cx, cy);
                };
                enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                    if (isA) ++count;
                    correction(true, [&](uint32_t k, float x, float y) { L.add(k, x, y); }, t, idb, isA, ax, ay,
                               dist, rs);
                });
                if (L.n > 0) {
                    uint32_t last = L.apply(st.p);
                    uint32_t done = L.n < 8 ? L.n : 8;
                    while (done < L.n) {
                        uint32_t best = 0xFFFFFFFFu;
                        float bx = 0.0f, by = 0.0f;
                        enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                            correction(
                                false,
                                [&](uint32_t k, float x, float y) {
                                    if (k > last && k < best) {
                                        best = k;
                                        bx = x;
                                        by = y;
                                    }
                                },
                                t, idb, isA, ax, ay, dist, rs);
                        });
                        if (best == 0xFFFFFFFFu) break;
                        st.p.x = st.p.x + bx;
                        st.p.y = st.p.y + by;
                        last = best;
                        ++done;
                    }
                }
            }
        }
        // Verlet (physics.rs:326-358)
        st.rot = d.b.rot[s];
        st.omega = d.b.omega[s];
        const float rot0 = st.rot, om0 = st.omega;
        if (fl & BF_STATIC) {
            d.b.pos_old[s] = st.p;
            d.b.vel[s] = make_float2(0.0f, 0.0f);
        } else {
            float2 acc = a.fresh ? d.b.acc[s] : make_float2(0.0f, 0.0f);
            const float gm = d.b.gmod[s];
            acc.x = acc.x + a.gx * gm;
            acc.y = acc.y + a.gy * gm;
            st.po = d.b.pos_old[s];
            const bool vr = fl & BF_HAS_VREQ;
            float2 vreq = vr ? d.b.vreq[s] : make_float2(0.0f, 0.0f);
            float tq = a.fresh ? d.b.torque[s] : 0.0f;
            float toi = (a.fresh || !(fl & BF_INERTIA_POS)) ? tq / d.b.inertia[s] : 0.0f;
            float ratio = a.ratio_rest;
            if ((int32_t)id == a.first_dyn) {
                ratio = a.h / d.accum->old_dt;
                d.accum->old_dt = a.h;
            }
            verlet_dynamic(st, acc, toi, vr, vreq, a.h, ratio);
            d.b.pos_old[s] = st.po;
            d.b.vel[s] = st.vel;
            if (__float_as_uint(st.rot) != __float_as_uint(rot0)) d.b.rot[s] = st.rot;
            if (__float_as_uint(st.omega) != __float_as_uint(om0)) d.b.omega[s] = st.omega;
            if (vr) d.b.bflags[s] = fl & ~BF_HAS_VREQ;
            if (a.fresh) d.b.torque[s] = 0.0f;
        }
        float2 x = ZERO_OFF ? body_xform_zero_off(st.rot, st.p) : body_xform(st.rot, st.p, d.c.off[s]);
        abs_out[s] = x;
        float2 x0 = d.x0[s];
        la.point(x.x, x.y, x0.x, x0.y, a.drift_x, a.drift_y);
        apply_constraints(st.p, d.con_pos, d.con_r, d.nk);
        d.b.pos[s] = st.p;
        la.contacts = count;
        la.push_outs = pushes;
        la.flags = flags;
    }
    flush_accum(la, d.accum, d.block_contacts, a.parity, true);
}

// Gather every per-body array into cell order: dst[s] = src[vals[s]].
__global__ __launch_bounds__(kThreads) void k_sphere_permute(DeviceWorld d, uint32_t carry_user) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.nb) return;
    const uint32_t o = d.vals[s];
    if (o >= d.nb) return;  // a permutation of [0, nb) by construction
    d.b_alt.pos[s] = d.b.pos[o];
    d.b_alt.pos_old[s] = d.b.pos_old[o];
    d.b_alt.vel[s] = d.b.vel[o];
    d.b_alt.rot[s] = d.b.rot[o];
    d.b_alt.omega[s] = d.b.omega[o];
    d.b_alt.inertia[s] = d.b.inertia[o];
    d.b_alt.mass[s] = d.b.mass[o];
    d.b_alt.gmod[s] = d.b.gmod[o];
    const uint32_t fl = d.b.bflags[o];
    d.b_alt.bflags[s] = fl;
    const uint32_t id = d.b.aid[o];
    d.b_alt.aid[s] = id;
    d.slot_of[id] = s;
    if (carry_user || (fl & BF_HAS_VREQ)) {
        d.b_alt.acc[s] = d.b.acc[o];
        d.b_alt.torque[s] = d.b.torque[o];
        d.b_alt.vreq[s] = d.b.vreq[o];
    }
    const float2 x = d.c.abs[o];
    d.c_alt.abs[s] = x;
    d.x0[s] = x;
    d.c_alt.off[s] = d.c.off[o];
    d.c_alt.radius[s] = d.c.radius[o];
    d.c_alt.groups[s] = d.c.groups[o];
    d.c_alt.cflags[s] = d.c.cflags[o];
}

}  // namespace

// ============================================================ launchers

hipError_t launch_rebuild_keys(DeviceWorld& d, float cell, uint32_t parity, bool sphere_path) {
    uint32_t n = d.nc;
    uint32_t blocks = div_up(n > 0 ? n : 1, kThreads);
    k_rebuild_keys<<<blocks, kThreads, 0, d.stream>>>(d.c.abs, n, d.accum, parity, cell, d.cells_cap, d.grid, d.keys,
                                                      d.vals, sphere_path ? nullptr : d.x0);
    return hipGetLastError();
}

hipError_t launch_rebuild_sort_and_cells(DeviceWorld& d) {
    hipError_t e = radix_sort_pairs(&d.keys, &d.vals, &d.keys_tmp, &d.vals_tmp, d.nc, d.key_bits, d.sort, d.stream);
    if (e != hipSuccess) return e;
    k_cell_start<<<div_up(d.cells_cap + 1, kThreads), kThreads, 0, d.stream>>>(d.keys, d.nc, d.grid, d.cell_start);
    return hipGetLastError();
}

hipError_t launch_sphere_permute(DeviceWorld& d, bool carry_user_inputs) {
    if (d.nb == 0) return hipSuccess;
    k_sphere_permute<<<div_up(d.nb, kThreads), kThreads, 0, d.stream>>>(d, carry_user_inputs ? 1u : 0u);
    hipError_t e = hipGetLastError();
    BodyArrays tb = d.b;
    d.b = d.b_alt;
    d.b_alt = tb;
    ColliderArrays tc = d.c;
    d.c = d.c_alt;
    d.c_alt = tc;
    return e;
}

hipError_t launch_general_substep(DeviceWorld& d, const SubstepArgs& a, uint32_t joint_iterations, bool groups,
                                  bool sensors) {
    if (d.nb == 0) return hipSuccess;
    const uint32_t blocks = div_up(d.nb, kThreads);
    if (a.has_springs) k_forces<<<blocks, kThreads, 0, d.stream>>>(d, a);
    if (a.collisions && d.nc > 0) {
        if (groups && sensors)
            k_contacts<true, true><<<blocks, kThreads, 0, d.stream>>>(d, a);
        else if (groups)
            k_contacts<true, false><<<blocks, kThreads, 0, d.stream>>>(d, a);
        else if (sensors)
            k_contacts<false, true><<<blocks, kThreads, 0, d.stream>>>(d, a);
        else
            k_contacts<false, false><<<blocks, kThreads, 0, d.stream>>>(d, a);
    }
    if (d.n_jcomp > 0 && joint_iterations > 0)
        k_joints<<<div_up(d.n_jcomp, 64), 64, 0, d.stream>>>(d, a, joint_iterations);
    k_integrate<<<blocks, kThreads, 0, d.stream>>>(d, a);
    return hipGetLastError();
}

hipError_t launch_sphere_substep(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors, bool zero_off,
                                 float2* abs_in, float2* abs_out) {
    if (d.nb == 0) return hipSuccess;
    const uint32_t blocks = div_up(d.nb, kThreads);
#define BG_SPHERE(G, S, Z) k_sphere_substep<G, S, Z><<<blocks, kThreads, 0, d.stream>>>(d, a, abs_in, abs_out)
    if (!groups && !sensors && zero_off)
        BG_SPHERE(false, false, true);
    else if (!groups && !sensors)
        BG_SPHERE(false, false, false);
    else if (groups && sensors)
        BG_SPHERE(true, true, false);
    else if (groups)
        BG_SPHERE(true, false, false);
    else
        BG_SPHERE(false, true, false);
#undef BG_SPHERE
    return hipGetLastError();
}

}  // namespace bg
