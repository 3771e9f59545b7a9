// step_common.hpp -- device helpers shared by the general and the fused
// sphere substep kernels (compiled with -ffp-contract=off, like every kernel
// source here: the reference never fuses a*b+c).
#pragma once

#include <float.h>

#include "common.hpp"
#include "device_state.hpp"

namespace bg {

constexpr int kThreads = 256;

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Renumbering them so that every XCD gets one contiguous
// range of slots keeps a body's grid neighbours (same and adjacent cell
// rows) in the L2 of the XCD that reads them.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    const uint32_t q = n / 8u, r = n % 8u;
    const uint32_t x = b % 8u, i = b / 8u;
    return x * q + (x < r ? x : r) + i;
}

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
// around (qx, qy), in slot order, four at a time: visit(t, r, nv) gets slots
// t .. t+nv-1 (nv in 1..4) with their records r[0..nv-1] (loaded together,
// so their loads overlap).  Cell coordinates are monotone in x, so the
// visited set is a superset of all snapshots within distance Q of the query
// point.
template <class F>
__device__ __forceinline__ void scan_stencil(const GridParams& g, const uint32_t* __restrict__ cell_start, uint32_t n,
                                             const float4* __restrict__ rec, float qx, float qy, float Q, F&& visit) {
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
        for (uint32_t t = t0; t < t1; t += 4) {
            float4 r[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) r[u] = rec[min(t + u, t1 - 1u)];
            visit(t, r, t1 - t < 4u ? t1 - t : 4u);
        }
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
// to one shared counter.  Blocks of 64 to kThreads threads.
static __device__ void flush_accum(const LocalAccum& l, StepAccum* g, unsigned long long* block_contacts, uint32_t parity,
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
        for (int w = 1; w < (int)(blockDim.x / BG_WAVE); ++w) {
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

// Position part of one fixed joint (physics.rs:434-460) on body positions pa,
// pb; the axis d = wb - wa goes to (dx, dy).  Returns false for coincident
// anchors, which the reference skips.
__device__ __forceinline__ bool joint_positions(float2& pa, float2& pb, float2 an_a, float2 an_b, float jdist,
                                                float4 jm, bool sa, bool sb, float& dx, float& dy,
                                                uint32_t& flags) {
    float wax = pa.x + an_a.x, way = pa.y + an_a.y;
    float wbx = pb.x + an_b.x, wby = pb.y + an_b.y;
    dx = wbx - wax;
    dy = wby - way;
    float dist = vlen(dx, dy);
    if (dist < 1e-6f) return false;
    float off_by = dist - jdist;
    float cx = (off_by * dx) / dist, cy = (off_by * dy) / dist;
    if (jm.w == 0.0f) flags |= ERR_JOINT_MASS;
    const float ims = jm.x;
    if (sa) {
        pb.x = pb.x - ims * cx;
        pb.y = pb.y - ims * cy;
    } else if (sb) {
        pa.x = pa.x + ims * cx;
        pa.y = pa.y + ims * cy;
    } else {
        const float ratio = jm.y, rbm = jm.z;
        pa.x = pa.x + ratio * cx;
        pa.y = pa.y + ratio * cy;
        pb.x = pb.x - rbm * cx;
        pb.y = pb.y - rbm * cy;
    }
    return true;
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
        if (fx <= 65536.0f && fy <= 65536.0f && (uint64_t)fx * (uint64_t)fy <= (uint64_t)cells_cap) {
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


}  // namespace bg
