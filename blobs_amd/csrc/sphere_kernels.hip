// sphere_kernels.hip -- the fused substep for worlds whose every body owns
// exactly one collider and that have no springs or joints (the plain large
// scene): gravity, contacts, Verlet, collider transform and circle
// constraints of physics.rs:401-421 in ONE kernel per substep.
//
// Layout: bodies sit in broadphase-cell order (permuted at every rebuild),
// in packed 16-byte records; the collider snapshot the contact pass reads
// (physics.rs:264, written by the previous substep's update_objects at
// :360-366) is ping-ponged so a substep can read all snapshots while
// writing its own.
//
// Broadphase between rebuilds.  The sorted grid holds the snapshots of the
// rebuild (x0).  A body whose snapshot moves more than `skin` from x0 (after
// removing the common gravity drift) is an outlier for the rest of the
// period.  In the substep where a body is an outlier at the end, it
//   - inserts itself into a hashed cell table of current snapshots, and
//   - pushes its slot onto the candidate list of every body whose x0 cell is
//     within r + r_max + skin of it (the bodies that can touch it next).
// A regular body then visits: the sorted grid with radius r + r_max + skin
// (skipping outliers) plus its candidate list; an outlier visits the sorted
// grid plus the outlier table.  Every pair of the reference predicate is
// visited exactly once from each side, whatever the bodies do; the skin and
// rebuild interval change only the speed.
#include "common.hpp"
#include "device_state.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

struct SphereSubstep {
    SphereArrays s;
    const float4* __restrict__ snap_in;
    float4* __restrict__ snap_out;
    const float2* __restrict__ x0;
    const uint32_t* __restrict__ cell_start;
    const GridParams* __restrict__ grid;
    const uint32_t* __restrict__ oh_head;  // outlier table read this substep
    const uint32_t* __restrict__ oh_next;
    uint32_t* oh_ins_head;                 // table this substep fills (read by the next one)
    uint32_t* oh_ins_next;
    uint32_t* oh_clr_head;                 // table cleared for the substep after next
    uint32_t oh_mask;
    uint32_t* xl_count_in;                 // candidate lists read (and emptied) this substep
    const uint32_t* __restrict__ xl_list_in;
    uint32_t* xl_count_out;                // candidate lists filled for the next substep
    uint32_t* xl_list_out;
    uint32_t nb;
    StepAccum* accum;
    unsigned long long* block_contacts;
    const float2* __restrict__ con_pos;
    const float* __restrict__ con_r;
    uint32_t nk;
};

__device__ __forceinline__ int32_t free_cell(float x, float o, float inv) {
    float f = floorf((x - o) * inv);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int32_t)f;
}

__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, uint32_t mask) {
    return (((uint32_t)cx * 0x9E3779B1u) ^ ((uint32_t)cy * 0x85EBCA77u)) & mask;
}

// Rows of the sorted grid overlapping the square of half-width Q around
// (qx, qy): calls row(t0, t1) with the slot range of each row.  Cell
// coordinates are monotone in x, so the rows cover every x0 within Q.
template <class F>
__device__ __forceinline__ void grid_rows(const GridParams& g, const uint32_t* __restrict__ cell_start, uint32_t n,
                                          float qx, float qy, float Q, F&& row) {
    const float R = Q + (1e-5f * (fabsf(qx) + fabsf(qy) + Q) + 1e-6f);
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
        const uint32_t rowc = (uint32_t)cy * (uint32_t)g.nx;
        const uint32_t t0 = cell_start[rowc + (uint32_t)cx0];
        uint32_t t1 = cell_start[rowc + (uint32_t)cx1 + 1u];
        t1 = t1 < n ? t1 : n;
        row(t0, t1);
    }
}

template <bool GROUPS, bool SENSORS, bool ZERO_OFF, bool CONSTRAINTS>
__global__ __launch_bounds__(kThreads) void k_sphere_substep(SphereSubstep k, SubstepArgs a) {
    LocalAccum la;
    la.init();
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kThreads + threadIdx.x;
    // clear the outlier table of the substep after next
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i <= k.oh_mask; i += gridDim.x * kThreads)
        k.oh_clr_head[i] = kNil;
    uint32_t new_outliers = 0;
    if (s < k.nb) {
        const uint32_t fl = k.s.bflags[s];
        const float4 pp = k.s.pp[s];
        BodyState st;
        st.p = make_float2(pp.x, pp.y);
        st.po = make_float2(pp.z, pp.w);
        const float4 own = k.snap_in[s];
        const uint32_t wa = __float_as_uint(own.w);
        const uint32_t id = wa & ~kSnapOutlier;
        const bool was_out = (wa & kSnapOutlier) != 0u;
        const float2 mgv = k.s.mg[s];
        const float mown = mgv.x;
        const uint32_t xl_n = k.xl_count_in[s];
        if (xl_n) k.xl_count_in[s] = 0u;
        uint32_t count = 0, pushes = 0, flags = 0;
        if (a.collisions && finitef(own.x) && finitef(own.y)) {
            const float2 xa = make_float2(own.x, own.y);
            const float ra = own.z;
            const GridParams g = *k.grid;
            uint2 ga = make_uint2(0, 0);
            if (GROUPS) ga = k.s.groups[s];
            const bool sen_a = SENSORS ? (fl & BF_SENSOR) != 0u : false;
            // reference predicate (physics.rs:255-269) against the slot-t snapshot
            auto test = [&](uint32_t t, const float4 sb, auto&& sink) {
                if (GROUPS) {
                    const uint2 gb = k.s.groups[t];
                    if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) return;
                }
                const uint32_t idb = __float_as_uint(sb.w) & ~kSnapOutlier;
                const bool isA = id > idb;
                const float ax = isA ? xa.x - sb.x : sb.x - xa.x;
                const float ay = isA ? xa.y - sb.y : sb.y - xa.y;
                const float rs = ra + sb.z;
                const float dot = ax * ax + ay * ay;
                if (dot > (rs * rs) * 1.00000095367431640625f) return;
                const float dist = sqrtf(dot);
                if (!(dist < rs)) return;
                sink(t, idb, isA, ax, ay, dist, rs);
            };
            auto enumerate = [&](auto&& sink) {
                // sorted grid: four candidates per load batch (clamped indices, no branches)
                grid_rows(g, k.cell_start, k.nb, xa.x - a.drift_prev_x, xa.y - a.drift_prev_y, ra + a.rmax + a.skin,
                          [&](uint32_t t0, uint32_t t1) {
                              for (uint32_t t = t0; t < t1; t += 4) {
                                  float4 sb[4];
#pragma unroll
                                  for (int u = 0; u < 4; ++u) sb[u] = k.snap_in[min(t + u, t1 - 1)];
#pragma unroll
                                  for (int u = 0; u < 4; ++u) {
                                      const uint32_t tu = t + u;
                                      if (tu < t1 && tu != s && !(__float_as_uint(sb[u].w) & kSnapOutlier))
                                          test(tu, sb[u], sink);
                                  }
                              }
                          });
                if (!was_out && xl_n <= kExtra) {
                    // outliers that pushed themselves onto this body's list
                    for (uint32_t i = 0; i < xl_n; ++i) {
                        const uint32_t t = k.xl_list_in[s * kExtra + i];
                        test(t, k.snap_in[t], sink);
                    }
                    return;
                }
                // outlier body, or a list that overflowed: the outlier table
                float R = ra + a.rmax;
                R = R + (1e-5f * (fabsf(xa.x) + fabsf(xa.y) + R) + 1e-6f);
                const int32_t cx0 = free_cell(xa.x - R, g.ox, g.inv_cell), cx1 = free_cell(xa.x + R, g.ox, g.inv_cell);
                const int32_t cy0 = free_cell(xa.y - R, g.oy, g.inv_cell), cy1 = free_cell(xa.y + R, g.oy, g.inv_cell);
                if (!(R < 1e30f) || cx1 - cx0 > 64 || cy1 - cy0 > 64) {
                    for (uint32_t h = 0; h <= k.oh_mask; ++h)  // huge radius: every bucket (superset)
                        for (uint32_t t = k.oh_head[h]; t != kNil; t = k.oh_next[t])
                            if (t != s) test(t, k.snap_in[t], sink);
                    return;
                }
                for (int32_t cy = cy0; cy <= cy1; ++cy)
                    for (int32_t cx = cx0; cx <= cx1; ++cx)
                        for (uint32_t t = k.oh_head[cell_hash(cx, cy, k.oh_mask)]; t != kNil; t = k.oh_next[t]) {
                            if (t == s) continue;
                            const float4 sb = k.snap_in[t];
                            // each outlier lives in exactly one cell: drop hash collisions
                            if (free_cell(sb.x, g.ox, g.inv_cell) != cx || free_cell(sb.y, g.oy, g.inv_cell) != cy)
                                continue;
                            test(t, sb, sink);
                        }
            };
            // correction of one contact for this body's side (physics.rs:272-300);
            // a push-out is folded into the same entry (order-dependent in the
            // reference, reproduced for the pair)
            auto correction = [&](bool first, auto&& put, uint32_t t, uint32_t idb, bool isA, float ax, float ay,
                                  float dist, float rs) {
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
                const bool sen_b = SENSORS ? (k.s.bflags[t] & BF_SENSOR) != 0u : false;
                if (sen_a || sen_b) {
                    if (pushed) put(idb, px, py);
                    return;
                }
                const float mpart = k.s.mg[t].x;
                const float mA = isA ? mown : mpart, mB = isA ? mpart : mown;
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
            ContactList<8, uint32_t> L;
            L.init();
            enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                if (isA) ++count;
                correction(true, [&](uint32_t kk, float x, float y) { L.add(kk, x, y); }, t, idb, isA, ax, ay, dist,
                           rs);
            });
            if (L.n > 0) {
                uint32_t last = L.apply(st.p);
                uint32_t done = L.n < 8 ? L.n : 8;
                while (done < L.n) {  // more than 8: next key by selection, same order
                    uint32_t best = kNil;
                    float bx = 0.0f, by = 0.0f;
                    enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs) {
                        correction(
                            false,
                            [&](uint32_t kk, float x, float y) {
                                if (kk > last && kk < best) {
                                    best = kk;
                                    bx = x;
                                    by = y;
                                }
                            },
                            t, idb, isA, ax, ay, dist, rs);
                    });
                    if (best == kNil) break;
                    st.p.x = st.p.x + bx;
                    st.p.y = st.p.y + by;
                    last = best;
                    ++done;
                }
            }
        }
        // update_objects (physics.rs:326-358)
        const float2 rw = k.s.rw[s];
        st.rot = rw.x;
        st.omega = rw.y;
        float2 vel;
        if (fl & BF_STATIC) {
            st.po = st.p;
            vel = make_float2(0.0f, 0.0f);
        } else {
            float2 acc = a.fresh ? k.s.acc[s] : make_float2(0.0f, 0.0f);
            acc.x = acc.x + a.gx * mgv.y;
            acc.y = acc.y + a.gy * mgv.y;
            const bool vr = (fl & BF_HAS_VREQ) != 0u;
            const float2 vreq = vr ? k.s.vreq[s] : make_float2(0.0f, 0.0f);
            const float tq = a.fresh ? k.s.torque[s] : 0.0f;
            const float toi = (a.fresh || !(fl & BF_INERTIA_POS)) ? tq / k.s.inertia[s] : 0.0f;
            float ratio = a.ratio_rest;
            if ((int32_t)id == a.first_dyn) {
                ratio = a.h / k.accum->old_dt;
                k.accum->old_dt = a.h;
            }
            verlet_dynamic(st, acc, toi, vr, vreq, a.h, ratio);
            vel = st.vel;
            if (__float_as_uint(st.rot) != __float_as_uint(rw.x) || __float_as_uint(st.omega) != __float_as_uint(rw.y))
                k.s.rw[s] = make_float2(st.rot, st.omega);
            if (vr) k.s.bflags[s] = fl & ~BF_HAS_VREQ;
            if (a.fresh) k.s.torque[s] = 0.0f;
        }
        k.s.vel[s] = vel;
        // collider transform from the pre-constraint position (physics.rs:360-366)
        const float2 x = ZERO_OFF ? body_xform_zero_off(st.rot, st.p) : body_xform_slow(st.rot, st.p, k.s.off[s]);
        const bool finite_x = finitef(x.x) && finitef(x.y);
        bool out = was_out;
        if (!out && finite_x) {
            const float2 x0 = k.x0[s];
            const float ex = (x.x - x0.x) - a.drift_x, ey = (x.y - x0.y) - a.drift_y;
            out = !(sqrtf(ex * ex + ey * ey) <= a.skin);
            new_outliers += out ? 1u : 0u;
        }
        k.snap_out[s] = make_float4(x.x, x.y, own.z, __uint_as_float(id | (out ? kSnapOutlier : 0u)));
        if (out && finite_x && a.collisions) {
            const GridParams g = *k.grid;
            // outlier table of the next substep
            const uint32_t h = cell_hash(free_cell(x.x, g.ox, g.inv_cell), free_cell(x.y, g.oy, g.inv_cell), k.oh_mask);
            const uint32_t old = atomicExch(&k.oh_ins_head[h], s);
            k.oh_ins_next[s] = old;
            // every body whose rebuild cell is within reach gets this slot on its list
            grid_rows(g, k.cell_start, k.nb, x.x - a.drift_x, x.y - a.drift_y, own.z + a.rmax + a.skin,
                      [&](uint32_t t0, uint32_t t1) {
                          for (uint32_t t = t0; t < t1; ++t) {
                              if (t == s) continue;
                              const uint32_t c = atomicAdd(&k.xl_count_out[t], 1u);
                              if (c < kExtra) k.xl_list_out[t * kExtra + c] = s;
                          }
                      });
        }
        la.point(x.x, x.y, x.x, x.y, 0.0f, 0.0f);
        if (CONSTRAINTS) apply_constraints(st.p, k.con_pos, k.con_r, k.nk);
        k.s.pp[s] = make_float4(st.p.x, st.p.y, st.po.x, st.po.y);
        la.contacts = count;
        la.push_outs = pushes;
        la.flags = flags;
    }
    // outlier count: one atomic per block that produced any
    const uint32_t no = wave_sum_u32(new_outliers);
    __shared__ uint32_t s_no[kThreads / BG_WAVE];
    if ((threadIdx.x & (BG_WAVE - 1)) == 0) s_no[threadIdx.x / BG_WAVE] = no;
    flush_accum(la, k.accum, k.block_contacts, a.parity, true);
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kThreads / BG_WAVE; ++w) tot += s_no[w];
        if (tot) atomicAdd(&k.accum->n_outliers, tot);
    }
}

struct SpherePermute {
    SphereArrays src, dst;
    const uint32_t* __restrict__ vals;
    float2* __restrict__ x0;
    uint32_t* heads[3];
    uint32_t oh_size;
    uint32_t* xl_count[2];
    uint32_t nb;
    uint32_t carry_user;
};

// Gather every per-body record into cell order: dst[s] = src[vals[s]]; the
// new snapshot is also the new x0, no body is an outlier, the outlier tables
// and candidate lists are empty.
__global__ __launch_bounds__(kThreads) void k_sphere_permute(SpherePermute k) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = s; i < 3u * k.oh_size; i += gridDim.x * blockDim.x) k.heads[i / k.oh_size][i % k.oh_size] = kNil;
    if (s >= k.nb) return;
    k.xl_count[0][s] = 0u;
    k.xl_count[1][s] = 0u;
    const uint32_t o = k.vals[s];
    if (o >= k.nb) return;  // a permutation of [0, nb) by construction
    k.dst.pp[s] = k.src.pp[o];
    k.dst.vel[s] = k.src.vel[o];
    k.dst.rw[s] = k.src.rw[o];
    k.dst.mg[s] = k.src.mg[o];
    const uint32_t fl = k.src.bflags[o];
    k.dst.bflags[s] = fl;
    k.dst.off[s] = k.src.off[o];
    k.dst.groups[s] = k.src.groups[o];
    k.dst.inertia[s] = k.src.inertia[o];
    if (k.carry_user) {
        k.dst.acc[s] = k.src.acc[o];
        k.dst.torque[s] = k.src.torque[o];
    }
    if (k.carry_user || (fl & BF_HAS_VREQ)) k.dst.vreq[s] = k.src.vreq[o];
    float4 sn = k.src.snap[o];
    sn.w = __uint_as_float(__float_as_uint(sn.w) & ~kSnapOutlier);
    k.dst.snap[s] = sn;
    k.x0[s] = make_float2(sn.x, sn.y);
}

}  // namespace

hipError_t launch_sphere_permute(DeviceWorld& d, bool carry_user_inputs) {
    if (d.nb == 0) return hipSuccess;
    SpherePermute k;
    k.src = d.s;
    k.dst = d.s_alt;
    k.vals = d.vals;
    k.x0 = d.x0;
    for (int i = 0; i < 3; ++i) k.heads[i] = d.oh_head[i];
    k.oh_size = d.oh_size;
    k.xl_count[0] = d.xl_count[0];
    k.xl_count[1] = d.xl_count[1];
    k.nb = d.nb;
    k.carry_user = carry_user_inputs ? 1u : 0u;
    k_sphere_permute<<<div_up(d.nb, kThreads), kThreads, 0, d.stream>>>(k);
    hipError_t e = hipGetLastError();
    SphereArrays t = d.s;
    d.s = d.s_alt;
    d.s_alt = t;
    return e;
}

hipError_t launch_sphere_substep(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors, bool zero_off,
                                 bool constraints) {
    if (d.nb == 0) return hipSuccess;
    SphereSubstep k;
    k.s = d.s;
    k.snap_in = d.s.snap;
    k.snap_out = d.snap_next;
    k.x0 = d.x0;
    k.cell_start = d.cell_start;
    k.grid = d.grid;
    const uint32_t j = a.sub;
    k.oh_head = d.oh_head[j % 3];
    k.oh_next = d.oh_next[j % 3];
    k.oh_ins_head = d.oh_head[(j + 1) % 3];
    k.oh_ins_next = d.oh_next[(j + 1) % 3];
    k.oh_clr_head = d.oh_head[(j + 2) % 3];
    k.oh_mask = d.oh_size - 1;
    k.xl_count_in = d.xl_count[j % 2];
    k.xl_list_in = d.xl_list[j % 2];
    k.xl_count_out = d.xl_count[(j + 1) % 2];
    k.xl_list_out = d.xl_list[(j + 1) % 2];
    k.nb = d.nb;
    k.accum = d.accum;
    k.block_contacts = d.block_contacts;
    k.con_pos = d.con_pos;
    k.con_r = d.con_r;
    k.nk = d.nk;
    const uint32_t blocks = div_up(d.nb, kThreads);
#define BG_SPHERE(G, S, Z, C) k_sphere_substep<G, S, Z, C><<<blocks, kThreads, 0, d.stream>>>(k, a)
    if (!groups && !sensors && zero_off && !constraints)
        BG_SPHERE(false, false, true, false);
    else if (!groups && !sensors && !constraints)
        BG_SPHERE(false, false, false, false);
    else if (!groups && !sensors)
        BG_SPHERE(false, false, false, true);
    else
        BG_SPHERE(true, true, false, true);
#undef BG_SPHERE
    hipError_t e = hipGetLastError();
    float4* t = d.s.snap;
    d.s.snap = d.snap_next;
    d.snap_next = t;
    return e;
}

}  // namespace bg
