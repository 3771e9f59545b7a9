// step_kernels.hip -- gfx950 kernels of one `Physics::step` substep.
//
// Reference stage order per substep (physics.rs:401-421): gravity, springs,
// contacts, fixed joints, Verlet update (+ collider transforms), circle
// constraints.  Two device paths implement it:
//
//  * general path (any world), per substep: k_keys_forces (the rebuild's
//    key pass beside gravity + springs per body), the scan and k_cell_scatter
//    (counting sort of the colliders by cell), k_contact_lists (per collider)
//    and k_contact_merge (per body; k_contacts for the bodies whose lists
//    overflowed), the ordered push-out pass (ordered_contacts.hip), k_joints
//    (one lane per independent joint component), k_integrate (Verlet, joint
//    rotations, collider transforms, constraints, per body);
//  * fused sphere path (one collider per body, no springs/joints): all stages
//    in one kernel per substep over bodies kept in broadphase-cell order
//    (sphere_kernels.hip).
//
// Contacts are order-independent until a push-out because the reference
// tests the collider snapshot written by the previous Verlet (physics.rs:264,
// 360-366), so the pair set and every correction vector depend only on the
// snapshot; only the summation order on a body position matters, and each
// body sums its own corrections in the reference's (i, j<i) lexicographic
// order.
#include <float.h>

#include "common.hpp"
#include "device_state.hpp"
#include "sort.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

constexpr uint32_t kContactThreads = 64;  // block size of k_contacts (block_contacts: one slot per 64 bodies)


// Cell key of collider i's snapshot on the grid fitted to the bounding box
// accumulated since the previous rebuild, and its rank among its cell's
// colliders (counting sort; order inside a cell is arbitrary, no result
// depends on it); x0 keeps the snapshot.  Thread 0 stores the grid and
// clears the box the coming substep accumulates.
__device__ __forceinline__ void rebuild_key_at(uint32_t i, const KeysArgs& k) {
    const GridParams g = grid_from_bbox(k.accum->bbox[k.parity], k.cell, k.cells_cap);
    if (i == 0) {
        *k.grid_out = g;
        k.accum->bbox[k.parity ^ 1][0] = 0xFFFFFFFFu;
        k.accum->bbox[k.parity ^ 1][1] = 0xFFFFFFFFu;
        k.accum->bbox[k.parity ^ 1][2] = 0u;
        k.accum->bbox[k.parity ^ 1][3] = 0u;
        k.accum->dmax = 0u;
        k.accum->n_outliers = 0u;
    }
    if (i < k.nc) {
        const float2 x = k.xy[i];
        int32_t cx = cell_coord(x.x, g.ox, g.inv_cell, g.nx);
        int32_t cy = cell_coord(x.y, g.oy, g.inv_cell, g.ny);
        const uint32_t kk = (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
        k.keys[i] = kk;
        k.rank[i] = atomicAdd(&k.cnt[kk], 1u);
        k.x0[i] = x;
    }
}

// Colliders into cell order: slot = cell_start[key] + rank (cell_start is the
// exclusive scan of the counts); the counts are cleared for the next rebuild.
// Beside the collider index, each slot gets the record the contact pass tests
// (position, radius, parent body), so a candidate is one load from a cell run.
__global__ __launch_bounds__(kThreads) void k_cell_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ rank, uint32_t nc,
                                                           const uint32_t* __restrict__ cell_start,
                                                           uint32_t* __restrict__ vals, uint32_t* cnt,
                                                           uint32_t n_cnt, const float2* __restrict__ abs,
                                                           const float* __restrict__ radius,
                                                           const int32_t* __restrict__ parent,
                                                           float4* __restrict__ rec,
                                                           const GridParams* __restrict__ grid) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // only the grid's cells were counted (and scanned): clear those
    const uint32_t n_used = grid->ncells + 1u < n_cnt ? grid->ncells + 1u : n_cnt;
    for (uint32_t c = i; c < n_used; c += gridDim.x * blockDim.x) cnt[c] = 0u;
    if (i < nc) {
        const uint32_t slot = cell_start[keys[i]] + rank[i];
        if (slot < nc) {  // consistent counts and scan never fail this
            vals[slot] = i;
            const float2 x = abs[i];
            rec[slot] = make_float4(x.x, x.y, radius[i], __int_as_float(parent[i]));
        }
    }
}

// ============================================================ general path

// Gravity + springs -> acc of body b (physics.rs:404-408, springs.rs:25-47).
__device__ __forceinline__ void forces_at(uint32_t b, const DeviceWorld& d, const SubstepArgs& a) {
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

// The substep's first launch on the general path: the collider keys of the
// rebuild (blocks [0, key_blocks), when the substep rebuilds) and gravity +
// springs (the rest).  The two are independent: keys read the collider
// snapshots, forces the body positions and velocities.
__global__ __launch_bounds__(kThreads) void k_keys_forces(KeysArgs k, uint32_t key_blocks, DeviceWorld d,
                                                         SubstepArgs a) {
    if (blockIdx.x < key_blocks) {
        rebuild_key_at(blockIdx.x * blockDim.x + threadIdx.x, k);
    } else if (a.has_springs) {
        forces_at((blockIdx.x - key_blocks) * blockDim.x + threadIdx.x, d, a);
    }
}

constexpr uint32_t kColStage = 8;  // colliders per body staged in LDS by k_contacts

// Contacts (physics.rs:241-317), one thread per body over its colliders, in
// blocks of kContactThreads.  With `list`, only the bodies listed there (count
// in cfb_n[0], grid-stride) and without counting or events: the bodies whose
// collider lists overflowed in k_contact_lists, which counted their contacts.
template <bool GROUPS, bool SENSORS>
__global__ __launch_bounds__(kContactThreads) void k_contacts(DeviceWorld d, SubstepArgs a, const uint32_t* list) {
    // the body's first kColStage colliders (x, y, radius, index), loaded together up front
    // so that the walk over its colliders does not wait on two dependent loads per collider
    __shared__ float4 s_col[kColStage][kContactThreads];
    LocalAccum la;
    la.init();
    const uint32_t lane = threadIdx.x;
    const uint32_t n_items = list ? d.cfb_n[0] : d.nb;
    const uint32_t stride = list ? gridDim.x * blockDim.x : n_items;  // without a list: one body per thread
    for (uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items; item += stride) {
        const uint32_t b = list ? list[item] : item;
        const GridParams g = *d.grid;
        const float D = read_dmax(d.accum);
        float2 p = d.b.pos[b];
        d.pos_pre[b] = p;  // the ordered pass re-sums from here if a push-out happens
        const float mown = d.b.mass[b];
        ContactList<16, uint64_t> L;
        L.init();
        const uint32_t k0 = d.bc_start[b], k1 = d.bc_start[b + 1];
        if (k1 > k0) {
            // an index past the end repeats the last collider (same slot, same values)
            uint32_t cc[kColStage];
#pragma unroll
            for (uint32_t u = 0; u < kColStage; ++u) cc[u] = d.bc_list[min(k0 + u, k1 - 1u)];
#pragma unroll
            for (uint32_t u = 0; u < kColStage; ++u) {
                const float2 x = d.c.abs[cc[u]];
                s_col[min(u, k1 - k0 - 1u)][lane] = make_float4(x.x, x.y, d.c.radius[cc[u]], __uint_as_float(cc[u]));
            }
        }
        // one enumeration of this body's contact corrections, in visit order
        auto enumerate = [&](auto&& emit) {
            for (uint32_t kk = k0; kk < k1; ++kk) {
                uint32_t c;
                float2 xa;
                float ra;
                if (kk - k0 < kColStage) {
                    const float4 sc = s_col[kk - k0][lane];
                    c = __float_as_uint(sc.w);
                    xa = make_float2(sc.x, sc.y);
                    ra = sc.z;
                } else {
                    c = d.bc_list[kk];
                    xa = d.c.abs[c];
                    ra = d.c.radius[c];
                }
                if (!finitef(xa.x) || !finitef(xa.y)) continue;
                uint2 ga = make_uint2(0, 0);
                if (GROUPS) ga = d.c.groups[c];
                const bool sen_a = SENSORS ? (d.c.cflags[c] & 1u) : false;
                const float Q = ra + a.rmax + D;
                scan_stencil(g, d.cell_start, d.nc, d.crec, xa.x - a.drift_prev_x, xa.y - a.drift_prev_y, Q,
                             [&](uint32_t t, const float4 (&rb)[4], uint32_t nv) {
                    // slot records (the sort ran on this substep's positions); the collider
                    // itself and its body's other colliders share the parent b.  All four
                    // are tested first, then the hits' collider indices and partner masses
                    // are loaded together, then the hits are taken in slot order.
                    // axis = A - B with A the higher collider index; b - a == -(a - b) up to
                    // the sign of a zero, so the unoriented difference decides and the axis
                    // is formed for a hit only
                    bool hit[4];
                    float dx[4], dy[4], dot[4], rs[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        const int32_t pj = __float_as_int(rb[u].w);
                        dx[u] = xa.x - rb[u].x;
                        dy[u] = xa.y - rb[u].y;
                        rs[u] = ra + rb[u].z;
                        dot[u] = dx[u] * dx[u] + dy[u] * dy[u];
                        hit[u] = u < nv && pj >= 0 && pj != (int32_t)b &&
                                 !(dot[u] > (rs[u] * rs[u]) * 1.00000095367431640625f) && sqrt_lt(dot[u], rs[u]);
                    }
                    uint32_t jj[4];
                    float mp[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        jj[u] = 0u;
                        mp[u] = 0.0f;
                        if (hit[u]) {
                            jj[u] = d.vals[t + u];
                            mp[u] = d.b.mass[__float_as_int(rb[u].w)];
                        }
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        if (!hit[u]) continue;
                        const uint32_t j = jj[u];
                        if (GROUPS) {
                            uint2 gb = d.c.groups[j];
                            if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) continue;
                        }
                        const bool isA = c > j;
                        const float ax = isA ? dx[u] : rb[u].x - xa.x, ay = isA ? dy[u] : rb[u].y - xa.y;
                        const uint64_t key = isA ? (((uint64_t)c << 32) | j) : (((uint64_t)j << 32) | c);
                        emit(key, isA, j, __float_as_int(rb[u].w), ax, ay, sqrtf(dot[u]), rs[u], sen_a, mp[u]);
                    }
                });
            }
        };
        uint32_t count = 0, flags = 0;
        // Correction vector of one contact for this body's side.  A pair whose
        // snapshots coincide takes the push-out branch (physics.rs:272-286), which
        // makes the rest of the pass order dependent: its row (the A collider) is
        // recorded once for k_contacts_ordered, which replays the affected pairs in
        // reference order and rewrites the positions they touch.  The value summed
        // here for such a pair is a placeholder.
        auto correction = [&](bool first, auto&& sink, uint64_t key, bool isA, uint32_t j, float mpart, float ax,
                              float ay, float dist, float rs, bool sen_a) {
            float px = 0.0f, py = 0.0f;
            const bool pushed = dist < 1e-6f;
            if (pushed) {
                ax = ax + 0.02f;
                dist = vlen(ax, ay);
                px = isA ? 0.01f : -0.01f;
                py = isA ? 0.0f : -0.0f;
                if (first) {
                    flags |= INFO_PUSH_OUT;
                    if (isA) {
                        const uint32_t row = (uint32_t)(key >> 32), bit = 1u << (row & 31u);
                        if (!(atomicOr(&d.ord_mark[row >> 5], bit) & bit)) d.ord_rows[atomicAdd(d.ord_n, 1u)] = row;
                    }
                }
            }
            const bool sen_b = SENSORS ? (d.c.cflags[j] & 1u) : false;
            if (sen_a || sen_b) {
                if (pushed) sink(key, px, py);
                return;
            }
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
                      bool sen_a, float mp) {
            if (isA && !list) ++count;
            correction(true, [&](uint64_t k, float x, float y) { L.add(k, x, y); }, key, isA, j, mp, ax, ay, dist,
                       rs, sen_a);
        });
        if (d.events && count > 0) {
            // collision events of the pairs whose A collider is this body's (physics.rs:288-311);
            // velocities are the bodies' calculated_velocity before the contact pass
            unsigned long long e = atomicAdd(&d.accum->event_count, (unsigned long long)count);
            const float2 va = d.b.vel[b];
            enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float ax, float ay, float dist, float rs,
                          bool sen_a, float) {
                if (!isA) return;
                if (e < d.event_cap) {
                    DevEvent ev;
                    ev.sub = a.sub_global;
                    ev.a = (uint32_t)(key >> 32);
                    ev.b = (uint32_t)key;
                    ev.pad = 0;
                    ev.va = va;
                    ev.vb = d.b.vel[pj];
                    d.events[e] = ev;
                }
                ++e;
            });
        }
        if (L.n > 0) {
            const uint32_t total = L.n;
            uint64_t last = L.apply(p);
            uint32_t done = total < 16 ? total : 16;
            // more corrections than the register list holds: each further pass over
            // the candidates collects the 16 smallest keys above the last one applied
            // and applies them, so the order stays ascending
            while (done < total) {
                L.init();
                enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float ax, float ay, float dist,
                              float rs, bool sen_a, float mp) {
                    correction(
                        false,
                        [&](uint64_t k, float x, float y) {
                            if (k > last) L.add(k, x, y);
                        },
                        key, isA, j, mp, ax, ay, dist, rs, sen_a);
                });
                if (L.n == 0) break;
                last = L.apply(p);
                done += L.n < 16 ? L.n : 16;
            }
            d.b.pos[b] = p;
        }
        la.contacts += count;
        la.flags |= flags;
    }
    flush_accum(la, d.accum, d.block_contacts, 0, false);
}

constexpr uint32_t kColList = 16;    // corrections a collider keeps (the smallest keys)
constexpr uint32_t kBodyMerge = 64;  // corrections a body merges in LDS

// Contacts, first half of the two-kernel form: one thread per collider walks
// its own stencil (physics.rs:248-300 for the pairs it is in), counts the
// pairs it is A in, sends their events, records push-out rows, and keeps its
// body's side of each correction -- the kColList smallest keys, in key
// order -- in clist[i * nc + c], with the full count in ccnt[c].
template <bool GROUPS, bool SENSORS>
__global__ __launch_bounds__(kContactThreads) void k_contact_lists(DeviceWorld d, SubstepArgs a) {
    LocalAccum la;
    la.init();
    // threads walk the colliders in cell order (the order the key pass just sorted them
    // into): a wave's colliders are neighbours, so their candidate records overlap in cache,
    // and each one's own record is the slot's (x, y, radius, parent), loaded coalesced
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot == 0) d.cfb_n[0] = 0u;  // k_contact_merge fills the fallback list after this kernel
    if (slot < d.nc) {
        const uint32_t c = d.vals[slot];
        const float4 own = d.crec[slot];
        const int32_t b = __float_as_int(own.w);
        const float2 xa = make_float2(own.x, own.y);
        uint32_t count = 0, flags = 0;
        // keyed by the partner collider j: for this collider c, the pair key (max, min) of
        // (c, j) orders as j does (every j < c before every j > c, each side by j)
        ContactList<kColList, uint32_t> L;
        L.init();
        if (b >= 0 && finitef(xa.x) && finitef(xa.y)) {
            const GridParams g = *d.grid;
            const float D = read_dmax(d.accum);
            const float ra = own.z;
            const float mown = d.b.mass[b];
            uint2 ga = make_uint2(0, 0);
            if (GROUPS) ga = d.c.groups[c];
            const bool sen_a = SENSORS ? (d.c.cflags[c] & 1u) : false;
            const float Q = ra + a.rmax + D;
            auto enumerate = [&](auto&& emit) {
                scan_stencil(g, d.cell_start, d.nc, d.crec, xa.x - a.drift_prev_x, xa.y - a.drift_prev_y, Q,
                             [&](uint32_t t, const float4 (&rb)[4], uint32_t nv) {
                    bool hit[4];
                    float dx[4], dy[4], dot[4], rs[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        const int32_t pj = __float_as_int(rb[u].w);
                        dx[u] = xa.x - rb[u].x;
                        dy[u] = xa.y - rb[u].y;
                        rs[u] = ra + rb[u].z;
                        dot[u] = dx[u] * dx[u] + dy[u] * dy[u];
                        hit[u] = u < nv && pj >= 0 && pj != b &&
                                 !(dot[u] > (rs[u] * rs[u]) * 1.00000095367431640625f) && sqrt_lt(dot[u], rs[u]);
                    }
                    uint32_t jj[4];
                    float mp[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        jj[u] = 0u;
                        mp[u] = 0.0f;
                        if (hit[u]) {
                            jj[u] = d.vals[t + u];
                            mp[u] = d.b.mass[__float_as_int(rb[u].w)];
                        }
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        if (!hit[u]) continue;
                        const uint32_t j = jj[u];
                        if (GROUPS) {
                            uint2 gb = d.c.groups[j];
                            if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) continue;
                        }
                        const bool isA = c > j;
                        const float ax = isA ? dx[u] : rb[u].x - xa.x, ay = isA ? dy[u] : rb[u].y - xa.y;
                        const uint64_t key = isA ? (((uint64_t)c << 32) | j) : (((uint64_t)j << 32) | c);
                        emit(key, isA, j, __float_as_int(rb[u].w), ax, ay, sqrtf(dot[u]), rs[u], mp[u]);
                    }
                });
            };
            enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float ax, float ay, float dist, float rs,
                          float mpart) {
                if (isA) ++count;
                // correction of this body's side (as k_contacts; a push-out only records its row)
                float px = 0.0f, py = 0.0f;
                const bool pushed = dist < 1e-6f;
                if (pushed) {
                    ax = ax + 0.02f;
                    dist = vlen(ax, ay);
                    px = isA ? 0.01f : -0.01f;
                    py = isA ? 0.0f : -0.0f;
                    flags |= INFO_PUSH_OUT;
                    if (isA) {
                        const uint32_t bit = 1u << (c & 31u);
                        if (!(atomicOr(&d.ord_mark[c >> 5], bit) & bit)) d.ord_rows[atomicAdd(d.ord_n, 1u)] = c;
                    }
                }
                const bool sen_b = SENSORS ? (d.c.cflags[j] & 1u) : false;
                if (sen_a || sen_b) {
                    if (pushed) L.add(j, px, py);
                    return;
                }
                const float mA = isA ? mown : mpart, mB = isA ? mpart : mown;
                const float ratio = 1.0f - mA / (mA + mB);
                const float nx = ax / dist, ny = ay / dist;
                if (nx != nx || ny != ny) flags |= ERR_NAN_NORMAL;
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
                L.add(j, cx, cy);
            });
            if (d.events && count > 0) {
                // collision events of the pairs whose A collider is this one (physics.rs:288-311)
                unsigned long long e = atomicAdd(&d.accum->event_count, (unsigned long long)count);
                const float2 va = d.b.vel[b];
                enumerate([&](uint64_t key, bool isA, uint32_t j, int32_t pj, float, float, float, float, float) {
                    if (!isA) return;
                    if (e < d.event_cap) {
                        DevEvent ev;
                        ev.sub = a.sub_global;
                        ev.a = (uint32_t)(key >> 32);
                        ev.b = (uint32_t)key;
                        ev.pad = 0;
                        ev.va = va;
                        ev.vb = d.b.vel[pj];
                        d.events[e] = ev;
                    }
                    ++e;
                });
            }
        }
        const uint32_t nk = L.n < kColList ? L.n : kColList;
#pragma unroll
        for (uint32_t i = 0; i < kColList; ++i)
            if (i < nk)
                d.clist[(size_t)i * d.nc + c] =
                    make_float4(__uint_as_float(c > L.key[i] ? c : L.key[i]), __uint_as_float(c > L.key[i] ? L.key[i] : c),
                                L.cx[i], L.cy[i]);
        d.ccnt[c] = L.n;
        la.contacts = count;
        la.flags = flags;
    }
    flush_accum(la, d.accum, d.block_contacts, 0, false);
}

// Second half: one thread per body gathers its colliders' corrections and
// adds them to the position in key order -- the reference's (i, j<i) order on
// this body.  Each collider's list is already in key order, so a body of at
// most kMergeRuns colliders stages its lists in LDS (all loads of a round in
// flight together) and merges them with the runs' heads in registers: one
// dependent LDS read per correction.  A body of more colliders sorts its
// staged corrections by insertion.  A body with a collider past kColList
// corrections, or more than kBodyMerge in all, goes to the fallback list
// (k_contacts in list mode re-walks it).
constexpr uint32_t kMergeRuns = 8;   // colliders a body merges as runs
constexpr uint32_t kMergeBatch = 16;  // corrections staged per round

__device__ __forceinline__ uint64_t merge_key(float4 v) {
    return ((uint64_t)__float_as_uint(v.x) << 32) | __float_as_uint(v.y);
}

__global__ __launch_bounds__(kContactThreads) void k_contact_merge(DeviceWorld d) {
    __shared__ float4 s_e[kBodyMerge][kContactThreads];
    const uint32_t lane = threadIdx.x;
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nb) return;
    float2 p = d.b.pos[b];
    d.pos_pre[b] = p;  // the ordered pass re-sums from here if a push-out happens
    const uint32_t k0 = d.bc_start[b], k1 = d.bc_start[b + 1];
    if (k1 - k0 <= kMergeRuns) {
        const uint32_t nr = k1 - k0;
        uint32_t cid[kMergeRuns], base[kMergeRuns + 1];
#pragma unroll
        for (uint32_t r = 0; r < kMergeRuns; ++r) cid[r] = r < nr ? d.bc_list[k0 + r] : 0u;
        bool fallback = false;
        base[0] = 0;
#pragma unroll
        for (uint32_t r = 0; r < kMergeRuns; ++r) {
            const uint32_t n = r < nr ? d.ccnt[cid[r]] : 0u;
            fallback |= n > kColList;
            base[r + 1] = base[r] + n;
        }
        const uint32_t total = base[kMergeRuns];
        if (fallback || total > kBodyMerge) {
            d.cfb_list[atomicAdd(d.cfb_n, 1u)] = b;
            return;
        }
        if (total == 0) return;
        // stage: entry e of the body is entry e - base[r] of run r (the last run starting at or
        // before e; an empty run starts where the next one does)
        for (uint32_t e0 = 0; e0 < total; e0 += kMergeBatch) {
            float4 v[kMergeBatch];
#pragma unroll
            for (uint32_t u = 0; u < kMergeBatch; ++u) {
                const uint32_t e = e0 + u;
                uint32_t c = cid[0], b0 = 0;
#pragma unroll
                for (uint32_t r = 1; r < kMergeRuns; ++r) {
                    const bool at = base[r] <= e;
                    c = at ? cid[r] : c;
                    b0 = at ? base[r] : b0;
                }
                v[u] = e < total ? d.clist[(size_t)(e - b0) * d.nc + c] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (uint32_t u = 0; u < kMergeBatch; ++u)
                if (e0 + u < total) s_e[e0 + u][lane] = v[u];
        }
        // k-way merge: each run's head key and correction in registers
        uint64_t hk[kMergeRuns];
        float2 hv[kMergeRuns];
        uint32_t pos[kMergeRuns];
#pragma unroll
        for (uint32_t r = 0; r < kMergeRuns; ++r) {
            pos[r] = base[r];
            const float4 v = base[r] < base[r + 1] ? s_e[base[r]][lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            hk[r] = base[r] < base[r + 1] ? merge_key(v) : ~0ull;
            hv[r] = make_float2(v.z, v.w);
        }
        for (uint32_t m = 0; m < total; ++m) {
            uint32_t rb = 0;
            uint64_t kb = hk[0];
            float2 vb = hv[0];
#pragma unroll
            for (uint32_t r = 1; r < kMergeRuns; ++r) {
                const bool lt = hk[r] < kb;
                rb = lt ? r : rb;
                kb = lt ? hk[r] : kb;
                vb = lt ? hv[r] : vb;
            }
            p.x = p.x + vb.x;
            p.y = p.y + vb.y;
#pragma unroll
            for (uint32_t r = 0; r < kMergeRuns; ++r) {
                if (r == rb) {
                    const uint32_t q = pos[r] + 1u;
                    pos[r] = q;
                    const float4 v = q < base[r + 1] ? s_e[q][lane] : make_float4(0.f, 0.f, 0.f, 0.f);
                    hk[r] = q < base[r + 1] ? merge_key(v) : ~0ull;
                    hv[r] = make_float2(v.z, v.w);
                }
            }
        }
        d.b.pos[b] = p;
        return;
    }
    uint32_t total = 0;
    bool fallback = false;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t n = d.ccnt[d.bc_list[k]];
        fallback |= n > kColList;
        total += n;
    }
    if (fallback || total > kBodyMerge) {
        d.cfb_list[atomicAdd(d.cfb_n, 1u)] = b;
        return;
    }
    if (total == 0) return;
    uint32_t m = 0;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t c = d.bc_list[k];
        const uint32_t n = d.ccnt[c];
        for (uint32_t i = 0; i < n; ++i) s_e[m++][lane] = d.clist[(size_t)i * d.nc + c];
    }
    // insertion sort by key (runs of one collider are already sorted)
    for (uint32_t i = 1; i < m; ++i) {
        const float4 v = s_e[i][lane];
        const uint64_t kv = merge_key(v);
        uint32_t j = i;
        while (j > 0) {
            const float4 u = s_e[j - 1][lane];
            if (merge_key(u) <= kv) break;
            s_e[j][lane] = u;
            --j;
        }
        s_e[j][lane] = v;
    }
    for (uint32_t i = 0; i < m; ++i) {
        const float4 v = s_e[i][lane];
        p.x = p.x + v.z;
        p.y = p.y + v.w;
    }
    d.b.pos[b] = p;
}

// One fixed joint (physics.rs:424-477) on body positions pa, pb and rotations
// ra, rb (read before any write of this joint).  Returns false for coincident
// anchors, which the reference skips; the caller then writes nothing.
// The joint's mass terms (physics.rs:447-460) depend on the two bodies'
// calculated_mass only, which no stage of a step changes: they come in
// precomputed (DeviceWorld::j_mass, same f32 operations) and stay off the chain.
// Rotation correction of one joint from its axis (physics.rs:462-469): the
// reference adds corr * dt to rotation a and subtracts it from rotation b.
// Positions never read rotations inside the joint stage (anchors are not
// rotated, Q7), so these can be evaluated after the position sweep.
__device__ __forceinline__ float joint_rotation_corr(float dx, float dy, float jtarget) {
    float angle_a = atan2_f32(dy, dx);
    float angle_b = -atan2_f32(dy, -dx);
    float diff = (angle_b - angle_a) - jtarget;
    return diff * 0.5f;
}

__device__ __forceinline__ void joint_rotate(float& ra, float& rb, float corr, float h, uint32_t& flags) {
    ra = ra + corr * h;
    rb = rb - corr * h;
    if (!finitef(ra) || !finitef(rb)) flags |= ERR_JOINT_ROT;
}

// One fixed joint, positions then rotations.
__device__ __forceinline__ bool joint_solve(float2& pa, float2& pb, float& ra, float& rb, float2 an_a, float2 an_b,
                                            float jdist, float jtarget, float4 jm, bool sa, bool sb, float h,
                                            uint32_t& flags) {
    float dx, dy;
    if (!joint_positions(pa, pb, an_a, an_b, jdist, jm, sa, sb, dx, dy, flags)) return false;
    joint_rotate(ra, rb, joint_rotation_corr(dx, dy, jtarget), h, flags);
    return true;
}

// Fixed joints (physics.rs:424-477), positions: Gauss-Seidel in arena order
// inside each connected component of the joint graph (components touch
// disjoint bodies, so they run in parallel, one lane each).  A block of 64
// components first stages, with coalesced loads, the constant records of all
// their joints (stored in component order at upload) and the positions of
// the bodies of its components of at most kJointBodies bodies and
// kJointJoints joints; each such component then iterates on those LDS copies
// (the sweep is a chain of dependent reads and writes, so each step costs an
// LDS round trip instead of a cache round trip), and the block writes its
// bodies back.  A larger component, or every component of a block whose
// joints or bodies exceed the block's LDS, iterates in global memory.
// Writes keep the reference's order (a's position, b's, a's rotation, b's),
// which decides a joint of a body to itself.  With `defer`,
// every solve's axis goes to jrec[iterations * jc_start[c] + it * nq + q] --
// record (dx, dy, target angle, 1), or 0 when the anchors coincide and the
// reference skips the joint -- and the rotations are applied later:
// k_integrate evaluates each of its body's solves' atan2 pairs and adds the
// corrections in sequence order.  Without it
// (joint_iterations changed since the upload) the rotations are applied here.
constexpr uint32_t kBlockJoints = 512;  // joints of a block's components staged in LDS
constexpr uint32_t kBlockBodies = 640;  // bodies of its small components staged in LDS

__global__ __launch_bounds__(64) void k_joints(DeviceWorld d, SubstepArgs a, uint32_t iterations, bool defer) {
    __shared__ float4 s_anchor[kBlockJoints];  // anchor a, anchor b
    __shared__ float4 s_jm[kBlockJoints];      // mass terms
    __shared__ float2 s_len[kBlockJoints];     // distance, target angle
    __shared__ uint32_t s_loc[kBlockJoints];   // local a | local b << 8
    __shared__ float2 s_pos[kBlockBodies];
    __shared__ float s_rot[kBlockBodies];
    __shared__ int32_t s_gid[kBlockBodies];    // body index
    __shared__ uint8_t s_stat[kBlockBodies];   // 1: static body
    const uint32_t lane = threadIdx.x;
    const uint32_t c0 = blockIdx.x * 64u, c1 = min(c0 + 64u, d.n_jcomp);
    const uint32_t comp = c0 + lane;
    const uint32_t q0 = d.jc_start[c0], q1 = d.jc_start[c1];
    const uint32_t b0 = d.jb_start[c0], b1 = d.jb_start[c1];
    const bool staged = q1 - q0 <= kBlockJoints && b1 - b0 <= kBlockBodies;  // the same for the block
    uint32_t flags = 0;
    if (staged) {
        for (uint32_t i = lane; i < q1 - q0; i += 64u) {
            s_anchor[i] = d.jq_anchor[q0 + i];
            s_jm[i] = d.jq_mass[q0 + i];
            s_len[i] = d.jq_len[q0 + i];
            s_loc[i] = d.j_local[q0 + i];
        }
        for (uint32_t i = lane; i < b1 - b0; i += 64u) {
            const int32_t g = d.jb_list[b0 + i];
            s_gid[i] = g;
            s_pos[i] = d.b.pos[g];
            if (!defer) s_rot[i] = d.b.rot[g];
            s_stat[i] = (d.b.bflags[g] & BF_STATIC) ? 1u : 0u;
        }
        __syncthreads();
    }
    if (comp < c1) {
        const uint32_t k0 = d.jc_start[comp], k1 = d.jc_start[comp + 1];
        const uint32_t lb = d.jb_start[comp] - b0, nbl = d.jb_start[comp + 1] - d.jb_start[comp];
        const uint32_t nq = k1 - k0;
        float4* rec = defer ? d.jrec + (size_t)iterations * k0 : nullptr;
        if (staged && nbl > 0) {
            const uint32_t lq = k0 - q0;
            for (uint32_t it = 0; it < iterations; ++it) {
                for (uint32_t q = 0; q < nq; ++q) {
                    const uint32_t loc = s_loc[lq + q];
                    const uint32_t la = lb + (loc & 0xFFu), lbb = lb + ((loc >> 8) & 0xFFu);
                    const float4 an = s_anchor[lq + q];
                    const float2 len = s_len[lq + q];
                    float2 pa = s_pos[la], pb = s_pos[lbb];
                    float dx, dy;
                    const bool moved = joint_positions(pa, pb, make_float2(an.x, an.y), make_float2(an.z, an.w),
                                                       len.x, s_jm[lq + q], s_stat[la] != 0u, s_stat[lbb] != 0u,
                                                       dx, dy, flags);
                    if (defer) rec[it * nq + q] = moved ? make_float4(dx, dy, len.y, 1.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
                    if (!moved) continue;
                    s_pos[la] = pa;
                    s_pos[lbb] = pb;
                    if (!defer) {
                        float ra = s_rot[la], rb = s_rot[lbb];
                        joint_rotate(ra, rb, joint_rotation_corr(dx, dy, len.y), a.h, flags);
                        s_rot[la] = ra;
                        s_rot[lbb] = rb;
                    }
                }
            }
        } else if (!staged || nbl == 0) {
            for (uint32_t it = 0; it < iterations; ++it) {
                for (uint32_t kk = k0; kk < k1; ++kk) {
                    const uint32_t j = d.jc_list[kk];
                    const int32_t ia = d.j_a[j], ib = d.j_b[j];
                    float2 pa = d.b.pos[ia], pb = d.b.pos[ib];
                    float dx, dy;
                    const bool moved =
                        joint_positions(pa, pb, d.j_anchor_a[j], d.j_anchor_b[j], d.j_dist[j], d.j_mass[j],
                                        d.b.bflags[ia] & BF_STATIC, d.b.bflags[ib] & BF_STATIC, dx, dy, flags);
                    if (defer)
                        rec[it * nq + (kk - k0)] =
                            moved ? make_float4(dx, dy, d.j_target[j], 1.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
                    if (!moved) continue;
                    d.b.pos[ia] = pa;
                    d.b.pos[ib] = pb;
                    if (!defer) {
                        float ra = d.b.rot[ia], rb = d.b.rot[ib];
                        joint_rotate(ra, rb, joint_rotation_corr(dx, dy, d.j_target[j]), a.h, flags);
                        d.b.rot[ia] = ra;
                        d.b.rot[ib] = rb;
                    }
                }
            }
        }
    }
    if (staged) {
        __syncthreads();
        for (uint32_t i = lane; i < b1 - b0; i += 64u) {
            const int32_t g = s_gid[i];
            d.b.pos[g] = s_pos[i];
            if (!defer) d.b.rot[g] = s_rot[i];
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
        const float rot_in = s.rot;
        if (a.joint_rot) {
            // this body's joint rotation corrections (physics.rs:468-469), in the order the
            // joint stage made them: iterations, then the component's joints in arena order
            const int32_t c = d.bj_comp[b];
            if (c >= 0) {
                const uint32_t k0 = d.jc_start[c], nq = d.jc_start[c + 1] - k0;
                const uint32_t e0 = d.bj_start[b], e1 = d.bj_start[b + 1];
                const float4* rec = d.jrec + (size_t)a.joint_iters * k0;
                for (uint32_t it = 0; it < a.joint_iters; ++it)
                    for (uint32_t e = e0; e < e1; ++e) {
                        const uint32_t ent = d.bj_list[e];
                        const float4 x = rec[it * nq + ((ent >> 1) - k0)];
                        if (x.w == 0.0f) continue;  // the reference skipped that solve
                        // both bodies of the solve evaluate its atan2 pair (the same f32 result)
                        const float dlt = joint_rotation_corr(x.x, x.y, x.z) * a.h;
                        s.rot = (ent & 1u) ? s.rot - dlt : s.rot + dlt;
                        if (!finitef(s.rot)) la.flags |= ERR_JOINT_ROT;
                    }
            }
        }
        if (fl & BF_STATIC) {  // :327-332
            d.b.pos_old[b] = s.p;
            d.b.vel[b] = make_float2(0.0f, 0.0f);
            if (__float_as_uint(s.rot) != __float_as_uint(rot_in)) d.b.rot[b] = s.rot;
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
            const float om0 = s.omega;
            verlet_dynamic(s, acc, toi, vr, vreq, a.h, ratio);
            d.b.pos_old[b] = s.po;
            d.b.vel[b] = s.vel;
            if (__float_as_uint(s.rot) != __float_as_uint(rot_in)) d.b.rot[b] = s.rot;
            if (__float_as_uint(s.omega) != __float_as_uint(om0)) d.b.omega[b] = s.omega;
            if (vr) d.b.bflags[b] = fl & ~BF_HAS_VREQ;
            if (a.fresh) d.b.torque[b] = 0.0f;
        }
        // collider transforms from the pre-constraint position (:360-366), four colliders per
        // round so that their loads overlap (an index past the end repeats the last collider:
        // the same store, and the bounding box and displacement maximum are idempotent)
        const uint32_t c0 = d.bc_start[b], c1 = d.bc_start[b + 1];
        for (uint32_t k0 = c0; k0 < c1; k0 += 4) {
            uint32_t c[4];
            float2 off[4], x0[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) c[u] = d.bc_list[min(k0 + u, c1 - 1u)];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                off[u] = d.c.off[c[u]];
                x0[u] = d.x0[c[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                const float2 x = body_xform(s.rot, s.p, off[u]);
                d.c.abs[c[u]] = x;
                la.point(x.x, x.y, x0[u].x, x0[u].y, a.drift_x, a.drift_y);
            }
        }
        apply_constraints(s.p, d.con_pos, d.con_r, d.nk);
        d.b.pos[b] = s.p;
    }
    flush_accum(la, d.accum, d.block_contacts, a.parity, true);
}

}  // namespace

// ============================================================ launchers

// Counting sort by cell: cell_start = exclusive scan of the counts of the
// fitted grid's cells (cell_start[ncells] = nc), then the scatter.
hipError_t launch_rebuild_sort_and_cells(DeviceWorld& d) {
    // the grid the key pass fitted holds ncells <= cells_cap cells: the scan covers those
    // (the count is read on the device), not the whole table
    const uint32_t* last =
        reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(d.grid) + offsetof(GridParams, ncells));
    hipError_t e = exclusive_scan_u32(d.cell_count, d.cell_start, d.cells_cap + 1, d.sort, d.stream, last);
    if (e != hipSuccess) return e;
    const uint32_t n = d.nc > 0 ? d.nc : 1u;
    k_cell_scatter<<<div_up(n, kThreads), kThreads, 0, d.stream>>>(d.keys, d.vals_tmp, d.nc, d.cell_start, d.vals,
                                                                   d.cell_count, d.cells_cap + 1, d.c.abs,
                                                                   d.c.radius, d.c.parent, d.crec, d.grid);
    return hipGetLastError();
}

hipError_t launch_general_substep(DeviceWorld& d, const SubstepArgs& a, uint32_t joint_iterations, bool groups,
                                  bool sensors, const KeysArgs* rebuild) {
    if (d.nb == 0) return hipSuccess;
    const uint32_t blocks = div_up(d.nb, kThreads);
    // the rebuild's keys and the substep's forces in one launch, then the rest of the rebuild
    const uint32_t key_blocks = rebuild ? div_up(rebuild->nc > 0 ? rebuild->nc : 1u, kThreads) : 0u;
    const uint32_t force_blocks = a.has_springs ? blocks : 0u;
    if (key_blocks + force_blocks > 0)
        k_keys_forces<<<key_blocks + force_blocks, kThreads, 0, d.stream>>>(rebuild ? *rebuild : KeysArgs{},
                                                                            key_blocks, d, a);
    if (rebuild) {
        hipError_t e = launch_rebuild_sort_and_cells(d);
        if (e != hipSuccess) return e;
    }
    if (a.collisions && d.nc > 0) {
        // One-wave blocks: a body's work varies with its collider count, so small blocks
        // keep the slow waves from holding whole workgroups.  With the collider lists
        // allocated: one thread per collider, a per-body merge, and the full per-body
        // walk only for the bodies the merge could not take.
        const uint32_t cb = div_up(d.nb, kContactThreads);
#define BG_CONTACTS(G, S)                                                                                     \
    do {                                                                                                      \
        if (d.clist) {                                                                                        \
            k_contact_lists<G, S><<<div_up(d.nc, kContactThreads), kContactThreads, 0, d.stream>>>(d, a);     \
            k_contact_merge<<<cb, kContactThreads, 0, d.stream>>>(d);                                         \
            k_contacts<G, S><<<std::min(cb, 256u), kContactThreads, 0, d.stream>>>(d, a, d.cfb_list);         \
        } else {                                                                                              \
            k_contacts<G, S><<<cb, kContactThreads, 0, d.stream>>>(d, a, nullptr);                            \
        }                                                                                                     \
    } while (0)
        if (groups && sensors)
            BG_CONTACTS(true, true);
        else if (groups)
            BG_CONTACTS(true, false);
        else if (sensors)
            BG_CONTACTS(false, true);
        else
            BG_CONTACTS(false, false);
#undef BG_CONTACTS
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launch_contacts_ordered(d, a, groups, sensors);
        if (e != hipSuccess) return e;
    }
    SubstepArgs ai = a;
    ai.joint_rot = 0u;
    if (d.n_jcomp > 0 && joint_iterations > 0) {
        const bool defer = d.jrec && joint_iterations == d.j_iters;
        k_joints<<<div_up(d.n_jcomp, 64), 64, 0, d.stream>>>(d, a, joint_iterations, defer);
        if (defer) {
            ai.joint_rot = 1u;
            ai.joint_iters = joint_iterations;
        }
    }
    k_integrate<<<blocks, kThreads, 0, d.stream>>>(d, ai);
    return hipGetLastError();
}

}  // namespace bg
