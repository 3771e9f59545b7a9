// diff_kernels.hip -- forward and adjoint of rollouts of single-sphere worlds
// with springs, one 1024-thread workgroup per world for a whole segment of
// steps (two bodies per thread; snapshot, velocities, the cell grid and the
// adjoint vectors live in LDS).
//
// Forward substep (physics.rs:401-421, same f32 operation order as the
// general path): gravity and springs per body in ascending spring index
// (springs.rs:25-47), contacts against the substep's snapshot summed in
// ascending partner index (physics.rs:241-317), Verlet (physics.rs:323-358),
// collider transform = position (zero offsets).  The grid is rebuilt in LDS
// every substep: bounding box, cell counts with LDS atomics, a block scan,
// and a scatter of body indices.
//
// Adjoint substep, exact derivative of that map with the substep's pair set
// held fixed (pairs with dist < 1e-6 take the push-out branch and are held
// constant).  With w_i = m_j / (m_i + m_j) and a = s_i - s_j, the
// correction on i is w_i f(a), f(a) = (r_i + r_j - |a|) a / |a|, whose
// Jacobian is J = (rs/|a|)(I - n n^T) - I; the spring force
// F = u k (L - rest) - c (rv . u) (u_x^2, u_y^2) (springs.rs:38-43, the
// damping vector multiplies the direction componentwise), rv = v_a - v_b,
// has dF/dd = k[(1 - rest/L)(I - u u^T) + u u^T] + (damping through u,
// du/dd = (I - u u^T)/L), dF/drv = -c (u_x^2, u_y^2) u^T and
// dF/dk = u (L - rest).  The velocity v = displacement / h of the previous
// Verlet is part of the adjoint state with damping.  A fixed joint moves
// p_a += w_a corr and p_b -= w_b corr, corr = (|d| - D) d/|d| with d the
// anchor axis; dcorr/dd = I - (D/|d|)(I - n n^T), and the reverse sweep
// walks the joints in the opposite order on the axes of the forward sweep.
#include "diff.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

constexpr int kWaves = kDiffThreads / BG_WAVE;

struct DiffGrid {
    float ox, oy, inv_cell;
    int32_t nx, ny;
    uint32_t ncells;
};

struct Ctx {
    float4* snap;      // x, y, radius, mass
    float2* vel;       // calculated_velocity
    uint32_t* cnt;     // cell counts, then cell starts
    uint16_t* sorted;  // body index in cell order
    uint32_t* red;     // [4][kWaves]
    double* dred;      // [2][kWaves]
    DiffGrid* grid;
    float2* lam_pp;  // adjoint only
    float2* lam_am;
};

struct Own {
    float2 p[kDiffPer], q[kDiffPer], v[kDiffPer];
    float m[kDiffPer], r[kDiffPer], gm[kDiffPer];
    uint32_t fl[kDiffPer], key[kDiffPer], rank[kDiffPer];
    uint32_t li[kDiffPer], gi[kDiffPer];
    bool live[kDiffPer];
};

__device__ __forceinline__ void load_own(Own& o, const DiffArgs& a, const DiffWorld& W, const float2* p,
                                         const float2* q, const float2* v) {
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        o.li[j] = threadIdx.x + j * kDiffThreads;
        o.live[j] = o.li[j] < W.nb;
        o.gi[j] = W.b0 + (o.live[j] ? o.li[j] : 0u);
        const uint32_t g = o.gi[j];
        o.p[j] = o.live[j] ? p[g] : make_float2(0.0f, 0.0f);
        o.q[j] = o.live[j] ? q[g] : make_float2(0.0f, 0.0f);
        o.v[j] = o.live[j] ? v[g] : make_float2(0.0f, 0.0f);
        o.m[j] = a.mass[g];
        o.r[j] = a.radius[g];
        o.gm[j] = a.gmod[g];
        o.fl[j] = a.flags[g];
        o.key[j] = 0;
        o.rank[j] = 0;
    }
}

__device__ __forceinline__ void publish(Ctx& c, const Own& o) {
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j)
        if (o.live[j]) {
            c.snap[o.li[j]] = make_float4(o.p[j].x, o.p[j].y, o.r[j], o.m[j]);
            c.vel[o.li[j]] = o.v[j];
        }
}

// Exclusive scan of cnt[0, n) by the block (n <= 5 * kDiffThreads).
__device__ void block_scan_cnt(Ctx& c, uint32_t n) {
    constexpr uint32_t kItems = (kDiffCells + 1 + kDiffThreads - 1) / kDiffThreads;
    const uint32_t base = threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        v[k] = (base + k < n) ? c.cnt[base + k] : 0u;
        sum += v[k];
    }
    const uint32_t incl = wave_incl_scan_u32(sum);
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    if (lane == BG_WAVE - 1) c.red[wid] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wid; ++w) off += c.red[w];
    uint32_t run = off + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        if (base + k < n) c.cnt[base + k] = run;
        run += v[k];
    }
    __syncthreads();
}

// Cell grid of the current snapshot in LDS (counting sort of body indices).
__device__ void build_grid(Ctx& c, Own& o, const DiffWorld& W) {
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    uint32_t b0 = 0xFFFFFFFFu, b1 = 0xFFFFFFFFu, b2 = 0u, b3 = 0u;
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        if (!o.live[j]) continue;
        const float4 s = c.snap[o.li[j]];
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
        c.red[wid] = b0;
        c.red[kWaves + wid] = b1;
        c.red[2 * kWaves + wid] = b2;
        c.red[3 * kWaves + wid] = b3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) {
            b0 = min(b0, c.red[w]);
            b1 = min(b1, c.red[kWaves + w]);
            b2 = max(b2, c.red[2 * kWaves + w]);
            b3 = max(b3, c.red[3 * kWaves + w]);
        }
        b0 = min(b0, c.red[0]);
        b1 = min(b1, c.red[kWaves]);
        b2 = max(b2, c.red[2 * kWaves]);
        b3 = max(b3, c.red[3 * kWaves]);
        float minx = 0.0f, miny = 0.0f, maxx = 0.0f, maxy = 0.0f;
        if (b0 <= b2 && b1 <= b3) {
            minx = ord2f(b0);
            miny = ord2f(b1);
            maxx = ord2f(b2);
            maxy = ord2f(b3);
        }
        float cell = 2.0f * W.rmax;
        if (!(cell > 0.0f) || !finitef(cell)) cell = 1.0f;
        const float ex = maxx - minx, ey = maxy - miny;
        int32_t nx = 1, ny = 1;
        for (;;) {
            const float fx = floorf(ex / cell) + 1.0f, fy = floorf(ey / cell) + 1.0f;
            if (fx <= 65536.0f && fy <= 65536.0f && (uint64_t)fx * (uint64_t)fy <= (uint64_t)kDiffCells) {
                nx = (int32_t)fx;
                ny = (int32_t)fy;
                break;
            }
            cell = cell * 2.0f;
        }
        c.grid->ox = minx;
        c.grid->oy = miny;
        c.grid->inv_cell = 1.0f / cell;
        c.grid->nx = nx;
        c.grid->ny = ny;
        c.grid->ncells = (uint32_t)nx * (uint32_t)ny;
    }
    __syncthreads();
    const DiffGrid g = *c.grid;
    for (uint32_t i = threadIdx.x; i <= g.ncells; i += kDiffThreads) c.cnt[i] = 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        if (!o.live[j]) continue;
        const float4 s = c.snap[o.li[j]];
        const int32_t cx = cell_coord(s.x, g.ox, g.inv_cell, g.nx);
        const int32_t cy = cell_coord(s.y, g.oy, g.inv_cell, g.ny);
        o.key[j] = (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
        o.rank[j] = atomicAdd(&c.cnt[o.key[j]], 1u);
    }
    __syncthreads();
    block_scan_cnt(c, g.ncells + 1);
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j)
        if (o.live[j]) c.sorted[c.cnt[o.key[j]] + o.rank[j]] = (uint16_t)o.li[j];
    __syncthreads();
}

// Bodies of the grid rows overlapping the square of half-width Q around (qx, qy).
template <class F>
__device__ __forceinline__ void grid_query(const Ctx& c, float qx, float qy, float Q, F&& visit) {
    const DiffGrid g = *c.grid;
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
        const uint32_t t0 = c.cnt[row + (uint32_t)cx0], t1 = c.cnt[row + (uint32_t)cx1 + 1u];
        for (uint32_t t = t0; t < t1; ++t) visit((uint32_t)c.sorted[t]);
    }
}

// Joint weights (physics.rs:450-460): a static -> b moves by ims corr; b
// static -> a does; otherwise a by ratio corr and b by (1 - ratio) corr.
__device__ __forceinline__ void joint_weights(const DiffArgs& a, const DiffWorld& W, uint32_t q, uint32_t la,
                                              uint32_t lb, bool& sa, bool& sb, float4& jm) {
    sa = (a.flags[W.b0 + la] & BF_STATIC) != 0;
    sb = (a.flags[W.b0 + lb] & BF_STATIC) != 0;
    jm = a.j_mass[q];
}

// Position sweep of one joint component (physics.rs:426-460) on the LDS
// positions P, joint_iterations times in arena order; with `axis`, every
// solve's anchor axis goes to axis[it * n_joints + q].
__device__ void joint_sweep(float2* P, uint32_t comp, const DiffArgs& a, const DiffWorld& W, float2* axis) {
    const uint32_t q0 = a.jc_start[comp], q1 = a.jc_start[comp + 1];
    uint32_t flags = 0;
    for (uint32_t it = 0; it < a.j_iters; ++it)
        for (uint32_t q = q0; q < q1; ++q) {
            const uint32_t ab = a.j_ab[q], la = ab & 0xFFFFu, lb = ab >> 16;
            bool sa, sb;
            float4 jm;
            joint_weights(a, W, q, la, lb, sa, sb, jm);
            const float4 an = a.j_anchor[q];
            float2 pa = P[la], pb = P[lb];
            float dx, dy;
            const bool moved = joint_positions(pa, pb, make_float2(an.x, an.y), make_float2(an.z, an.w), a.j_dist[q],
                                               jm, sa, sb, dx, dy, flags);
            if (axis) axis[(size_t)it * a.n_joints + q] = make_float2(dx, dy);
            if (!moved) continue;
            P[la] = pa;
            P[lb] = pb;
        }
}

// Adjoint of joint_sweep on the LDS position adjoints L, from the axes the
// forward sweep recorded: last solve first.
__device__ void joint_sweep_reverse(float2* L, uint32_t comp, const DiffArgs& a, const DiffWorld& W,
                                    const float2* axis) {
    const uint32_t q0 = a.jc_start[comp], q1 = a.jc_start[comp + 1];
    for (uint32_t it = a.j_iters; it-- > 0;)
        for (uint32_t q = q1; q-- > q0;) {
            const float2 d = axis[(size_t)it * a.n_joints + q];
            const float dist = vlen(d.x, d.y);
            if (dist < 1e-6f) continue;  // the reference skips the joint
            const uint32_t ab = a.j_ab[q], la = ab & 0xFFFFu, lb = ab >> 16;
            bool sa, sb;
            float4 jm;
            joint_weights(a, W, q, la, lb, sa, sb, jm);
            const float wa = sa ? 0.0f : (sb ? jm.x : jm.y);
            const float wb = sa ? jm.x : (sb ? 0.0f : jm.z);
            const float2 la2 = L[la], lb2 = L[lb];
            const float lcx = wa * la2.x - wb * lb2.x, lcy = wa * la2.y - wb * lb2.y;
            const float nx = d.x / dist, ny = d.y / dist;
            const float nm = nx * lcx + ny * lcy;
            const float f = a.j_dist[q] / dist;
            const float ldx = lcx - f * (lcx - nx * nm), ldy = lcy - f * (lcy - ny * nm);
            L[la] = make_float2(la2.x - ldx, la2.y - ldy);
            L[lb] = make_float2(lb2.x + ldx, lb2.y + ldy);
        }
}

// One forward substep of the world; with `rec`, records the snapshot,
// contact partners, velocities (damped springs) and joint-stage positions
// (joints) at tape slot `k`.
__device__ void substep_forward(Ctx& c, Own& o, const DiffArgs& a, const DiffWorld& W, uint32_t gsub, bool rec,
                                uint32_t k) {
    const float h = a.h;
    float2* tape_s = rec ? a.tape_s : nullptr;
    uint16_t* tape_nbr = rec ? a.tape_nbr : nullptr;
    float2* tape_v = rec ? a.tape_v : nullptr;
    if (a.collisions) build_grid(c, o, W);
    float2 pnew[kDiffPer], acc[kDiffPer];
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        pnew[j] = o.p[j];
        acc[j] = make_float2(0.0f, 0.0f);
        if (!o.live[j]) continue;
        const uint32_t li = o.li[j], gi = o.gi[j];
        const float4 own = c.snap[li];
        if (tape_s) tape_s[(size_t)k * a.nb_total + gi] = make_float2(own.x, own.y);
        if (tape_v) tape_v[(size_t)k * a.nb_total + gi] = c.vel[li];
        // gravity + springs (physics.rs:369-375, springs.rs:25-47), ascending spring index
        if (!(o.fl[j] & BF_STATIC)) {
            float2 ac = gsub == 0 ? a.acc0[gi] : make_float2(0.0f, 0.0f);
            ac.x = ac.x + a.gx * o.gm[j];
            ac.y = ac.y + a.gy * o.gm[j];
            const float m = o.m[j];
            for (uint32_t e = a.inc_start[gi]; e < a.inc_start[gi + 1]; ++e) {
                const uint32_t s = a.inc_list[e];
                const uint32_t ia = (uint32_t)a.sp_a[s] - W.b0, ib = (uint32_t)a.sp_b[s] - W.b0;
                const float4 pa = c.snap[ia], pb = c.snap[ib];
                const float2 va = c.vel[ia], vb = c.vel[ib];
                const float dx = pb.x - pa.x, dy = pb.y - pa.y;
                const float dist = vlen(dx, dy);
                const float ux = dx / dist, uy = dy / dist;
                const float rvx = va.x - vb.x, rvy = va.y - vb.y;
                const float dot = rvx * ux + rvy * uy;
                const float cd = a.sp_c[s] * dot;
                const float dfx = cd * ux, dfy = cd * uy;
                const float mag = a.sp_k[s] * (dist - a.sp_rest[s]);
                const float fx = ux * (mag - dfx), fy = uy * (mag - dfy);
                if (li == ia) {
                    ac.x = ac.x + fx / m;
                    ac.y = ac.y + fy / m;
                } else {
                    ac.x = ac.x + (-fx) / m;
                    ac.y = ac.y + (-fy) / m;
                }
            }
            acc[j] = ac;
        }
        // contacts against the snapshot (physics.rs:241-317), ascending partner index
        uint32_t nrec = 0;
        uint16_t rec[kDiffTapeNbr - 1];
#pragma unroll
        for (uint32_t t = 0; t + 1 < kDiffTapeNbr; ++t) rec[t] = 0xFFFFu;
        if (a.collisions && finitef(own.x) && finitef(own.y)) {
            const float ra = own.z, ma = own.w;
            auto enumerate = [&](auto&& sink) {
                grid_query(c, own.x, own.y, ra + W.rmax, [&](uint32_t lj) {
                    if (lj == li) return;
                    const float4 sb = c.snap[lj];
                    const bool isA = li > lj;
                    const float ax = isA ? own.x - sb.x : sb.x - own.x;
                    const float ay = isA ? own.y - sb.y : sb.y - own.y;
                    const float rs = ra + sb.z;
                    const float dot = ax * ax + ay * ay;
                    if (dot > (rs * rs) * 1.00000095367431640625f) return;
                    const float dist = sqrtf(dot);
                    if (!(dist < rs)) return;
                    sink(lj, sb.w, isA, ax, ay, dist, rs);
                });
            };
            auto correction = [&](auto&& put, uint32_t lj, float mpart, bool isA, float ax, float ay, float dist,
                                  float rs) {
                float px = 0.0f, py = 0.0f;
                const bool pushed = dist < 1e-6f;
                if (pushed) {
                    ax = ax + 0.02f;
                    dist = vlen(ax, ay);
                    px = isA ? 0.01f : -0.01f;
                    py = isA ? 0.0f : -0.0f;
                }
                const float mA = isA ? ma : mpart, mB = isA ? mpart : ma;
                const float ratio = 1.0f - mA / (mA + mB);
                const float nx = ax / dist, ny = ay / dist;
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
                put(lj, cx, cy);
            };
            ContactList<8, uint32_t> L;
            L.init();
            enumerate([&](uint32_t lj, float mpart, bool isA, float ax, float ay, float dist, float rs) {
                correction([&](uint32_t kk, float x, float y) { L.add(kk, x, y); }, lj, mpart, isA, ax, ay, dist,
                           rs);
            });
            nrec = L.n;
#pragma unroll
            for (uint32_t t = 0; t + 1 < kDiffTapeNbr; ++t)
                if (t < L.n) rec[t] = (uint16_t)L.key[t];
            if (L.n > 0) {
                float2 p = pnew[j];
                uint32_t last = L.apply(p);
                uint32_t done = L.n < 8 ? L.n : 8;
                while (done < L.n) {  // more than 8 contacts: next partner by selection, same order
                    uint32_t best = kNil;
                    float bx = 0.0f, by = 0.0f;
                    enumerate([&](uint32_t lj, float mpart, bool isA, float ax, float ay, float dist, float rs) {
                        correction(
                            [&](uint32_t kk, float x, float y) {
                                if (kk > last && kk < best) {
                                    best = kk;
                                    bx = x;
                                    by = y;
                                }
                            },
                            lj, mpart, isA, ax, ay, dist, rs);
                    });
                    if (best == kNil) break;
                    p.x = p.x + bx;
                    p.y = p.y + by;
                    last = best;
                    ++done;
                }
                pnew[j] = p;
            }
        }
        if (tape_nbr) {
            uint16_t* r = tape_nbr + ((size_t)k * a.nb_total + gi) * kDiffTapeNbr;
            r[0] = (uint16_t)(nrec < 0xFFFFu ? nrec : 0xFFFFu);
#pragma unroll
            for (uint32_t t = 0; t + 1 < kDiffTapeNbr; ++t) r[t + 1] = rec[t];
        }
    }
    // fixed joints, positions (physics.rs:424-460): the LDS velocity array is free until
    // publish, so it holds the positions while the component lanes sweep
    if (a.j_iters > 0 && W.njc > 0) {
        __syncthreads();  // every spring has read the velocities
#pragma unroll
        for (uint32_t j = 0; j < kDiffPer; ++j)
            if (o.live[j]) {
                c.vel[o.li[j]] = pnew[j];
                if (rec) a.tape_pj[(size_t)k * a.nb_total + o.gi[j]] = pnew[j];
            }
        __syncthreads();
        for (uint32_t comp = W.jc0 + threadIdx.x; comp < W.jc0 + W.njc; comp += kDiffThreads)
            joint_sweep(c.vel, comp, a, W, nullptr);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kDiffPer; ++j)
            if (o.live[j]) pnew[j] = c.vel[o.li[j]];
    }
    // update_objects (physics.rs:326-358); the snapshot is the new position
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        if (!o.live[j]) continue;
        if (o.fl[j] & BF_STATIC) {
            o.p[j] = pnew[j];
            o.q[j] = pnew[j];
            o.v[j] = make_float2(0.0f, 0.0f);
            continue;
        }
        BodyState st;
        st.p = pnew[j];
        st.po = o.q[j];
        st.rot = 0.0f;
        st.omega = 0.0f;
        const bool vr = gsub == 0 && (o.fl[j] & BF_HAS_VREQ);
        const float2 vq = vr ? a.vreq[o.gi[j]] : make_float2(0.0f, 0.0f);
        const float ratio = ((int32_t)o.li[j] == W.first_dyn && gsub == 0) ? h / W.old_dt0 : h / h;
        verlet_dynamic(st, acc[j], 0.0f, vr, vq, h, ratio);
        o.p[j] = st.p;
        o.q[j] = st.po;
        o.v[j] = st.vel;
    }
    __syncthreads();  // everyone is done reading the snapshot
    publish(c, o);
    __syncthreads();
}

// Adjoint of substep `gsub` (tape slot k).  In: lp, lq, lv = d loss / d
// (position, position_old, calculated_velocity) after the substep.  Out: the
// same before it; gv += d loss / d velocity_request at the substep that
// consumes it.
__device__ void substep_reverse(Ctx& c, const Own& o, const DiffArgs& a, const DiffWorld& W, uint32_t gsub,
                                uint32_t k, float2* lp, float2* lq, float2* lv, float2* gv) {
    const float h = a.h;
    const bool joints = a.j_iters > 0 && W.njc > 0;
    float2 lpp[kDiffPer], lqin[kDiffPer];
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        lpp[j] = make_float2(0.0f, 0.0f);
        lqin[j] = make_float2(0.0f, 0.0f);
        if (!o.live[j]) continue;
        const size_t t = (size_t)k * a.nb_total + o.gi[j];
        const float2 s = a.tape_s[t];
        c.snap[o.li[j]] = make_float4(s.x, s.y, o.r[j], o.m[j]);
        if (joints) c.vel[o.li[j]] = a.tape_pj[t];  // positions entering the joint stage
        else if (a.tape_v) c.vel[o.li[j]] = a.tape_v[t];
        float2 lam_a = make_float2(0.0f, 0.0f);
        if (o.fl[j] & BF_STATIC) {  // p'' = p', q'' = p', v'' = 0
            lpp[j] = make_float2(lp[j].x + lq[j].x, lp[j].y + lq[j].y);
        } else {
            // dx = (p' - q) r; p'' = p' + dx + a h^2, q'' = p', v'' = dx / h
            const float ratio = ((int32_t)o.li[j] == W.first_dyn && gsub == 0) ? h / W.old_dt0 : h / h;
            const float ldx = lp[j].x + lv[j].x / h, ldy = lp[j].y + lv[j].y / h;
            if (gsub == 0 && (o.fl[j] & BF_HAS_VREQ)) {  // q = p' - vreq h: dx = vreq h r
                lpp[j] = make_float2(lp[j].x + lq[j].x, lp[j].y + lq[j].y);
                gv[j].x += h * ratio * ldx;
                gv[j].y += h * ratio * ldy;
            } else {
                lpp[j] = make_float2((lp[j].x + lq[j].x) + ratio * ldx, (lp[j].y + lq[j].y) + ratio * ldy);
                lqin[j] = make_float2(-ratio * ldx, -ratio * ldy);
            }
            lam_a = make_float2(h * h * lp[j].x / o.m[j], h * h * lp[j].y / o.m[j]);
        }
        c.lam_pp[o.li[j]] = lpp[j];
        c.lam_am[o.li[j]] = lam_a;
    }
    __syncthreads();
    if (joints) {
        // replay each component's sweep from the recorded positions for its axes, then walk
        // it backwards on the position adjoints
        for (uint32_t comp = W.jc0 + threadIdx.x; comp < W.jc0 + W.njc; comp += kDiffThreads) {
            joint_sweep(c.vel, comp, a, W, a.j_axis);
            joint_sweep_reverse(c.lam_pp, comp, a, W, a.j_axis);
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kDiffPer; ++j)
            if (o.live[j]) {
                lpp[j] = c.lam_pp[o.li[j]];
                if (a.tape_v) c.vel[o.li[j]] = a.tape_v[(size_t)k * a.nb_total + o.gi[j]];
            }
        __syncthreads();
    }
    float2 lvin[kDiffPer];
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        lvin[j] = make_float2(0.0f, 0.0f);
        if (!o.live[j]) continue;
        const uint32_t li = o.li[j], gi = o.gi[j];
        const float4 si = c.snap[li];
        float2 ls = make_float2(0.0f, 0.0f);
        // contacts, pair set of the forward substep
        if (a.collisions && finitef(si.x) && finitef(si.y)) {
            auto pair = [&](uint32_t lj) {
                const float4 sj = c.snap[lj];
                const float ax = si.x - sj.x, ay = si.y - sj.y;
                const float d = sqrtf(ax * ax + ay * ay);
                if (!(d >= 1e-6f)) return;  // push-out branch: held constant
                const float rs = si.z + sj.z;
                const float wi = sj.w / (si.w + sj.w), wj = si.w / (si.w + sj.w);
                const float2 lj2 = c.lam_pp[lj];
                const float mux = wi * lpp[j].x - wj * lj2.x, muy = wi * lpp[j].y - wj * lj2.y;
                const float nx = ax / d, ny = ay / d;
                const float nm = nx * mux + ny * muy;
                const float f = rs / d;
                ls.x += f * (mux - nx * nm) - mux;
                ls.y += f * (muy - ny * nm) - muy;
            };
            const uint16_t* r = a.tape_nbr + ((size_t)k * a.nb_total + gi) * kDiffTapeNbr;
            const uint32_t n = r[0];
            if (n < kDiffTapeNbr) {
                for (uint32_t t = 0; t < n; ++t) pair(r[1 + t]);
            } else {  // more partners than the tape holds: the same predicate over the world
                for (uint32_t lj = 0; lj < W.nb; ++lj) {
                    if (lj == li) continue;
                    const float4 sb = c.snap[lj];
                    const bool isA = li > lj;
                    const float ax = isA ? si.x - sb.x : sb.x - si.x;
                    const float ay = isA ? si.y - sb.y : sb.y - si.y;
                    const float rs = si.z + sb.z;
                    const float dot = ax * ax + ay * ay;
                    if (dot > (rs * rs) * 1.00000095367431640625f) continue;
                    if (!(sqrtf(dot) < rs)) continue;
                    pair(lj);
                }
            }
        }
        // springs
        for (uint32_t e = a.inc_start[gi]; e < a.inc_start[gi + 1]; ++e) {
            const uint32_t s = a.inc_list[e];
            const uint32_t ia = (uint32_t)a.sp_a[s] - W.b0, ib = (uint32_t)a.sp_b[s] - W.b0;
            const float4 pa = c.snap[ia], pb = c.snap[ib];
            const float dx = pb.x - pa.x, dy = pb.y - pa.y;
            const float L = sqrtf(dx * dx + dy * dy);
            const float ux = dx / L, uy = dy / L;
            const float2 la = c.lam_am[ia], lb = c.lam_am[ib];
            const float lfx = la.x - lb.x, lfy = la.y - lb.y;
            const float uf = ux * lfx + uy * lfy;
            const float kk = a.sp_k[s], rest = a.sp_rest[s];
            const float t = kk * (1.0f - rest / L);
            float kx = t * (lfx - ux * uf) + kk * ux * uf;
            float ky = t * (lfy - uy * uf) + kk * uy * uf;
            float rvx = 0.0f, rvy = 0.0f;  // d loss / d (v_a - v_b)
            const float cc = a.sp_c[s];
            if (cc != 0.0f) {
                const float2 va = c.vel[ia], vb = c.vel[ib];
                const float vx = va.x - vb.x, vy = va.y - vb.y;
                const float dot = vx * ux + vy * uy;
                const float wl = ux * ux * lfx + uy * uy * lfy;
                // d loss / d u of -c dot (u_x^2, u_y^2), then through du/dd = (I - u u^T)/L
                const float gux = -cc * (vx * wl + 2.0f * dot * ux * lfx);
                const float guy = -cc * (vy * wl + 2.0f * dot * uy * lfy);
                const float gu = ux * gux + uy * guy;
                kx += (gux - ux * gu) / L;
                ky += (guy - uy * gu) / L;
                rvx = -cc * wl * ux;
                rvy = -cc * wl * uy;
            }
            if (li == ia) {
                ls.x -= kx;
                ls.y -= ky;
                lvin[j].x += rvx;
                lvin[j].y += rvy;
                a.grad_k[s] += (L - rest) * uf;
            } else {
                ls.x += kx;
                ls.y += ky;
                lvin[j].x -= rvx;
                lvin[j].y -= rvy;
            }
        }
        lp[j] = make_float2(lpp[j].x + ls.x, lpp[j].y + ls.y);
        lq[j] = lqin[j];
    }
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) lv[j] = lvin[j];
    __syncthreads();  // snapshot and adjoint vectors are rewritten by the next substep
}

// Centroid of the world's positions (f64 sums) by the block.
__device__ double2 block_centroid(Ctx& c, const Own& o, const DiffWorld& W) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j)
        if (o.live[j]) {
            sx += (double)o.p[j].x;
            sy += (double)o.p[j].y;
        }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off, BG_WAVE);
        sy += __shfl_xor(sy, off, BG_WAVE);
    }
    const int lane = threadIdx.x & (BG_WAVE - 1), wid = threadIdx.x / BG_WAVE;
    if (lane == 0) {
        c.dred[wid] = sx;
        c.dred[kWaves + wid] = sy;
    }
    __syncthreads();
    double tx = 0.0, ty = 0.0;
    for (int w = 0; w < kWaves; ++w) {
        tx += c.dred[w];
        ty += c.dred[kWaves + w];
    }
    __syncthreads();
    const double n = W.nb > 0 ? (double)W.nb : 1.0;
    return make_double2(tx / n, ty / n);
}

#define BG_DIFF_SHARED                                 \
    __shared__ float4 s_snap[kDiffMaxBodies];          \
    __shared__ float2 s_vel[kDiffMaxBodies];           \
    __shared__ uint32_t s_cnt[kDiffCells + 1];         \
    __shared__ uint16_t s_sorted[kDiffMaxBodies];      \
    __shared__ uint32_t s_red[4 * kWaves];             \
    __shared__ double s_dred[2 * kWaves];              \
    __shared__ DiffGrid s_grid;                        \
    Ctx c;                                             \
    c.snap = s_snap;                                   \
    c.vel = s_vel;                                     \
    c.cnt = s_cnt;                                     \
    c.sorted = s_sorted;                               \
    c.red = s_red;                                     \
    c.dred = s_dred;                                   \
    c.grid = &s_grid;                                  \
    c.lam_pp = nullptr;                                \
    c.lam_am = nullptr;

__global__ __launch_bounds__(kDiffThreads) void k_diff_forward(DiffArgs a, uint32_t n_steps, uint32_t ck_every) {
    BG_DIFF_SHARED
    const DiffWorld W = a.worlds[blockIdx.x];
    Own o;
    load_own(o, a, W, a.ck_p, a.ck_q, a.ck_v);
    publish(c, o);
    __syncthreads();
    for (uint32_t step = 0; step < n_steps; ++step) {
        if (step > 0 && step % ck_every == 0) {
            const size_t base = (size_t)(step / ck_every) * a.nb_total;
#pragma unroll
            for (uint32_t j = 0; j < kDiffPer; ++j)
                if (o.live[j]) {
                    a.ck_p[base + o.gi[j]] = o.p[j];
                    a.ck_q[base + o.gi[j]] = o.q[j];
                    a.ck_v[base + o.gi[j]] = o.v[j];
                }
        }
        for (uint32_t sub = 0; sub < a.substeps; ++sub)
            substep_forward(c, o, a, W, step * a.substeps + sub, false, 0);
    }
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j)
        if (o.live[j]) a.final_p[o.gi[j]] = o.p[j];
    const double2 cen = block_centroid(c, o, W);
    if (threadIdx.x == 0) {
        const double dx = cen.x - (double)W.tx, dy = cen.y - (double)W.ty;
        a.loss[blockIdx.x] = (float)(dx * dx + dy * dy);
    }
}

__global__ __launch_bounds__(kDiffThreads) void k_diff_backward(DiffArgs a, uint32_t seg, uint32_t first_step,
                                                                uint32_t seg_steps, uint32_t last) {
    BG_DIFF_SHARED
    __shared__ float2 s_lam_pp[kDiffMaxBodies];
    __shared__ float2 s_lam_am[kDiffMaxBodies];
    c.lam_pp = s_lam_pp;
    c.lam_am = s_lam_am;
    const DiffWorld W = a.worlds[blockIdx.x];
    Own o;
    const size_t ck = (size_t)seg * a.nb_total;
    load_own(o, a, W, a.ck_p + ck, a.ck_q + ck, a.ck_v + ck);
    publish(c, o);
    __syncthreads();
    // recompute the segment from its checkpoint, recording the tape
    const uint32_t nsub = seg_steps * a.substeps;
    for (uint32_t k = 0; k < nsub; ++k)
        substep_forward(c, o, a, W, first_step * a.substeps + k, true, k);
    float2 lp[kDiffPer], lq[kDiffPer], lv[kDiffPer], gv[kDiffPer];
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j)  // the loss does not read the final velocities
        lv[j] = (o.live[j] && !last) ? a.lam_v[o.gi[j]] : make_float2(0.0f, 0.0f);
    if (last && !a.dl_dfinal) {  // d |centroid - target|^2 / d p_i = 2 (centroid - target) / n
        const double2 cen = block_centroid(c, o, W);
        const float gx = (float)(2.0 * (cen.x - (double)W.tx) / (double)(W.nb > 0 ? W.nb : 1));
        const float gy = (float)(2.0 * (cen.y - (double)W.ty) / (double)(W.nb > 0 ? W.nb : 1));
#pragma unroll
        for (uint32_t j = 0; j < kDiffPer; ++j) {
            lp[j] = make_float2(gx, gy);
            lq[j] = make_float2(0.0f, 0.0f);
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kDiffPer; ++j) {
            lp[j] = o.live[j] ? (last ? a.dl_dfinal[o.gi[j]] : a.lam_p[o.gi[j]]) : make_float2(0.0f, 0.0f);
            lq[j] = (o.live[j] && !last) ? a.lam_q[o.gi[j]] : make_float2(0.0f, 0.0f);
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) gv[j] = make_float2(0.0f, 0.0f);
    for (uint32_t k = nsub; k-- > 0;) substep_reverse(c, o, a, W, first_step * a.substeps + k, k, lp, lq, lv, gv);
#pragma unroll
    for (uint32_t j = 0; j < kDiffPer; ++j) {
        if (!o.live[j]) continue;
        a.lam_p[o.gi[j]] = lp[j];
        a.lam_q[o.gi[j]] = lq[j];
        a.lam_v[o.gi[j]] = lv[j];
        if (seg == 0) a.grad_vreq[o.gi[j]] = gv[j];
    }
}

}  // namespace

hipError_t launch_diff_forward(const DiffArgs& a, uint32_t n_steps, uint32_t ck_every, hipStream_t s) {
    if (a.n_worlds == 0) return hipSuccess;
    k_diff_forward<<<a.n_worlds, kDiffThreads, 0, s>>>(a, n_steps, ck_every);
    return hipGetLastError();
}

hipError_t launch_diff_backward_segment(const DiffArgs& a, uint32_t seg, uint32_t seg_first_step,
                                        uint32_t seg_steps, bool last_segment, hipStream_t s) {
    if (a.n_worlds == 0) return hipSuccess;
    k_diff_backward<<<a.n_worlds, kDiffThreads, 0, s>>>(a, seg, seg_first_step, seg_steps, last_segment ? 1u : 0u);
    return hipGetLastError();
}

}  // namespace bg
