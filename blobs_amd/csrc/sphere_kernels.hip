// sphere_kernels.hip -- the fused substep for worlds whose every body owns
// exactly one collider and that have no springs or joints (the plain large
// scene, and batches of such worlds): gravity, contacts, Verlet, collider
// transform and circle constraints of physics.rs:401-421 in ONE kernel per
// substep, over every world of a batch at once.
//
// Layout: bodies sit in broadphase-cell order in packed 16-byte records; the
// collider snapshot the contact pass reads (physics.rs:264, written by the
// previous substep's update_objects at :360-366) is ping-ponged so a
// substep can read every snapshot while writing its own.  World w owns the
// cell keys [cell_base[w], cell_base[w+1]), so sorting by key keeps its
// bodies in the fixed slot range they were uploaded to; a query never
// leaves its world's cells, so pairs never cross worlds.
//
// Broadphase.  A rebuild fits each world's grid to the bounding box of its
// snapshots over the previous period (clipped to the centroid +- 4 sigma;
// bodies outside are clamped into border cells), counts bodies per cell with atomic
// ranks, scans the counts into the cell table and scatters the snapshot
// records and the permutation.  The first substep after it gathers the body
// records through the permutation and writes them back in cell order, so
// the reordering costs no separate pass.  Between rebuilds the sorted grid
// keeps the rebuild snapshots (x0); a body whose snapshot moves more than
// `skin` from x0 (after removing its world's common motion: the gravity
// drift plus the mean velocity of its dynamic bodies at the rebuild) becomes an
// outlier and is, from then on, re-inserted every substep into a coarse
// hashed cell table with an occupancy bitmap.  A query scans the sorted grid
// with radius r + r_max + skin (skipping outliers) and the outlier table
// with radius r + r_max, so every pair of the reference predicate is
// visited exactly once from each side, whatever the bodies do; skin, cell
// size and rebuild interval change only the speed.
#include <hip/hip_ext.h>

#include "common.hpp"
#include "device_state.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

constexpr int kCoarseShift = 2;  // outlier table cells are 4x4 grid cells
constexpr int kFlatRows = 6;      // grid rows a query walks as one flat candidate list
constexpr uint32_t kFlatMax = 128;  // candidates of a flat list (two 64-bit masks)
constexpr int kContacts = 4;      // contacts summed from registers; more take a selection pass in key order
constexpr int kHitBatch = 2;      // marked candidates whose snapshot and mass load together
constexpr int kSubstepWaves = 6;  // waves per SIMD the substep kernel is compiled for
constexpr int kBatchWaves = 5;    // the same for batches (world lookups: at six it spills)
constexpr int kGatherWaves = 4;   // the same for the first substep after a rebuild (no spills)
constexpr int kSubstepThreads = 64;  // threads per block of the substep kernel (one wave)
constexpr uint32_t kStatBlocks = 1024;  // blocks of a single world's statistics pass
static_assert(kSubstepThreads >= 64 && kSubstepThreads % 64 == 0,
              "block_contacts holds one slot per 64 bodies: substep blocks are whole waves");

struct SphereSubstep {
    SphereArrays src;  // body records read (through perm on the first substep after a rebuild)
    SphereArrays dst;  // body records written, slot order
    const uint32_t* __restrict__ perm;  // new slot -> old slot (first substep after a rebuild)
    const float4* __restrict__ snap_in;
    float4* __restrict__ snap_out;
    const float2* __restrict__ x0;
    const uint32_t* __restrict__ cell_start;
    const uint32_t* __restrict__ wid;
    const uint32_t* __restrict__ cell_base;
    const GridParams* __restrict__ grids;
    float* wold_dt;
    const float2* __restrict__ wvel;
    const uint32_t* __restrict__ oh_head;  // outlier table read this substep
    const uint32_t* __restrict__ oh_next;
    const uint32_t* __restrict__ oh_bits;
    uint32_t* oh_ins_head;                 // table this substep fills (read by the next one)
    uint32_t* oh_ins_next;
    uint32_t* oh_ins_bits;
    uint32_t* oh_clr_head;                 // table cleared for the substep after next
    uint32_t* oh_clr_bits;
    const uint32_t* __restrict__ oh_wave;  // per wave: an outlier may touch its bodies this substep
    uint32_t* oh_ins_wave;                 // the same for the next substep
    uint32_t* oh_clr_wave;
    uint32_t n_waves;
    uint32_t oh_mask;
    uint32_t nb;
    StepAccum* accum;
    unsigned long long* block_contacts;
    const float2* __restrict__ con_pos;
    const float* __restrict__ con_r;
    uint32_t nk;
    DevEvent* events;   // null: no event recording
    uint32_t event_cap;
    float2* vel_out;    // where this substep's velocities go
    uint32_t* nbr;      // partner lists (two 16-bit slot deltas a word): written by the first
                        // substep after a rebuild, read by the rest
    uint8_t* nbr_len;
    const uint8_t* __restrict__ skip;  // batch: per world, the persistent step runs it (or null)
};

__device__ __forceinline__ int32_t free_cell(float x, float o, float inv) {
    float f = floorf((x - o) * inv);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int32_t)f;
}

__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, uint32_t w, uint32_t mask) {
    return (((uint32_t)cx * 0x9E3779B1u) ^ ((uint32_t)cy * 0x85EBCA77u) ^ (w * 0xC2B2AE3Du)) & mask;
}

// Rows of the sorted grid overlapping the square of half-width Q around
// (qx, qy): calls row(t0, t1) with the slot range of each row.  Cell
// coordinates are monotone in x, so the rows cover every x0 within Q.
// Cell rectangle of the square of half-width Q around (qx, qy), padded for
// rounding; false (and the whole grid) when the query is not finite.
__device__ __forceinline__ bool query_rect(const GridParams& g, float qx, float qy, float Q, int32_t& cx0,
                                           int32_t& cx1, int32_t& cy0, int32_t& cy1) {
    const float R = Q + (1e-5f * (fabsf(qx) + fabsf(qy) + Q) + 1e-6f);
    if (R < 1e30f && finitef(qx) && finitef(qy)) {
        cx0 = cell_coord(qx - R, g.ox, g.inv_cell, g.nx);
        cx1 = cell_coord(qx + R, g.ox, g.inv_cell, g.nx);
        cy0 = cell_coord(qy - R, g.oy, g.inv_cell, g.ny);
        cy1 = cell_coord(qy + R, g.oy, g.inv_cell, g.ny);
        return true;
    }
    cx0 = 0;
    cx1 = g.nx - 1;
    cy0 = 0;
    cy1 = g.ny - 1;
    return false;
}

template <class F>
__device__ __forceinline__ void grid_rows(const GridParams& g, const uint32_t* __restrict__ cell_start, uint32_t n,
                                          float qx, float qy, float Q, F&& row) {
    int32_t cx0, cx1, cy0, cy1;
    query_rect(g, qx, qy, Q, cx0, cx1, cy0, cy1);
    for (int32_t cy = cy0; cy <= cy1; ++cy) {
        const uint32_t rowc = (uint32_t)cy * (uint32_t)g.nx;
        const uint32_t t0 = cell_start[rowc + (uint32_t)cx0];
        uint32_t t1 = cell_start[rowc + (uint32_t)cx1 + 1u];
        t1 = t1 < n ? t1 : n;
        row(t0, t1);
    }
}

template <bool GROUPS, bool SENSORS, bool ZERO_OFF, bool CONSTRAINTS, bool GATHER, bool BATCHED>
__global__ __launch_bounds__(kSubstepThreads) __attribute__((amdgpu_waves_per_eu(GATHER ? kGatherWaves : (BATCHED ? kBatchWaves : kSubstepWaves)))) void k_sphere_substep(SphereSubstep k, SubstepArgs a) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kSubstepThreads + threadIdx.x;
    // clear the outlier table of the substep after next
    for (uint32_t i = blockIdx.x * kSubstepThreads + threadIdx.x; i <= k.oh_mask; i += gridDim.x * kSubstepThreads) {
        k.oh_clr_head[i] = kNil;
        if ((i & 31u) == 0u) k.oh_clr_bits[i >> 5] = 0u;
    }
    for (uint32_t i = blockIdx.x * kSubstepThreads + threadIdx.x; i < k.n_waves; i += gridDim.x * kSubstepThreads)
        k.oh_clr_wave[i] = 0u;
    // world of this slot (lanes past the end take the last body's world)
    const uint32_t w = BATCHED ? k.wid[s < k.nb ? s : k.nb - 1] : 0u;
    uint32_t new_outliers = 0, count = 0, pushes = 0, flags = 0;
    if (s < k.nb && !(BATCHED && k.skip && k.skip[w])) {
        const uint32_t g_own = GATHER ? k.perm[s] : s;
        const float4 own = k.snap_in[s];
        const uint32_t wa = __float_as_uint(own.w);
        // A single world's snapshot words carry the body flags: a steady substep reads them
        // there instead of loading the flag record.
        const bool snap_fl = !BATCHED && a.snap_flags != 0u;
        const uint32_t fl = (snap_fl && !GATHER) ? ((wa >> kSnapFlagShift) & kSnapFlagMask) : k.src.bflags[g_own];
        const float4 pp = k.src.pp[g_own];
        BodyState st;
        st.p = make_float2(pp.x, pp.y);
        st.po = make_float2(pp.z, pp.w);
        const uint32_t idm = snap_fl ? kSnapIdMask : ~kSnapOutlier;  // id bits of a snapshot word
        const uint32_t id = wa & idm;
        const bool was_out = (wa & kSnapOutlier) != 0u;
        // With every mass exactly mass_per_r times the radius and every gravity_mod 1 (both
        // checked at upload), a steady substep needs nothing from the mass record: the own
        // mass is mass_per_r times the snapshot radius, and g * 1 is g exactly.  (Batches
        // load it: measured 0.4% faster there.)
        const bool no_mg = !GATHER && !BATCHED && a.mass_per_r != 0.0f && a.gmod_one != 0u;
        const float2 mgv = no_mg ? make_float2(0.0f, 1.0f) : k.src.mg[g_own];
        // loaded now, used after the contact pass: their latency overlaps the candidate walk.
        // A body whose rotation cannot change in this substep skips it.
        const bool still = ZERO_OFF && !GATHER && !a.fresh && (fl & BF_ROT_STILL) != 0u;
        const float2 rw = still ? make_float2(0.0f, 0.0f) : k.src.rw[g_own];
        const float2 x0 = k.x0[s];
        const uint32_t nlen = GATHER ? (uint32_t)kNbrNone : (uint32_t)k.nbr_len[s];
        // an outlier of the previous substep may touch a body of this wave
        const uint32_t near_out = k.oh_wave[__builtin_amdgcn_readfirstlane(s / BG_WAVE)];
        // the first eight partner-list entries, loaded with the record (the rest on demand)
        uint32_t lw[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) lw[j] = GATHER ? 0u : k.nbr[j * k.nb + s];
        const float mown = no_mg ? a.mass_per_r * own.z : mgv.x;
        if (a.collisions && finitef(own.x) && finitef(own.y)) {
            const float2 xa = make_float2(own.x, own.y);
            const float ra = own.z;
            const GridParams g = k.grids[w];
            const uint32_t* __restrict__ cs = k.cell_start + k.cell_base[w];
            uint2 ga = make_uint2(0, 0);
            if (GROUPS) ga = k.src.groups[g_own];
            const bool sen_a = SENSORS ? (fl & BF_SENSOR) != 0u : false;
            const uint32_t nout = __hip_atomic_load(&k.accum->n_outliers, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // common motion of the world since the rebuild at the previous snapshot (the
            // same expression as the skin test that wrote it)
            const float2 vw = k.wvel[w];
            const float dpx = a.drift_prev_x + vw.x * a.t_prev, dpy = a.drift_prev_y + vw.y * a.t_prev;
            // a partner's mass is mass_per_r times the radius in its snapshot (exact: a power of
            // two times a normal value, checked for every body at upload)
            const bool by_r = a.mass_per_r != 0.0f;
            // reference predicate (physics.rs:255-269) against the slot-t snapshot; the
            // partner's mass is `mp` when the caller loaded it (have_m), else loaded on a hit
            auto test = [&](uint32_t t, const float4 sb, float mp, bool have_m, auto&& sink) {
                if (GROUPS) {
                    const uint2 gb = k.src.groups[GATHER ? k.perm[t] : t];
                    if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) return;
                }
                // axis = A - B with A the higher collider index; b - a == -(a - b) exactly,
                // so the unoriented difference decides and is flipped only for a hit
                const float dx = xa.x - sb.x, dy = xa.y - sb.y;
                const float rs = ra + sb.z;
                const float dot = dx * dx + dy * dy;
                if (dot > (rs * rs) * 1.00000095367431640625f) return;
                const float dist = sqrtf(dot);
                if (!(dist < rs)) return;
                const uint32_t idb = __float_as_uint(sb.w) & idm;
                const bool isA = id > idb;
                const float mpart = have_m ? mp
                                           : (by_r ? a.mass_per_r * sb.z : k.src.mg[GATHER ? k.perm[t] : t].x);
                sink(t, idb, isA, isA ? dx : -dx, isA ? dy : -dy, dist, rs, mpart);
            };
            // Partner lists.  Two non-outliers are each within `skin` of their rebuild
            // snapshot (after the common motion, which cancels between them), so bodies in
            // contact now were within r_a + r_b + 2 skin at the rebuild: the first substep
            // after a rebuild collects those partners from its (wider) grid walk, and the
            // later substeps test only them.  Outlier partners come from the outlier table;
            // an outlier, or a body whose list did not fit, walks the grid.
            const float list_reach = 2.0f * a.skin + (1e-4f * (ra + a.rmax + 2.0f * a.skin) + 1e-5f);
            const float Qc = ra + a.rmax + (GATHER ? list_reach : a.skin);
            const bool use_list = !GATHER && nlen <= kNbrCap && !was_out;
            // Sorted-grid candidates (first substep after a rebuild).  With at most kFlatRows
            // rows and kFlatMax candidates (the common case) they form one flat list: a cheap
            // pass marks the candidates within the squared contact distance in a bit mask, and
            // only marked ones take the exact test and the correction -- so a wave pays the
            // correction path once per hit of its busiest lane, not once per candidate.
            int32_t qcx0 = 0, qcx1 = 0, qcy0 = 0, qcy1 = 0;
            const bool rect_ok = GATHER && query_rect(g, xa.x - dpx, xa.y - dpy, Qc, qcx0, qcx1, qcy0, qcy1);
            uint32_t rlo[kFlatRows], rend[kFlatRows];  // slot range start, flat end of each row
            uint32_t n_flat = 0;
            bool flat = GATHER && rect_ok && qcy1 - qcy0 < kFlatRows;
            if (flat) {
#pragma unroll
                for (int r = 0; r < kFlatRows; ++r) {
                    const int32_t cy = qcy0 + r;
                    uint32_t lo = 0, hi = 0;
                    if (cy <= qcy1) {
                        const uint32_t rowc = (uint32_t)cy * (uint32_t)g.nx;
                        lo = cs[rowc + (uint32_t)qcx0];
                        hi = cs[rowc + (uint32_t)qcx1 + 1u];
                        hi = hi < k.nb ? hi : k.nb;
                        hi = hi > lo ? hi : lo;
                    }
                    rlo[r] = lo;
                    n_flat += hi - lo;
                    rend[r] = n_flat;
                }
                flat = n_flat <= kFlatMax;
            }
            auto flat_slot = [&](uint32_t i) {
                uint32_t t = rlo[0] + i;
#pragma unroll
                for (int r = 1; r < kFlatRows; ++r)
                    if (i >= rend[r - 1]) t = rlo[r] + (i - rend[r - 1]);
                return t;
            };
            const uint32_t nl = use_list ? nlen : 0u;
            // Batches: list words 4 to 9 (entries 8-19) of a list longer than eight entries load
            // with the first eight entries' snapshots, so their rounds skip the word load.
            constexpr bool kPreWords = BATCHED && !GATHER;
            uint32_t lw4 = 0u, lw5 = 0u, lw6 = 0u, lw7 = 0u, lw8 = 0u, lw9 = 0u;
            // slot of list entry i (< kNbrCap): the first eight (batches: twenty) from registers
            auto list_slot = [&](uint32_t i) {
                const uint32_t j = i >> 1;
                const uint32_t a0 = (j & 1u) ? lw[1] : lw[0], a1 = (j & 1u) ? lw[3] : lw[2];
                const uint32_t wd = j < 4u ? ((j & 2u) ? a1 : a0)
                                           : (kPreWords && j == 4u ? lw4
                                              : (kPreWords && j == 5u ? lw5
                                              : (kPreWords && j == 6u ? lw6
                                              : (kPreWords && j == 7u ? lw7
                                              : (kPreWords && j == 8u ? lw8
                                              : (kPreWords && j == 9u ? lw9 : k.nbr[j * k.nb + s]))))));
                const int32_t dt = (int32_t)(int16_t)(uint16_t)((i & 1u) ? (wd >> 16) : (wd & 0xFFFFu));
                const uint32_t t = (uint32_t)((int32_t)s + dt);
                return t < k.nb ? t : k.nb - 1u;
            };
            // candidates (flat walk or partner list) within the squared contact distance
            uint64_t near = 0, near_hi = 0;  // candidates 0..63 and 64..127
            uint64_t lst = 0, lst_hi = 0;    // flat candidates within the list reach (first substep after a rebuild)
            if (flat) {
                for (uint32_t f = 0; f < n_flat; f += 4) {
                    uint32_t tt[4];
                    float4 sb[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        tt[u] = flat_slot(min(f + (uint32_t)u, n_flat - 1));
                        sb[u] = k.snap_in[tt[u]];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float dx = xa.x - sb[u].x, dy = xa.y - sb[u].y;
                        const float rs = ra + sb[u].z;
                        const float d2 = dx * dx + dy * dy;
                        const bool base = f + u < n_flat && tt[u] != s && !(__float_as_uint(sb[u].w) & kSnapOutlier);
                        const bool cand = base && !(d2 > (rs * rs) * 1.00000095367431640625f);
                        const float rl = rs + list_reach;
                        const bool in_list = base && d2 <= rl * rl;
                        const uint32_t b = f + (uint32_t)u;
                        const uint64_t bit = 1ull << (b & 63u);
                        if (b < 64u) {
                            near |= cand ? bit : 0ull;
                            lst |= in_list ? bit : 0ull;
                        } else {
                            near_hi |= cand ? bit : 0ull;
                            lst_hi |= in_list ? bit : 0ull;
                        }
                    }
                }
            } else if (!GATHER) {
                // The first eight entries (the first four words, already in registers): their
                // snapshot loads issue together, without a branch between them.  An entry past
                // the list loads the body's own snapshot (coalesced) and is masked out.
                {
                    uint32_t tt[8];
                    float4 sb[8];
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) tt[u] = u < nl ? list_slot(u) : s;
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) sb[u] = k.snap_in[tt[u]];
                    if (kPreWords) {
                        lw4 = nl > 8u ? k.nbr[4u * k.nb + s] : 0u;
                        lw5 = nl > 10u ? k.nbr[5u * k.nb + s] : 0u;
                        lw6 = nl > 12u ? k.nbr[6u * k.nb + s] : 0u;
                        lw7 = nl > 14u ? k.nbr[7u * k.nb + s] : 0u;
                        lw8 = nl > 16u ? k.nbr[8u * k.nb + s] : 0u;
                        lw9 = nl > 18u ? k.nbr[9u * k.nb + s] : 0u;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) {
                        const float dx = xa.x - sb[u].x, dy = xa.y - sb[u].y;
                        const float rs = ra + sb[u].z;
                        const bool cand = u < nl && !(__float_as_uint(sb[u].w) & kSnapOutlier) &&
                                          !(dx * dx + dy * dy > (rs * rs) * 1.00000095367431640625f);
                        near |= cand ? (1ull << u) : 0ull;
                    }
                }
                // the rest, four a round
#pragma unroll
                for (uint32_t f = 8; f < kNbrCap; f += 4) {
                    if (f < nl) {
                        uint32_t tt[4];
                        float4 sb[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4; ++u) {
                            tt[u] = f + u < nl ? list_slot(f + u) : s;
                            sb[u] = k.snap_in[tt[u]];
                        }
#pragma unroll
                        for (uint32_t u = 0; u < 4; ++u) {
                            const float dx = xa.x - sb[u].x, dy = xa.y - sb[u].y;
                            const float rs = ra + sb[u].z;
                            const bool cand = f + u < nl && !(__float_as_uint(sb[u].w) & kSnapOutlier) &&
                                              !(dx * dx + dy * dy > (rs * rs) * 1.00000095367431640625f);
                            near |= cand ? (1ull << (f + u)) : 0ull;
                        }
                    }
                }
            }
            if (GATHER) {
                // this body's partner list for the substeps up to the next rebuild
                uint32_t n_l = flat ? (uint32_t)(__popcll(lst) + __popcll(lst_hi)) : (uint32_t)kNbrNone;
                if (n_l > kNbrCap) n_l = kNbrNone;
                if (n_l != kNbrNone) {
                    uint32_t q = 0, word = 0;
                    uint64_t m = lst, m2 = lst_hi;
                    while ((m | m2) != 0) {
                        const uint32_t i = m ? (uint32_t)__builtin_ctzll(m) : 64u + (uint32_t)__builtin_ctzll(m2);
                        if (m) {
                            m &= m - 1;
                        } else {
                            m2 &= m2 - 1;
                        }
                        const int32_t dt = (int32_t)flat_slot(i) - (int32_t)s;
                        if (dt < -32768 || dt > 32767) {
                            n_l = kNbrNone;
                            break;
                        }
                        const uint32_t h16 = (uint32_t)(uint16_t)(int16_t)dt;
                        if (q & 1u) {
                            k.nbr[(q >> 1) * k.nb + s] = word | (h16 << 16);
                        } else {
                            word = h16;
                        }
                        ++q;
                    }
                    if (n_l != kNbrNone && (q & 1u)) k.nbr[(q >> 1) * k.nb + s] = word;
                }
                k.nbr_len[s] = (uint8_t)n_l;
            }
            auto enumerate = [&](auto&& sink) {
                if (flat || use_list) {
                    uint64_t m = near, m2 = near_hi;
                    // marked candidates kHitBatch at a time: their snapshots (and masses, unless
                    // they follow from the radius) load together
                    while ((m | m2) != 0) {
                        uint32_t tt[kHitBatch];
                        bool ok[kHitBatch];
#pragma unroll
                        for (int u = 0; u < kHitBatch; ++u) {
                            ok[u] = (m | m2) != 0;
                            const uint32_t i = m ? (uint32_t)__builtin_ctzll(m) : (m2 ? 64u + (uint32_t)__builtin_ctzll(m2) : 0u);
                            if (m) {
                                m &= m - 1;
                            } else {
                                m2 &= m2 - 1;
                            }
                            tt[u] = ok[u] ? (GATHER ? flat_slot(i) : list_slot(i)) : s;
                        }
                        float4 sb[kHitBatch];
                        float mp[kHitBatch];
#pragma unroll
                        for (int u = 0; u < kHitBatch; ++u) {
                            sb[u] = k.snap_in[tt[u]];
                            mp[u] = by_r ? 0.0f : k.src.mg[GATHER ? k.perm[tt[u]] : tt[u]].x;
                        }
#pragma unroll
                        for (int u = 0; u < kHitBatch; ++u)
                            if (ok[u]) test(tt[u], sb[u], by_r ? a.mass_per_r * sb[u].z : mp[u], true, sink);
                    }
                } else {
                    grid_rows(g, cs, k.nb, xa.x - dpx, xa.y - dpy, Qc,
                              [&](uint32_t t0, uint32_t t1) {
                                  for (uint32_t t = t0; t < t1; ++t) {
                                      const float4 sb = k.snap_in[t];
                                      if (t != s && !(__float_as_uint(sb.w) & kSnapOutlier))
                                          test(t, sb, 0.0f, false, sink);
                                  }
                              });
                }
                // A non-outlier skips the table unless its wave was flagged: every outlier
                // flags the waves holding bodies it could touch (see its insertion below).
                if (nout == 0u || (near_out == 0u && !was_out)) return;
                // outlier table: coarse cells of this world around the current snapshot
                float R = ra + a.rmax;
                R = R + (1e-5f * (fabsf(xa.x) + fabsf(xa.y) + R) + 1e-6f);
                const int32_t cx0 = free_cell(xa.x - R, g.ox, g.inv_cell) >> kCoarseShift;
                const int32_t cx1 = free_cell(xa.x + R, g.ox, g.inv_cell) >> kCoarseShift;
                const int32_t cy0 = free_cell(xa.y - R, g.oy, g.inv_cell) >> kCoarseShift;
                const int32_t cy1 = free_cell(xa.y + R, g.oy, g.inv_cell) >> kCoarseShift;
                if (!(R < 1e30f) || cx1 - cx0 > 64 || cy1 - cy0 > 64) {
                    for (uint32_t h = 0; h <= k.oh_mask; ++h)  // huge radius: every bucket (superset)
                        for (uint32_t t = k.oh_head[h]; t != kNil; t = k.oh_next[t])
                            if (t != s && (!BATCHED || k.wid[t] == w)) test(t, k.snap_in[t], 0.0f, false, sink);
                    return;
                }
                for (int32_t cy = cy0; cy <= cy1; ++cy)
                    for (int32_t cx = cx0; cx <= cx1; ++cx) {
                        const uint32_t h = cell_hash(cx, cy, w, k.oh_mask);
                        if (!((k.oh_bits[h >> 5] >> (h & 31u)) & 1u)) continue;
                        for (uint32_t t = k.oh_head[h]; t != kNil; t = k.oh_next[t]) {
                            if (t == s) continue;
                            if (BATCHED && k.wid[t] != w) continue;
                            const float4 sb = k.snap_in[t];
                            // each outlier lives in one coarse cell: drop hash collisions
                            if ((free_cell(sb.x, g.ox, g.inv_cell) >> kCoarseShift) != cx ||
                                (free_cell(sb.y, g.oy, g.inv_cell) >> kCoarseShift) != cy)
                                continue;
                            test(t, sb, 0.0f, false, sink);
                        }
                    }
            };
            // correction of one contact for this body's side (physics.rs:272-300);
            // a push-out is folded into the same entry (order-dependent in the
            // reference, reproduced for the pair)
            auto correction = [&](bool first, auto&& put, uint32_t t, uint32_t idb, bool isA, float ax, float ay,
                                  float dist, float rs, float mpart) {
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
                const uint32_t tp = GATHER ? k.perm[t] : t;
                const bool sen_b = SENSORS ? (k.src.bflags[tp] & BF_SENSOR) != 0u : false;
                if (sen_a || sen_b) {
                    if (pushed) put(idb, px, py);
                    return;
                }
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
            ContactList<kContacts, uint32_t> L;
            L.init();
            enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs, float mp) {
                if (isA) ++count;
                correction(true, [&](uint32_t kk, float x, float y) { L.add(kk, x, y); }, t, idb, isA, ax, ay, dist,
                           rs, mp);
            });
            if (k.events && count > 0) {
                // collision events of the pairs this body is A in, with both bodies'
                // velocities before the contact pass (physics.rs:288-311)
                unsigned long long e = atomicAdd(&k.accum->event_count, (unsigned long long)count);
                const float2 va = k.src.vel[g_own];
                enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs, float) {
                    if (!isA) return;
                    if (e < k.event_cap) {
                        DevEvent ev;
                        ev.sub = a.sub_global;
                        ev.a = id;
                        ev.b = idb;
                        ev.pad = 0;
                        ev.va = va;
                        ev.vb = k.src.vel[GATHER ? k.perm[t] : t];
                        k.events[e] = ev;
                    }
                    ++e;
                });
            }
            if (L.n > 0) {
                uint32_t last = L.apply(st.p);
                uint32_t done = L.n < (uint32_t)kContacts ? L.n : (uint32_t)kContacts;
                while (done < L.n) {  // more than kContacts: next key by selection, same order
                    uint32_t best = kNil;
                    float bx = 0.0f, by = 0.0f;
                    enumerate([&](uint32_t t, uint32_t idb, bool isA, float ax, float ay, float dist, float rs,
                                  float mp) {
                        correction(
                            false,
                            [&](uint32_t kk, float x, float y) {
                                if (kk > last && kk < best) {
                                    best = kk;
                                    bx = x;
                                    by = y;
                                }
                            },
                            t, idb, isA, ax, ay, dist, rs, mp);
                    });
                    if (best == kNil) break;
                    st.p.x = st.p.x + bx;
                    st.p.y = st.p.y + by;
                    last = best;
                    ++done;
                }
            }
        } else if (GATHER) {
            k.nbr_len[s] = kNbrNone;
        }
        // update_objects (physics.rs:326-358)
        st.rot = rw.x;
        st.omega = rw.y;
        float2 vel;
        uint32_t fl_out = fl;
        float tq_out = 0.0f;
        if (fl & BF_STATIC) {
            st.po = st.p;
            vel = make_float2(0.0f, 0.0f);
            if (GATHER) tq_out = k.src.torque[g_own];
        } else {
            float2 acc = a.fresh ? k.src.acc[g_own] : make_float2(0.0f, 0.0f);
            acc.x = acc.x + a.gx * mgv.y;
            acc.y = acc.y + a.gy * mgv.y;
            const bool vr = (fl & BF_HAS_VREQ) != 0u;
            const float2 vreq = vr ? k.src.vreq[g_own] : make_float2(0.0f, 0.0f);
            const float tq = a.fresh ? k.src.torque[g_own] : 0.0f;
            const float toi = (a.fresh || !(fl & BF_INERTIA_POS)) ? tq / k.src.inertia[g_own] : 0.0f;
            float ratio = a.ratio_rest;
            if (fl & BF_FIRST_DYN) {  // Q1: this world's old_dt
                ratio = a.h / k.wold_dt[w];
                k.wold_dt[w] = a.h;
            }
            verlet_dynamic(st, acc, toi, vr, vreq, a.h, ratio);
            vel = st.vel;
            if (vr) fl_out = fl & ~BF_HAS_VREQ;
        }
        if (ZERO_OFF && !still) {
            const bool now_still = (fl & BF_STATIC) != 0u ||
                                   (__float_as_uint(st.omega) == 0u && __float_as_uint(st.rot) != 0x80000000u &&
                                    finitef(st.rot) && (fl & BF_INERTIA_POS) != 0u);
            fl_out = now_still ? (fl_out | BF_ROT_STILL) : (fl_out & ~BF_ROT_STILL);
        }
        if (GATHER) {
            // first substep after a rebuild: every record moves to its new slot
            k.dst.rw[s] = make_float2(st.rot, st.omega);
            k.dst.mg[s] = mgv;
            k.dst.bflags[s] = fl_out;
            k.dst.inertia[s] = k.src.inertia[g_own];
            k.dst.torque[s] = tq_out;
            k.dst.off[s] = k.src.off[g_own];
            if (GROUPS) k.dst.groups[s] = k.src.groups[g_own];
            if (fl_out & BF_HAS_VREQ) k.dst.vreq[s] = k.src.vreq[g_own];
        } else {
            if (!still && (__float_as_uint(st.rot) != __float_as_uint(rw.x) ||
                           __float_as_uint(st.omega) != __float_as_uint(rw.y)))
                k.dst.rw[s] = make_float2(st.rot, st.omega);
            if (fl_out != fl) k.dst.bflags[s] = fl_out;
            if (a.fresh && !(fl & BF_STATIC)) k.dst.torque[s] = 0.0f;
        }
        if (a.write_vel) k.vel_out[s] = vel;
        // collider transform from the pre-constraint position (physics.rs:360-366)
        // (a still body's rotation is needed only for a position with a zero coordinate)
        const float rot_x = (still && !(st.p.x != 0.0f && st.p.y != 0.0f)) ? k.src.rw[g_own].x : st.rot;
        const float2 x = ZERO_OFF ? body_xform_zero_off(rot_x, st.p) : body_xform_slow(st.rot, st.p, k.src.off[g_own]);
        const bool finite_x = finitef(x.x) && finitef(x.y);
        bool out = was_out;
        if (a.collisions && !out && finite_x) {
            const float2 vw = k.wvel[w];
            const float dx = a.drift_x + vw.x * a.t_cur, dy = a.drift_y + vw.y * a.t_cur;
            const float ex = (x.x - x0.x) - dx, ey = (x.y - x0.y) - dy;
            out = !(ex * ex + ey * ey <= a.skin * a.skin);
            new_outliers += out ? 1u : 0u;
        }
        const uint32_t fl_bits = snap_fl ? (fl_out & kSnapFlagMask) << kSnapFlagShift : 0u;
        k.snap_out[s] = make_float4(x.x, x.y, own.z, __uint_as_float(id | fl_bits | (out ? kSnapOutlier : 0u)));
        if (out && finite_x) {
            // outlier table of the next substep
            const GridParams g = k.grids[w];
            const uint32_t h = cell_hash(free_cell(x.x, g.ox, g.inv_cell) >> kCoarseShift,
                                         free_cell(x.y, g.oy, g.inv_cell) >> kCoarseShift, w, k.oh_mask);
            const uint32_t old = atomicExch(&k.oh_ins_head[h], s);
            k.oh_ins_next[s] = old;
            atomicOr(&k.oh_ins_bits[h >> 5], 1u << (h & 31u));
            // Flag the waves whose bodies this outlier could touch next substep.  A body
            // in contact is within 2 r_max of it; a non-outlier is within `skin` of its
            // rebuild position plus the world's common motion, so its rebuild cell lies
            // within 2 r_max + skin of this position with that motion removed (the skin
            // test's expression) -- and the grid holds bodies by rebuild cell.
            const float2 vw = k.wvel[w];
            const float ddx = a.drift_x + vw.x * a.t_cur, ddy = a.drift_y + vw.y * a.t_cur;
            int32_t fx0, fx1, fy0, fy1;
            query_rect(g, x.x - ddx, x.y - ddy, 2.0f * a.rmax + a.skin, fx0, fx1, fy0, fy1);
            const uint32_t* __restrict__ csw = k.cell_start + k.cell_base[w];
            for (int32_t cy = fy0; cy <= fy1; ++cy) {
                const uint32_t rowc = (uint32_t)cy * (uint32_t)g.nx;
                const uint32_t t0 = csw[rowc + (uint32_t)fx0];
                uint32_t t1 = csw[rowc + (uint32_t)fx1 + 1u];
                t1 = t1 < k.nb ? t1 : k.nb;
                for (uint32_t wv = t0 / BG_WAVE; t0 < t1 && wv <= (t1 - 1u) / BG_WAVE; ++wv) k.oh_ins_wave[wv] = 1u;
            }
        }
        if (CONSTRAINTS) apply_constraints(st.p, k.con_pos, k.con_r, k.nk);
        k.dst.pp[s] = make_float4(st.p.x, st.p.y, st.po.x, st.po.y);
    }
    // counters: two wave sums, then one atomic per wave and quantity that is non-zero
    // (contacts into this block's slot, summed on the host when queried)
    if (flags) atomicOr(&k.accum->error_flags, flags);
    const uint32_t c_sum = wave_sum_u32(count);
    const uint32_t po = wave_sum_u32((pushes << 16) | new_outliers);
    if ((threadIdx.x & (BG_WAVE - 1)) == 0) {
        if (c_sum) atomicAdd(&k.block_contacts[blockIdx.x], (unsigned long long)c_sum);
        if (po >> 16) atomicAdd(&k.accum->push_outs, po >> 16);
        if (po & 0xFFFFu) atomicAdd(&k.accum->n_outliers, po & 0xFFFFu);
    }
}

// ---- rebuild: one-pass counting sort by cell ------------------------------

// Per-world sums over the current snapshot for the next grid and drift:
// [0,1] velocity of dynamic bodies, [2] their count, [3,4] position,
// [5,6] squared position, [7] count, and the bounding box (ordered-int
// atomics).  Each block reduces in LDS and adds
// into one of `rep` replicas of its world's sums (chosen by block index),
// so one world's blocks do not all hit the same words.  Float atomics: the sums' rounding
// order varies, which moves grid bounds and the drift model only -- never
// a result (any grid and drift give the same pair set).
template <bool BATCHED>
__global__ __launch_bounds__(kThreads) void k_world_stats(const float4* __restrict__ snap,
                                                          const float2* __restrict__ vel,
                                                          const uint32_t* __restrict__ bflags, uint32_t nb,
                                                          const uint32_t* __restrict__ wid, float* wstat,
                                                          uint32_t rep, uint32_t* wbbox,
                                                          const uint8_t* __restrict__ skip) {
    __shared__ float s_part[kThreads / BG_WAVE][8];
    __shared__ uint32_t s_bb[kThreads / BG_WAVE][4];
    __shared__ uint32_t s_w[kThreads / BG_WAVE];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t w = BATCHED ? wid[i < nb ? i : nb - 1] : 0u;
    // slots of worlds the persistent step runs are not the grid path's
    const bool live = i < nb && !(BATCHED && skip && skip[w]);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = 0.0f;
    uint32_t b0 = 0xFFFFFFFFu, b1 = 0xFFFFFFFFu, b2 = 0u, b3 = 0u;  // f2ord bounding box
    auto add = [&](uint32_t j) {
        const float4 sn = snap[j];
        if (finitef(sn.x) && finitef(sn.y)) {
            const uint32_t ox = f2ord(sn.x), oy = f2ord(sn.y);
            b0 = min(b0, ox);
            b2 = max(b2, ox);
            b1 = min(b1, oy);
            b3 = max(b3, oy);
            if (!(bflags[j] & BF_STATIC)) {
                const float2 vv = vel[j];
                if (finitef(vv.x) && finitef(vv.y)) {
                    v[0] += vv.x;
                    v[1] += vv.y;
                    v[2] += 1.0f;
                }
            }
            v[3] += sn.x;
            v[4] += sn.y;
            v[5] += sn.x * sn.x;
            v[6] += sn.y * sn.y;
            v[7] += 1.0f;
        }
    };
    if (BATCHED) {
        if (live) add(i);
    } else {
        // one world: a grid-stride pass (a few bodies a thread), so fewer blocks add atomics
        for (uint32_t j = i; j < nb; j += gridDim.x * blockDim.x) add(j);
    }
    const int lane = threadIdx.x & (BG_WAVE - 1), wv = threadIdx.x / BG_WAVE;
    const uint32_t w0 = __shfl(w, 0, BG_WAVE);
    const bool wave_uniform = __all(w == w0);
    if (!wave_uniform) {
        // a wave across worlds (small worlds): one wave reduction per world in it, and
        // one lane issues that world's update
        uint64_t left = __ballot(1);
        while (left != 0) {
            const int l0 = __builtin_ctzll(left);
            const uint32_t wc = __shfl(w, l0, BG_WAVE);
            const bool mine = w == wc;
            float r[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                r[t] = mine ? v[t] : 0.0f;
                for (int o = 32; o > 0; o >>= 1) r[t] += __shfl_xor(r[t], o, BG_WAVE);
            }
            const uint32_t c0 = wave_min_u32(mine ? b0 : 0xFFFFFFFFu), c1 = wave_min_u32(mine ? b1 : 0xFFFFFFFFu);
            const uint32_t c2 = wave_max_u32(mine ? b2 : 0u), c3 = wave_max_u32(mine ? b3 : 0u);
            if (lane == l0 && r[7] > 0.0f) {
                float* dc = wstat + ((size_t)wc * rep + (blockIdx.x & (rep - 1))) * 8;
                for (int t = 0; t < 8; ++t) atomicAdd(&dc[t], r[t]);
                uint32_t* bb = wbbox + ((size_t)wc * rep + (blockIdx.x & (rep - 1))) * 4;
                atomic_min_if_less(&bb[0], c0);
                atomic_min_if_less(&bb[1], c1);
                atomic_max_if_greater(&bb[2], c2);
                atomic_max_if_greater(&bb[3], c3);
            }
            left &= ~__ballot(mine);
        }
    } else {
#pragma unroll
        for (int t = 0; t < 8; ++t)
            for (int o = 32; o > 0; o >>= 1) v[t] += __shfl_xor(v[t], o, BG_WAVE);
        b0 = wave_min_u32(b0);
        b1 = wave_min_u32(b1);
        b2 = wave_max_u32(b2);
        b3 = wave_max_u32(b3);
    }
    if (lane == 0) {
        s_w[wv] = wave_uniform ? w : kNil;
        for (int t = 0; t < 8; ++t) s_part[wv][t] = v[t];
        s_bb[wv][0] = b0;
        s_bb[wv][1] = b1;
        s_bb[wv][2] = b2;
        s_bb[wv][3] = b3;
    }
    __syncthreads();
    // the block's waves: those sharing the first wave's world are combined into one
    // update, each other uniform wave updates its own world; lane q of the first wave
    // issues quantity q (12 quantities: 8 sums, 4 box bounds) so the atomics overlap
    if (threadIdx.x < 12) {
        const int q = threadIdx.x;
        const uint32_t rsel = blockIdx.x & (rep - 1);
        const uint32_t wb = s_w[0];
        float facc = 0.0f;
        uint32_t uacc = q < 10 ? 0xFFFFFFFFu : 0u;
        bool have = false;
        for (int k = 0; k < kThreads / BG_WAVE; ++k) {
            if (s_w[k] == kNil || !(s_part[k][7] > 0.0f)) continue;
            if (s_w[k] == wb) {
                have = true;
                if (q < 8) facc += s_part[k][q];
                else if (q < 10) uacc = min(uacc, s_bb[k][q - 8]);
                else uacc = max(uacc, s_bb[k][q - 8]);
            } else {
                const uint32_t wk = s_w[k];
                if (q < 8) atomicAdd(&wstat[((size_t)wk * rep + rsel) * 8 + q], s_part[k][q]);
                else if (q < 10) atomic_min_if_less(&wbbox[((size_t)wk * rep + rsel) * 4 + (q - 8)], s_bb[k][q - 8]);
                else atomic_max_if_greater(&wbbox[((size_t)wk * rep + rsel) * 4 + (q - 8)], s_bb[k][q - 8]);
            }
        }
        if (have) {
            if (q < 8) atomicAdd(&wstat[((size_t)wb * rep + rsel) * 8 + q], facc);
            else if (q < 10) atomic_min_if_less(&wbbox[((size_t)wb * rep + rsel) * 4 + (q - 8)], uacc);
            else atomic_max_if_greater(&wbbox[((size_t)wb * rep + rsel) * 4 + (q - 8)], uacc);
        }
    }
}

// Grid of every world over its snapshots: their bounding box clipped to the
// centroid +- 4 sigma so a few far-flung bodies (clamped into border cells)
// do not coarsen the cells of the rest; the mean velocity of its dynamic
// bodies (drift model).
// One wave per world (the replica sums are reduced across its lanes).
__global__ __launch_bounds__(BG_WAVE) void k_world_grids(uint32_t n_worlds, uint32_t* wbbox, float cell, const uint32_t* __restrict__ cell_base,
                                                         const float* __restrict__ wstat, uint32_t rep,
                                                         GridParams* grids, float2* wvel, StepAccum* accum) {
    const uint32_t w = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (w == 0 && lane == 0) {
        accum->dmax = 0u;
        accum->n_outliers = 0u;
    }
    if (w >= n_worlds) return;
    float st[8];
    for (int t = 0; t < 8; ++t) st[t] = 0.0f;
    uint32_t bb[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u};
    for (uint32_t r = lane; r < rep; r += BG_WAVE) {
        for (int t = 0; t < 8; ++t) st[t] += wstat[((size_t)w * rep + r) * 8 + t];
        uint32_t* src = wbbox + ((size_t)w * rep + r) * 4;
        bb[0] = min(bb[0], src[0]);
        bb[1] = min(bb[1], src[1]);
        bb[2] = max(bb[2], src[2]);
        bb[3] = max(bb[3], src[3]);
        src[0] = src[1] = 0xFFFFFFFFu;  // empty again for the next rebuild's statistics
        src[2] = src[3] = 0u;
    }
    for (int t = 0; t < 8; ++t)
        for (int o = 32; o > 0; o >>= 1) st[t] += __shfl_xor(st[t], o, BG_WAVE);
    bb[0] = wave_min_u32(bb[0]);
    bb[1] = wave_min_u32(bb[1]);
    bb[2] = wave_max_u32(bb[2]);
    bb[3] = wave_max_u32(bb[3]);
    if (lane != 0) return;
    if (st[7] > 0.0f && bb[0] <= bb[2] && bb[1] <= bb[3]) {
        const float n = st[7];
        const float mx = st[3] / n, my = st[4] / n;
        const float sx = sqrtf(fmaxf(st[5] / n - mx * mx, 0.0f)), sy = sqrtf(fmaxf(st[6] / n - my * my, 0.0f));
        const float mg = 2.0f * cell;
        if (finitef(mx) && finitef(sx)) {
            bb[0] = max(bb[0], f2ord(mx - 4.0f * sx - mg));
            bb[2] = min(bb[2], f2ord(mx + 4.0f * sx + mg));
        }
        if (finitef(my) && finitef(sy)) {
            bb[1] = max(bb[1], f2ord(my - 4.0f * sy - mg));
            bb[3] = min(bb[3], f2ord(my + 4.0f * sy + mg));
        }
        if (bb[0] > bb[2]) bb[0] = bb[2];
        if (bb[1] > bb[3]) bb[1] = bb[3];
    }
    grids[w] = grid_from_bbox(bb, cell, cell_base[w + 1] - cell_base[w]);
    wvel[w] = st[2] > 0.0f ? make_float2(st[0] / st[2], st[1] / st[2]) : make_float2(0.0f, 0.0f);
}

// Cell key of every snapshot and its rank inside the cell.
template <bool BATCHED>
__global__ __launch_bounds__(kThreads) void k_sphere_count(const float4* __restrict__ snap, uint32_t nb,
                                                           const uint32_t* __restrict__ wid,
                                                           const GridParams* __restrict__ grids,
                                                           const uint32_t* __restrict__ cell_base,
                                                           uint32_t* __restrict__ cnt, uint32_t* __restrict__ key,
                                                           uint32_t* __restrict__ rank,
                                                           const uint8_t* __restrict__ skip) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < nb;
    uint32_t kk = kNil;
    if (live) {
        const uint32_t w = BATCHED ? wid[i] : 0u;
        if (BATCHED && skip && skip[w]) {
            // a world the persistent step runs: counted in its first cell so that the scan
            // keeps its slot range (the scatter leaves its slots alone)
            kk = cell_base[w];
        } else {
            const GridParams g = grids[w];
            const float4 sn = snap[i];
            const int32_t cx = cell_coord(sn.x, g.ox, g.inv_cell, g.nx);
            const int32_t cy = cell_coord(sn.y, g.oy, g.inv_cell, g.ny);
            kk = cell_base[w] + (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
        }
        key[i] = kk;
    }
    // Slots are still in the previous rebuild's cell order and bodies move little, so
    // neighbouring lanes often share a cell: each run of equal keys takes its ranks with
    // one atomic (the run's first lane), the others offset from it.  Order inside a cell
    // is arbitrary.  A lane past the end carries kNil, which starts a run of its own.
    const uint32_t lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t prev = __shfl_up(kk, 1, BG_WAVE);
    const bool head = lane == 0u || prev != kk;
    const uint64_t heads = __ballot(head);
    const uint64_t le = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);
    const uint32_t my_head = 63u - (uint32_t)__builtin_clzll((heads & le) | 1ull);  // this lane's run start
    const uint64_t above = heads & ~le;  // heads after this lane
    const uint32_t run_end = above ? (uint32_t)__builtin_ctzll(above) : 64u;
    uint32_t base = 0;
    if (head && live) base = atomicAdd(&cnt[kk], run_end - lane);
    base = __shfl(base, (int)my_head, BG_WAVE);
    if (live) rank[i] = base + (lane - my_head);
}

struct SphereScatter {
    const float4* __restrict__ snap;
    float4* __restrict__ snap_sorted;
    float2* __restrict__ x0;
    uint32_t* __restrict__ perm;
    const uint32_t* __restrict__ key;
    const uint32_t* __restrict__ rank;
    const uint32_t* __restrict__ cell_start;
    uint32_t* cnt;
    uint32_t* heads[3];
    uint32_t* bits[3];
    uint32_t* waves[3];
    uint32_t n_waves;
    uint32_t oh_size;
    uint32_t nb;
    float* wstat;
    uint32_t n_wstat;
    const uint32_t* __restrict__ wid;  // batch only
    const uint8_t* __restrict__ skip;  // per world: the persistent step runs it (or null)
};

// Each snapshot to slot cell_start[key] + rank; the permutation records
// where its body record is; x0 = the new snapshot; no body is an outlier;
// counters and outlier tables are reset for the next period.
__global__ __launch_bounds__(kThreads) void k_sphere_scatter(SphereScatter k) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t c = i; c < k.n_wstat; c += stride) k.wstat[c] = 0.0f;
    for (uint32_t h = i; h < 3u * k.oh_size; h += stride) {
        const uint32_t t = h / k.oh_size, e = h % k.oh_size;
        k.heads[t][e] = kNil;
        if ((e & 31u) == 0u) k.bits[t][e >> 5] = 0u;
    }
    for (uint32_t h = i; h < 3u * k.n_waves; h += stride) k.waves[h / k.n_waves][h % k.n_waves] = 0u;
    if (i >= k.nb) return;
    // every body counted itself in its key's cell: clearing those cells leaves every count
    // zero for the next rebuild (the scan has read them)
    const uint32_t key = k.key[i];
    k.cnt[key] = 0u;
    if (k.skip && k.skip[k.wid[i]]) return;
    const uint32_t slot = k.cell_start[key] + k.rank[i];
    if (slot >= k.nb) return;  // consistent counts and scan never get here
    float4 sn = k.snap[i];
    sn.w = __uint_as_float(__float_as_uint(sn.w) & ~kSnapOutlier);
    k.snap_sorted[slot] = sn;
    k.x0[slot] = make_float2(sn.x, sn.y);
    k.perm[slot] = i;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill_u32<<<(uint32_t)blocks, 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_sphere_rebuild(DeviceWorld& d, float cell, uint32_t parity) {
    if (d.nb == 0) return hipSuccess;
    const uint32_t blocks = div_up(d.nb, kThreads);
    if (d.n_worlds > 1)
        k_world_stats<true><<<blocks, kThreads, 0, d.stream>>>(d.s.snap, d.s.vel, d.s.bflags, d.nb, d.wid, d.wstat,
                                                               d.stat_rep, d.wbbox, d.p_skip);
    else
        k_world_stats<false><<<std::min(blocks, kStatBlocks), kThreads, 0, d.stream>>>(
            d.s.snap, d.s.vel, d.s.bflags, d.nb, d.wid, d.wstat, d.stat_rep, d.wbbox, nullptr);
    k_world_grids<<<d.n_worlds, BG_WAVE, 0, d.stream>>>(d.n_worlds, d.wbbox, cell,
                                                                           d.cell_base, d.wstat, d.stat_rep,
                                                                           d.grids, d.wvel, d.accum);
    if (d.n_worlds > 1)
        k_sphere_count<true><<<blocks, kThreads, 0, d.stream>>>(d.s.snap, d.nb, d.wid, d.grids, d.cell_base,
                                                                d.cell_count, d.keys, d.vals, d.p_skip);
    else
        k_sphere_count<false><<<blocks, kThreads, 0, d.stream>>>(d.s.snap, d.nb, d.wid, d.grids, d.cell_base,
                                                                 d.cell_count, d.keys, d.vals, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (d.n_worlds > 1) {
        // a batch's cell table: only the cells each world's new grid uses
        ScanSegments g;
        g.tile_seg = d.tile_seg;
        g.seg_base = d.cell_base;
        g.seg_used = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(d.grids) +
                                                       offsetof(GridParams, ncells));
        static_assert(sizeof(GridParams) % sizeof(uint32_t) == 0, "GridParams holds 32-bit fields");
        g.seg_used_stride = sizeof(GridParams) / sizeof(uint32_t);
        g.n_seg = d.n_worlds;
        e = exclusive_scan_u32_segments(d.cell_count, d.cell_start, d.total_cells + 1, d.sort, d.stream, g);
    } else {
        // one world: the cells its new grid uses and the end sentinel (its count read on the device)
        const uint32_t* last =
            reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(d.grids) + offsetof(GridParams, ncells));
        e = exclusive_scan_u32(d.cell_count, d.cell_start, d.total_cells + 1, d.sort, d.stream, last);
    }
    if (e != hipSuccess) return e;
    SphereScatter k;
    k.snap = d.s.snap;
    k.snap_sorted = d.snap_next;
    k.x0 = d.x0;
    k.perm = d.perm;
    k.key = d.keys;
    k.rank = d.vals;
    k.cell_start = d.cell_start;
    k.cnt = d.cell_count;
    for (int t = 0; t < 3; ++t) {
        k.heads[t] = d.oh_head[t];
        k.bits[t] = d.oh_bits[t];
        k.waves[t] = d.oh_wave[t];
    }
    k.oh_size = d.oh_size;
    k.n_waves = d.n_waves;
    k.nb = d.nb;
    k.wstat = d.wstat;
    k.n_wstat = 8u * d.stat_rep * d.n_worlds;
    k.wid = d.wid;
    k.skip = d.n_worlds > 1 ? d.p_skip : nullptr;
    k_sphere_scatter<<<blocks, kThreads, 0, d.stream>>>(k);
    e = hipGetLastError();
    // the sorted snapshot becomes the current one; records follow on the next substep
    float4* t = d.s.snap;
    d.s.snap = d.snap_next;
    d.snap_next = t;
    d.gather_pending = true;
    return e;
}

hipError_t launch_sphere_substep(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors, bool zero_off,
                                 bool constraints) {
    if (d.nb == 0) return hipSuccess;
    SphereSubstep k;
    const bool gather = d.gather_pending;
    const bool batched = d.n_worlds > 1;
    k.src = d.s;
    k.dst = gather ? d.s_alt : d.s;
    k.dst.snap = d.s.snap;  // the snapshot is ping-ponged separately
    k.perm = d.perm;
    k.snap_in = d.s.snap;
    k.snap_out = d.snap_next;
    k.x0 = d.x0;
    k.cell_start = d.cell_start;
    k.wid = d.wid;
    k.cell_base = d.cell_base;
    k.grids = d.grids;
    k.wold_dt = d.wold_dt;
    k.wvel = d.wvel;
    const uint32_t j = a.sub;
    k.oh_head = d.oh_head[j % 3];
    k.oh_next = d.oh_next[j % 3];
    k.oh_bits = d.oh_bits[j % 3];
    k.oh_ins_head = d.oh_head[(j + 1) % 3];
    k.oh_ins_next = d.oh_next[(j + 1) % 3];
    k.oh_ins_bits = d.oh_bits[(j + 1) % 3];
    k.oh_clr_head = d.oh_head[(j + 2) % 3];
    k.oh_clr_bits = d.oh_bits[(j + 2) % 3];
    k.oh_wave = d.oh_wave[j % 3];
    k.oh_ins_wave = d.oh_wave[(j + 1) % 3];
    k.oh_clr_wave = d.oh_wave[(j + 2) % 3];
    k.n_waves = d.n_waves;
    k.oh_mask = d.oh_size - 1;
    k.nb = d.nb;
    k.accum = d.accum;
    k.block_contacts = d.block_contacts;
    k.con_pos = d.con_pos;
    k.con_r = d.con_r;
    k.nk = d.nk;
    k.events = (d.events && a.collisions) ? d.events : nullptr;
    k.event_cap = d.event_cap;
    // with events the old velocities stay readable: the new ones go to vel_next
    k.vel_out = k.events ? d.vel_next : k.dst.vel;
    k.nbr = d.nbr;
    k.nbr_len = d.nbr_len;
    k.skip = d.p_skip;
    const uint32_t blocks = div_up(d.nb, kSubstepThreads);
#define BG_SPHERE(G, S, Z, C, P, B)                                                                        \
    do {                                                                                                  \
        if (d.prof_start)                                                                                 \
            hipExtLaunchKernelGGL(k_sphere_substep<G, S, Z, C, P, B>, dim3(blocks), dim3(kSubstepThreads), 0, \
                                  d.stream, d.prof_start, d.prof_stop, 0, k, a);                          \
        else                                                                                              \
            k_sphere_substep<G, S, Z, C, P, B><<<blocks, kSubstepThreads, 0, d.stream>>>(k, a);         \
    } while (0)
#define BG_SPHERE_PB(G, S, Z, C)                 \
    do {                                         \
        if (gather && batched)                   \
            BG_SPHERE(G, S, Z, C, true, true);   \
        else if (gather)                         \
            BG_SPHERE(G, S, Z, C, true, false);  \
        else if (batched)                        \
            BG_SPHERE(G, S, Z, C, false, true);  \
        else                                     \
            BG_SPHERE(G, S, Z, C, false, false); \
    } while (0)
    if (!groups && !sensors && zero_off && !constraints)
        BG_SPHERE_PB(false, false, true, false);
    else if (!groups && !sensors && !constraints)
        BG_SPHERE_PB(false, false, false, false);
    else if (!groups && !sensors)
        BG_SPHERE_PB(false, false, false, true);
    else
        BG_SPHERE_PB(true, true, false, true);
#undef BG_SPHERE_PB
#undef BG_SPHERE
    hipError_t e = hipGetLastError();
    if (gather) {
        // records now live in the alternate buffers, in cell order
        float4* snap_cur = d.s.snap;
        SphereArrays t = d.s;
        d.s = d.s_alt;
        d.s_alt = t;
        d.s.snap = snap_cur;
        d.s_alt.snap = nullptr;
        d.gather_pending = false;
    }
    if (k.events) {  // the current velocities are in vel_next now
        float2* cur = k.vel_out;
        d.vel_next = d.s.vel;
        d.s.vel = cur;
    }
    float4* t = d.s.snap;
    d.s.snap = d.snap_next;
    d.snap_next = t;
    return e;
}

}  // namespace bg
