// ordered_contacts.hip -- the push-out branch of brute_force_collisions
// (physics.rs:272-286) in the reference's exact order, on the general path.
//
// k_contacts evaluates every pair on the substep's collider snapshot, which
// is the reference's result as long as no push-out happens: the reference
// reads the snapshot, never writes it.  A push-out is the one exception.  It
// moves both bodies by -+0.01 in x and REWRITES both colliders' snapshots as
// `position + offset.translation` from the bodies' live positions (which hold
// every correction applied so far in the pass), and every later pair of the
// (i, j<i) loop reads the rewritten values.
//
// Only a pair whose snapshots coincide can start that chain: before the
// first push-out nothing has been rewritten.  k_contacts therefore records
// the rows (A colliders) of coincident pairs, and this kernel runs after it.
// With no row it exits at once.  Otherwise one thread replays, in key order
// (A, B) with A the higher collider index, every pair whose outcome can
// differ from the snapshot evaluation -- "dirty" pairs: a coincident pair, or
// a pair one of whose colliders was rewritten at a smaller key -- and logs
// their corrections.  A pair that is not dirty is "clean": its correction is
// the snapshot one k_contacts already used.
//
// Rows are replayed in increasing order from a min-heap.  A rewrite at key
// (x, y) can only make pairs (r, c) with r > x dirty (for r == x the row
// itself re-gathers its candidates around the new snapshot), so the heap
// never needs a row below the one being replayed.  A body's live position
// at key K is its pre-contact position plus, in key order, every clean
// correction below K (re-derived from the snapshot exactly as k_contacts
// does) and every logged one below K.  At the end each body touched by a
// dirty pair gets that sum over all keys; the contact counter and the
// collision events are corrected for dirty pairs whose contact status
// changed.
#include "common.hpp"
#include "device_state.hpp"
#include "step_common.hpp"

namespace bg {

namespace {

constexpr unsigned long long kNever = ~0ull;

__device__ __forceinline__ unsigned long long pair_key(uint32_t a, uint32_t b) {
    return a > b ? (((unsigned long long)a << 32) | b) : (((unsigned long long)b << 32) | a);
}

struct Ordered {
    DeviceWorld d;
    float rmax;
    uint32_t n_moved, n_log, n_aff, heap_n;
    uint32_t flags;
    uint32_t push_outs;
    long long count_delta;
};

__device__ bool mark_bit(uint32_t* bits, uint32_t i) {
    const uint32_t m = 1u << (i & 31u);
    const uint32_t old = bits[i >> 5];
    bits[i >> 5] = old | m;
    return (old & m) == 0u;
}
__device__ void clear_bit(uint32_t* bits, uint32_t i) { bits[i >> 5] &= ~(1u << (i & 31u)); }

__device__ void heap_push(Ordered& o, uint32_t v) {
    uint32_t* h = o.d.ord_heap;
    uint32_t i = o.heap_n++;
    while (i > 0) {
        const uint32_t p = (i - 1) >> 1;
        if (h[p] <= v) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = v;
}

__device__ uint32_t heap_pop(Ordered& o) {
    uint32_t* h = o.d.ord_heap;
    const uint32_t top = h[0];
    const uint32_t v = h[--o.heap_n];
    uint32_t i = 0;
    for (;;) {
        uint32_t c = 2 * i + 1;
        if (c >= o.heap_n) break;
        if (c + 1 < o.heap_n && h[c + 1] < h[c]) ++c;
        if (v <= h[c]) break;
        h[i] = h[c];
        i = c;
    }
    if (o.heap_n > 0) h[i] = v;
    return top;
}

// Collider c's snapshot as the loop sees it now: rewritten, or the substep's.
__device__ __forceinline__ float2 cur_abs(const Ordered& o, uint32_t c) {
    return o.d.mv_key[c] != kNever ? o.d.mv_abs[c] : o.d.c.abs[c];
}

// Colliders whose substep snapshot lies in the cells around q (half-width R,
// padded for rounding), in any order.  Non-finite queries see nothing: no
// pair with a non-finite snapshot is a contact.
template <class F>
__device__ void grid_visit(const Ordered& o, float2 q, float R, F&& f) {
    if (!finitef(q.x) || !finitef(q.y) || !finitef(R)) return;
    const GridParams g = *o.d.grid;
    R = R + (1e-5f * (fabsf(q.x) + fabsf(q.y) + R) + 1e-6f);
    const int32_t cx0 = cell_coord(q.x - R, g.ox, g.inv_cell, g.nx), cx1 = cell_coord(q.x + R, g.ox, g.inv_cell, g.nx);
    const int32_t cy0 = cell_coord(q.y - R, g.oy, g.inv_cell, g.ny), cy1 = cell_coord(q.y + R, g.oy, g.inv_cell, g.ny);
    for (int32_t cy = cy0; cy <= cy1; ++cy) {
        const uint32_t row = (uint32_t)cy * (uint32_t)g.nx;
        const uint32_t t0 = o.d.cell_start[row + (uint32_t)cx0];
        uint32_t t1 = o.d.cell_start[row + (uint32_t)cx1 + 1u];
        t1 = t1 < o.d.nc ? t1 : o.d.nc;
        for (uint32_t t = t0; t < t1; ++t) f(o.d.vals[t]);
    }
}

// Position-independent part of the reference predicate (physics.rs:251-261):
// both parents exist, the groups test passes, the parents differ.
template <bool GROUPS>
__device__ __forceinline__ bool pair_allowed(const Ordered& o, uint32_t a, uint32_t b) {
    const int32_t pa = o.d.c.parent[a], pb = o.d.c.parent[b];
    if (pa < 0 || pb < 0) return false;
    if (GROUPS) {
        const uint2 ga = o.d.c.groups[a], gb = o.d.c.groups[b];
        if (!((ga.x & gb.y) != 0u && (gb.x & ga.y) != 0u)) return false;
    }
    return pa != pb;
}

template <bool SENSORS>
__device__ __forceinline__ bool is_sensor(const Ordered& o, uint32_t c) {
    return SENSORS ? (o.d.c.cflags[c] & 1u) != 0u : false;
}

// Contact correction of both bodies for axis = A - B (physics.rs:292-299,
// mass_ratio :319-321): A gets + (ratio*delta)*n, B gets - ((1-ratio)*delta)*n.
__device__ __forceinline__ void pair_correction(float ax, float ay, float dist, float rs, float ma, float mb,
                                                float2& ca, float2& cb, uint32_t& flags) {
    const float nx = ax / dist, ny = ay / dist;
    if (nx != nx || ny != ny) flags |= ERR_NAN_NORMAL;
    const float delta = rs - dist;
    const float ratio = 1.0f - ma / (ma + mb);
    const float ka = ratio * delta;
    const float kb = (1.0f - ratio) * delta;
    ca = make_float2(ka * nx, ka * ny);
    cb = make_float2(-(kb * nx), -(kb * ny));
}

// Live position of body b before key K (K = kNever: after the whole pass).
template <bool GROUPS, bool SENSORS>
__device__ float2 live_pos(Ordered& o, int32_t b, unsigned long long K) {
    const DeviceWorld& d = o.d;
    OrderedTerm* t = d.ord_terms;
    uint32_t n = 0;
    auto add = [&](const OrderedTerm& e) {
        if (n == kOrdTerms) {
            o.flags |= ERR_ORDERED_CAP;
            return;
        }
        // insertion into the key-sorted list (short lists: a body's contacts)
        uint32_t i = n++;
        while (i > 0 && t[i - 1].key > e.key) {
            t[i] = t[i - 1];
            --i;
        }
        t[i] = e;
    };
    // clean pairs, from the substep snapshot (the same arithmetic as k_contacts)
    for (uint32_t k = d.bc_start[b]; k < d.bc_start[b + 1]; ++k) {
        const uint32_t c = d.bc_list[k];
        const float2 xc = d.c.abs[c];
        const float rc = d.c.radius[c];
        grid_visit(o, xc, rc + o.rmax, [&](uint32_t y) {
            if (y == c || !pair_allowed<GROUPS>(o, c, y)) return;
            const unsigned long long key = pair_key(c, y);
            if (key >= K) return;
            if (d.mv_key[c] < key || d.mv_key[y] < key) return;  // dirty: logged
            const float2 xy = d.c.abs[y];
            const bool isA = c > y;
            const float ax = isA ? xc.x - xy.x : xy.x - xc.x, ay = isA ? xc.y - xy.y : xy.y - xc.y;
            const float dist = vlen(ax, ay);
            const float rs = rc + d.c.radius[y];
            if (!(dist < rs)) return;
            if (dist < 1e-6f) return;                               // coincident: dirty, logged
            if (is_sensor<SENSORS>(o, c) || is_sensor<SENSORS>(o, y)) return;
            const float mown = d.b.mass[b], mpart = d.b.mass[d.c.parent[y]];
            float2 ca, cb;
            uint32_t fl = 0;
            pair_correction(ax, ay, dist, rs, isA ? mown : mpart, isA ? mpart : mown, ca, cb, fl);
            OrderedTerm e;
            e.key = key;
            e.push = make_float2(0.0f, 0.0f);
            e.c = isA ? ca : cb;
            e.flags = OT_CORR;
            e.pad = 0;
            add(e);
        });
    }
    // dirty pairs, from the log (kept in key order)
    for (uint32_t i = 0; i < o.n_log; ++i) {
        const OrderedPair& p = d.ord_log[i];
        if (p.key >= K) break;
        if (p.pa != b && p.pb != b) continue;
        const bool a_side = p.pa == b;
        OrderedTerm e;
        e.key = p.key;
        e.flags = 0;
        e.pad = 0;
        e.push = a_side ? make_float2(0.01f, 0.0f) : make_float2(-0.01f, -0.0f);
        e.c = a_side ? p.ca : p.cb;
        if (p.flags & OP_PUSHED) e.flags |= OT_PUSH;
        if ((p.flags & OP_CONTACT) && !(p.flags & OP_SENSOR)) e.flags |= OT_CORR;
        if (e.flags) add(e);
    }
    float2 pos = d.pos_pre[b];
    for (uint32_t i = 0; i < n; ++i) {
        if (t[i].flags & OT_PUSH) {  // rbd.position +-= push_out (physics.rs:275-276)
            pos.x = pos.x + t[i].push.x;
            pos.y = pos.y + t[i].push.y;
        }
        if (t[i].flags & OT_CORR) {  // physics.rs:298-299
            pos.x = pos.x + t[i].c.x;
            pos.y = pos.y + t[i].c.y;
        }
    }
    return pos;
}

// Rows r > x whose collider's substep snapshot lies within reach of q: the
// pairs (r, c) of a collider c now at q become dirty.
__device__ void enqueue_near(Ordered& o, uint32_t x, float2 q, float rc) {
    grid_visit(o, q, rc + o.rmax, [&](uint32_t r) {
        if (r > x && mark_bit(o.d.ord_mark, r)) heap_push(o, r);
    });
}

__device__ void note_body(Ordered& o, int32_t b) {
    if (mark_bit(o.d.aff_mark, (uint32_t)b)) o.d.aff_list[o.n_aff++] = (uint32_t)b;
}

template <bool GROUPS, bool SENSORS>
__device__ void replay_row(Ordered& o, const SubstepArgs& a, uint32_t x) {
    const DeviceWorld& d = o.d;
    if (d.c.parent[x] < 0) return;
    const float rx = d.c.radius[x];
    uint32_t lo = 0;  // candidates y >= lo (after a rewrite of x: above the pair that made it)
    for (;;) {
        // candidates: substep-snapshot neighbours of x's substep snapshot (pairs that were
        // contacts) and of its current one, plus every rewritten collider
        uint32_t* cand = d.ord_cand;
        uint32_t n = 0;
        auto take = [&](uint32_t y) {
            if (y >= x || y < lo) return;
            if (n == kOrdCand) {
                o.flags |= ERR_ORDERED_CAP;
                return;
            }
            cand[n++] = y;
        };
        grid_visit(o, d.c.abs[x], rx + o.rmax, take);
        if (d.mv_key[x] != kNever) grid_visit(o, d.mv_abs[x], rx + o.rmax, take);
        for (uint32_t i = 0; i < o.n_moved; ++i) take(d.mv_list[i]);
        for (uint32_t i = 1; i < n; ++i) {  // ascending
            const uint32_t v = cand[i];
            uint32_t j = i;
            while (j > 0 && cand[j - 1] > v) {
                cand[j] = cand[j - 1];
                --j;
            }
            cand[j] = v;
        }
        bool rewritten = false;
        for (uint32_t i = 0; i < n && !rewritten; ++i) {
            const uint32_t y = cand[i];
            if (i > 0 && cand[i - 1] == y) continue;
            if (!pair_allowed<GROUPS>(o, x, y)) continue;
            const unsigned long long key = pair_key(x, y);
            const float ry = d.c.radius[y];
            const float rs = rx + ry;
            // the pair on the substep snapshot (what k_contacts counted)
            const float2 x0 = d.c.abs[x], y0 = d.c.abs[y];
            const float d0 = vlen(x0.x - y0.x, x0.y - y0.y);
            const bool orig = d0 < rs;
            const bool dirty = d.mv_key[x] < key || d.mv_key[y] < key || (orig && d0 < 1e-6f);
            if (!dirty) continue;
            // the pair as the ordered loop sees it (physics.rs:264-269)
            const float2 xa = cur_abs(o, x), xb = cur_abs(o, y);
            float ax = xa.x - xb.x, ay = xa.y - xb.y;
            float dist = vlen(ax, ay);
            const int32_t pa = d.c.parent[x], pb = d.c.parent[y];
            const bool contact = dist < rs;
            if (!contact && !orig) continue;
            OrderedPair p;
            p.key = key;
            p.pa = pa;
            p.pb = pb;
            p.flags = (orig ? OP_ORIG : 0u) | (contact ? OP_CONTACT : 0u);
            p.pad = 0;
            p.ca = p.cb = make_float2(0.0f, 0.0f);
            if (contact && dist < 1e-6f) {  // push-out (physics.rs:272-286)
                float2 la = live_pos<GROUPS, SENSORS>(o, pa, key);
                float2 lb = live_pos<GROUPS, SENSORS>(o, pb, key);
                la.x = la.x + 0.01f;
                la.y = la.y + 0.0f;
                lb.x = lb.x - 0.01f;
                lb.y = lb.y - 0.0f;
                const float2 oa = d.c.off[x], ob = d.c.off[y];
                const float2 na = make_float2(la.x + oa.x, la.y + oa.y);
                const float2 nb = make_float2(lb.x + ob.x, lb.y + ob.y);
                for (int s = 0; s < 2; ++s) {
                    const uint32_t c = s == 0 ? x : y;
                    const float2 q = s == 0 ? na : nb;
                    if (d.mv_key[c] == kNever) {
                        d.mv_key[c] = key;
                        d.mv_list[o.n_moved++] = c;
                        enqueue_near(o, x, d.c.abs[c], d.c.radius[c]);  // its former contacts
                    }
                    d.mv_abs[c] = q;
                    enqueue_near(o, x, q, d.c.radius[c]);
                }
                ax = na.x - nb.x;
                ay = na.y - nb.y;
                dist = vlen(ax, ay);
                p.flags |= OP_PUSHED;
                ++o.push_outs;
                rewritten = true;  // the rest of row x reads x's new snapshot
                lo = y + 1;
            }
            const bool sensor = is_sensor<SENSORS>(o, x) || is_sensor<SENSORS>(o, y);
            if (sensor) p.flags |= OP_SENSOR;
            if (contact && !sensor)
                pair_correction(ax, ay, dist, rs, d.b.mass[pa], d.b.mass[pb], p.ca, p.cb, o.flags);
            if (o.n_log == d.ord_log_cap) {
                o.flags |= ERR_ORDERED_CAP;
                return;
            }
            d.ord_log[o.n_log++] = p;
            note_body(o, pa);
            note_body(o, pb);
            if (contact != orig) {
                // k_contacts counted and sent the pair on the substep snapshot
                o.count_delta += contact ? 1 : -1;
                if (d.events) {
                    const unsigned long long e = atomicAdd(&d.accum->event_count, 1ull);
                    if (e < d.event_cap) {
                        DevEvent ev;
                        ev.sub = a.sub_global;
                        ev.a = x;
                        ev.b = y;
                        ev.pad = contact ? 0u : 1u;  // 1: withdraws k_contacts' event of the pair
                        ev.va = d.b.vel[pa];
                        ev.vb = d.b.vel[pb];
                        d.events[e] = ev;
                    }
                }
            }
        }
        if (!rewritten) return;
    }
}

template <bool GROUPS, bool SENSORS>
__global__ __launch_bounds__(64) void k_contacts_ordered(DeviceWorld dw, SubstepArgs a) {
    const uint32_t n_rows = dw.ord_n[0];
    if (n_rows == 0 || threadIdx.x != 0) return;
    Ordered o;
    o.d = dw;
    o.rmax = a.rmax;
    o.n_moved = o.n_log = o.n_aff = o.heap_n = 0;
    o.flags = 0;
    o.push_outs = 0;
    o.count_delta = 0;
    const uint32_t nr = n_rows < dw.nc ? n_rows : dw.nc;
    for (uint32_t i = 0; i < nr; ++i) heap_push(o, dw.ord_rows[i]);
    while (o.heap_n > 0 && !(o.flags & ERR_ORDERED_CAP)) {
        const uint32_t x = heap_pop(o);
        clear_bit(dw.ord_mark, x);
        replay_row<GROUPS, SENSORS>(o, a, x);
    }
    if (!(o.flags & ERR_ORDERED_CAP)) {
        // final positions of the bodies of dirty pairs (the sums read pos_pre, never pos)
        for (uint32_t i = 0; i < o.n_aff; ++i) {
            const int32_t b = (int32_t)dw.aff_list[i];
            dw.b.pos[b] = live_pos<GROUPS, SENSORS>(o, b, kNever);
        }
    }
    // reset the scratch for the next substep (graph-safe: no memset)
    while (o.heap_n > 0) clear_bit(dw.ord_mark, heap_pop(o));
    for (uint32_t i = 0; i < o.n_moved; ++i) dw.mv_key[dw.mv_list[i]] = kNever;
    for (uint32_t i = 0; i < o.n_aff; ++i) clear_bit(dw.aff_mark, dw.aff_list[i]);
    dw.ord_n[0] = 0;
    if (o.count_delta != 0) atomicAdd(&dw.block_contacts[0], (unsigned long long)o.count_delta);
    if (o.push_outs) atomicAdd(&dw.accum->push_outs, o.push_outs);
    if (o.flags) atomicOr(&dw.accum->error_flags, o.flags);
}

}  // namespace

hipError_t launch_contacts_ordered(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors) {
    if (d.nc == 0 || !d.ord_n) return hipSuccess;
    if (groups && sensors)
        k_contacts_ordered<true, true><<<1, 64, 0, d.stream>>>(d, a);
    else if (groups)
        k_contacts_ordered<true, false><<<1, 64, 0, d.stream>>>(d, a);
    else if (sensors)
        k_contacts_ordered<false, true><<<1, 64, 0, d.stream>>>(d, a);
    else
        k_contacts_ordered<false, false><<<1, 64, 0, d.stream>>>(d, a);
    return hipGetLastError();
}

}  // namespace bg
