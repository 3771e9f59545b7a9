// batch.cpp -- batches of independent single-sphere worlds, stepped as one
// device world (one launch per substep covers every world), hipGraph replay
// of multi-step chunks, and the broadphase autotuner.
//
// A batch copies its source worlds' bodies and colliders, world after world
// in each world's own iteration order, so world w owns the dense range
// [batch_start_[w], batch_start_[w+1]).  Every world keeps its own grid,
// Physics::old_dt (physics.rs:67, 338-340) and time; contacts never cross
// worlds.  Results are bit-identical to stepping each world on its own.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "world.hpp"

namespace bg {

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess) return fail(BG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

void World::dense_slots(std::vector<uint32_t>& bodies, std::vector<uint32_t>& cols) const {
    bodies.clear();
    cols.clear();
    for (uint32_t s = 0; s < bodies_.capacity_slots(); ++s)
        if (bodies_.occupied(s)) bodies.push_back(s);
    for (uint32_t s = 0; s < colliders_.capacity_slots(); ++s)
        if (colliders_.occupied(s)) cols.push_back(s);
}

// The fused sphere path's condition, checked on the host arenas: the k-th
// collider in iteration order is the only collider of the k-th body, and
// there are no springs or joints.
bool World::sphere_eligible() const {
    if (springs_.len() != 0 || joints_.len() != 0) return false;
    std::vector<uint32_t> bs, cs;
    dense_slots(bs, cs);
    if (bs.size() != cs.size()) return false;
    for (size_t k = 0; k < bs.size(); ++k) {
        const HostBody& hb = bodies_.at_slot(bs[k]);
        const bg_collider& c = colliders_.at_slot(cs[k]).f;
        const bg_handle ch = colliders_.handle_at(cs[k]);
        if (!c.has_parent || c.parent != bodies_.handle_at(bs[k])) return false;
        for (bg_handle h : hb.colliders)
            if (h != ch) return false;
        if (hb.colliders.empty()) return false;
    }
    return true;
}

int World::init_batch(World* const* src, uint32_t n) {
    if (n == 0) return fail(BG_ERR_INVALID_ARGUMENT, "a batch needs at least one world");
    if (bodies_.len() || colliders_.len() || batch_size())
        return fail(BG_ERR_INVALID_ARGUMENT, "a batch is built into an empty world");
    const bg_params& p0 = src[0]->params_;
    for (uint32_t w = 0; w < n; ++w) {
        World* s = src[w];
        if (!s) return fail(BG_ERR_INVALID_ARGUMENT, "null world in batch");
        if (s->sync_host()) return fail(BG_ERR_HIP, "download of world " + std::to_string(w) + " failed");
        const bg_params& p = s->params_;
        if (p.substeps != p0.substeps || p.gravity[0] != p0.gravity[0] || p.gravity[1] != p0.gravity[1] ||
            p.collisions_enabled != p0.collisions_enabled || p.use_spatial_hash != p0.use_spatial_hash)
            return fail(BG_ERR_INVALID_ARGUMENT,
                        "worlds of a batch share substeps, gravity and collisions_enabled (world " +
                            std::to_string(w) + " differs)");
        if (!s->constraints_.empty())
            return fail(BG_ERR_INVALID_ARGUMENT, "worlds of a batch have no circle constraints");
        if (!s->sphere_eligible())
            return fail(BG_ERR_INVALID_ARGUMENT, "world " + std::to_string(w) +
                                                     " is not a single-sphere world (one collider per body in the "
                                                     "same order, no springs or joints)");
    }
    params_ = p0;
    params_.time = 0.0;
    params_.accumulator = 0.0;
    // A batch's cells default to 1.5 x 2 r_max: over 850 steps of 4,096 uneven worlds (which
    // fall and spread) 2.28 G sphere-steps/s against 2.23 G at 1.25 and 2.08 G at 1.5 with a
    // skin of 0.35; the autotuner's trials, taken at the start, do not see the spreading.
    cell_scale_ = 1.5f;
    batch_start_.assign(1, 0u);
    batch_old_dt_.clear();
    batch_time_.clear();
    std::vector<uint32_t> bs, cs;
    size_t total = 0;
    for (uint32_t w = 0; w < n; ++w) total += src[w]->bodies_.len();
    bodies_.reserve(total);
    colliders_.reserve(total);
    for (uint32_t w = 0; w < n; ++w) {
        const World* s = src[w];
        s->dense_slots(bs, cs);
        for (size_t k = 0; k < bs.size(); ++k) {
            const HostBody& sb = s->bodies_.at_slot(bs[k]);
            const size_t dup = sb.colliders.size();
            HostBody hb;
            hb.f = sb.f;
            const bg_handle bh = bodies_.insert(hb);
            HostCollider hc = s->colliders_.at_slot(cs[k]);
            hc.f.parent = bh;
            const bg_handle ch = colliders_.insert(hc);
            bodies_.get(bh)->colliders.assign(dup, ch);  // the handle twice (Q2), mass already set up
        }
        batch_start_.push_back(bodies_.len());
        batch_old_dt_.push_back(s->params_.old_dt);
        batch_time_.push_back(s->params_.time);
    }
    mark_dirty();
    return BG_OK;
}

int World::batch_writeback(World* const* dst, uint32_t n) {
    if (n != batch_size()) return fail(BG_ERR_INVALID_ARGUMENT, "writeback needs the batch's worlds, in order");
    if (sync_host()) return BG_ERR_HIP;
    std::vector<uint32_t> my_b, my_c, bs, cs;
    dense_slots(my_b, my_c);
    for (uint32_t w = 0; w < n; ++w) {
        World* d = dst[w];
        if (!d) return fail(BG_ERR_INVALID_ARGUMENT, "null world in writeback");
        d->dense_slots(bs, cs);
        const uint32_t b0 = batch_start_[w], nw = batch_start_[w + 1] - b0;
        if (bs.size() != nw || cs.size() != nw)
            return fail(BG_ERR_INVALID_ARGUMENT, "world " + std::to_string(w) + " changed since the batch was built");
    }
    for (uint32_t w = 0; w < n; ++w) {
        World* d = dst[w];
        if (d->sync_host()) return fail(BG_ERR_HIP, "download of world " + std::to_string(w) + " failed");
        d->dense_slots(bs, cs);
        const uint32_t b0 = batch_start_[w];
        for (size_t k = 0; k < bs.size(); ++k) {
            const bg_rigid_body& f = bodies_.at_slot(my_b[b0 + k]).f;
            bg_rigid_body& g = d->bodies_.at_slot(bs[k]).f;
            memcpy(g.position, f.position, sizeof(g.position));
            memcpy(g.position_old, f.position_old, sizeof(g.position_old));
            memcpy(g.calculated_velocity, f.calculated_velocity, sizeof(g.calculated_velocity));
            memcpy(g.acceleration, f.acceleration, sizeof(g.acceleration));
            g.rotation = f.rotation;
            g.angular_velocity = f.angular_velocity;
            g.torque = f.torque;
            g.has_velocity_request = f.has_velocity_request;
            const bg_collider& c = colliders_.at_slot(my_c[b0 + k]).f;
            bg_collider& e = d->colliders_.at_slot(cs[k]).f;
            memcpy(e.absolute_translation, c.absolute_translation, sizeof(e.absolute_translation));
            memcpy(e.absolute_matrix, c.absolute_matrix, sizeof(e.absolute_matrix));
        }
        d->params_.old_dt = batch_old_dt_[w];
        d->params_.time = batch_time_[w];
        d->mark_dirty();  // its device copy (if any) is older than the host mirror now
    }
    return BG_OK;
}

namespace {

// Host-side state a captured chunk depends on: the chunk may be replayed
// only from the state it was captured in.
struct LaunchState {
    const void* p[8];
    uint32_t u[5];
    bool operator==(const LaunchState& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
};

}  // namespace

int World::step_chunks(double delta, uint32_t chunk_steps, uint32_t n_chunks) {
    if (chunk_steps == 0 || n_chunks == 0) return BG_OK;
    int st = prepare_step();
    if (st) return st;
    DeviceWorld& d = d_;
    auto state = [&]() {
        LaunchState s;
        memset(&s, 0, sizeof(s));
        s.p[0] = d.s.pp;
        s.p[1] = d.s.snap;
        s.p[2] = d.snap_next;
        s.p[3] = d.s_alt.pp;
        s.p[4] = d.b.pos;
        s.p[5] = d.keys;
        s.p[6] = p_src_.pp;  // the persistent launch reads its worlds from here
        s.p[7] = p_src_.snap;
        s.u[0] = bbox_parity_;
        s.u[1] = steps_since_rebuild_;
        s.u[2] = (need_rebuild_ ? 1u : 0u) | (d.gather_pending ? 2u : 0u) | (fresh_ ? 4u : 0u);
        s.u[3] = sub_since_rebuild_;
        s.u[4] = rebuild_interval_;
        return s;
    };
    const bool prof = profiling_;
    uint32_t done = 0;
    if (event_cap_ > 0) {
        // events are stamped with the substep index the host passes: a replayed graph would
        // repeat the stamps, so a world that records events runs its chunks eagerly
        st = launch_steps(delta, chunk_steps * n_chunks);
        if (st) return st;
        advance_time(delta, chunk_steps * n_chunks);
        host_stale_ = true;
        return check_and_process();
    }
    // the first chunk after an upload runs eagerly (user inputs are consumed once)
    if (fresh_) {
        st = launch_steps(delta, chunk_steps);
        if (st) return st;
        advance_time(delta, chunk_steps);
        ++done;
    }
    // Capture graphs until the launch state repeats (never more than the chunks left), then replay.
    // Capturing runs no kernel; it advances the host mirror exactly as the replay will advance the device.
    std::vector<hipGraphExec_t> execs;
    std::vector<hipGraph_t> graphs;
    auto release = [&]() {
        for (hipGraphExec_t x : execs) (void)hipGraphExecDestroy(x);
        for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
    };
    // one chunk's launches recorded into a graph (kept if `keep`)
    auto capture = [&](bool keep) -> int {
        hipGraph_t g = nullptr;
        profiling_ = false;  // no event records inside a capture
        hipError_t e = hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal);
        int r = BG_OK;
        if (e == hipSuccess) {
            r = launch_steps(delta, chunk_steps);
            e = hipStreamEndCapture(d.stream, &g);
        }
        profiling_ = prof;
        if (r == BG_OK && e != hipSuccess) r = fail(BG_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(e));
        if (r == BG_OK && keep) {
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                r = fail(BG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            } else {
                graphs.push_back(g);
                execs.push_back(x);
                return BG_OK;
            }
        }
        if (g) (void)hipGraphDestroy(g);
        return r;
    };
    int result = BG_OK;
    const uint32_t todo = n_chunks - done;
    bool cyclic = false;
    if (todo > 0) {
        const LaunchState s0 = state();
        const uint32_t max_graphs = std::min<uint32_t>(4u, todo);
        while (result == BG_OK && execs.size() < max_graphs) {
            result = capture(true);
            if (result == BG_OK && state() == s0) cyclic = true;
            if (cyclic) break;
        }
        const uint32_t m = (uint32_t)execs.size();
        for (uint32_t c = 0; c < todo && result == BG_OK; ++c) {
            if (c < m || cyclic) {
                hipError_t e = hipGraphLaunch(execs[c % m], d.stream);
                if (e != hipSuccess) result = fail(BG_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
            } else {
                result = launch_steps(delta, chunk_steps);  // past a non-repeating capture: eager
            }
            if (result == BG_OK) advance_time(delta, chunk_steps);
        }
        // cyclic: the host mirror is back at the first graph's start; bring it to where the
        // device stopped (todo % m graphs into the cycle)
        if (cyclic && result == BG_OK)
            for (uint32_t r = 0; r < todo % m && result == BG_OK; ++r) result = capture(false);
    }
    hipError_t e = hipStreamSynchronize(d.stream);
    release();
    if (e != hipSuccess && result == BG_OK) result = fail(BG_ERR_HIP, hipGetErrorString(e));
    host_stale_ = true;
    if (result != BG_OK) return result;
    return check_and_process();
}

int World::set_persistent(bool enable) {
    if (enable == persist_enabled_) return BG_OK;
    if (sync_host()) return BG_ERR_HIP;  // the state moves through the host mirror
    persist_enabled_ = enable;
    mark_dirty();
    return BG_OK;
}

int World::autotune(double delta, uint32_t steps_per_trial, float* cell_scale, float* skin, uint32_t* interval) {
    if (steps_per_trial == 0) steps_per_trial = 1;
    int st = prepare_step();
    if (st) return st;
    // a batch the persistent step can run: the grid candidates are timed on the grid path
    // alone, then the persistent schedule with the best of them (each switch re-uploads
    // through the host mirror)
    const bool try_persist = persist_;
    if (try_persist) {
        st = set_persistent(false);
        if (st == BG_OK) st = prepare_step();
        if (st) return st;
    }
    struct Cand {
        float cs, sk;
        uint32_t iv;
    };
    // cell scale x skin x rebuild interval
    std::vector<Cand> cands;
    for (float cs : {1.0f, 1.25f, 1.5f, 2.0f})
        for (float sk : {0.25f, 0.35f, 0.5f})
            for (uint32_t iv : {1u, 2u, 3u}) cands.push_back({cs, sk, iv});
    // consume the upload's one-off inputs before timing
    if (fresh_) {
        st = launch_steps(delta, 1);
        if (st) return st;
        advance_time(delta, 1);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best_ms = 0.0f;
    Cand best = cands[0];
    bool have = false;
    for (const Cand& c : cands) {
        st = set_broadphase(c.cs, c.sk, c.iv);
        if (st) break;
        // one untimed step so the new grid exists, then the timed trial
        st = launch_steps(delta, 1);
        if (st) break;
        advance_time(delta, 1);
        if (hipEventRecord(e0, d_.stream) != hipSuccess) break;
        st = launch_steps(delta, steps_per_trial);
        if (st) break;
        advance_time(delta, steps_per_trial);
        if (hipEventRecord(e1, d_.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (!have || ms < best_ms) {
            best_ms = ms;
            best = c;
            have = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    host_stale_ = true;
    if (st) return st;
    if (!have) return fail(BG_ERR_HIP, "autotune: event timing failed");
    st = set_broadphase(best.cs, best.sk, best.iv);
    if (st) return st;
    if (try_persist) {
        st = set_persistent(true);
        if (st == BG_OK) st = prepare_step();
        if (st) return st;
        float persist_ms = -1.0f;
        hipEvent_t p0 = nullptr, p1 = nullptr;
        HIP_TRY(hipEventCreate(&p0));
        HIP_TRY(hipEventCreate(&p1));
        st = launch_steps(delta, 1);  // untimed: consumes the upload's one-off inputs, first grid
        if (st == BG_OK) advance_time(delta, 1);
        hipError_t e = hipSuccess;
        if (st == BG_OK) e = hipEventRecord(p0, d_.stream);
        if (st == BG_OK && e == hipSuccess) st = launch_steps(delta, steps_per_trial);
        if (st == BG_OK && e == hipSuccess) {
            advance_time(delta, steps_per_trial);
            e = hipEventRecord(p1, d_.stream);
            if (e == hipSuccess) e = hipEventSynchronize(p1);
            if (e == hipSuccess) e = hipEventElapsedTime(&persist_ms, p0, p1);
        }
        (void)hipEventDestroy(p0);
        (void)hipEventDestroy(p1);
        host_stale_ = true;
        if (st) return st;
        if (e != hipSuccess) return fail(BG_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        if (persist_ms > best_ms) {
            st = set_persistent(false);
            if (st) return st;
        }
    }
    if (cell_scale) *cell_scale = best.cs;
    if (skin) *skin = best.sk;
    if (interval) *interval = best.iv;
    return check_and_process();
}

}  // namespace bg
