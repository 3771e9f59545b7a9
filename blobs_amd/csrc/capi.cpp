// capi.cpp -- extern "C" entry points declared in include/blobs_gpu.h.
// Thin: argument checks, exception containment, dispatch to bg::World.
#include <string.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/blobs_gpu.h"
#include "diff.hpp"
#include "world.hpp"

struct bg_world {
    bg::World impl;
    bg_world(float gx, float gy, bool sh) : impl(gx, gy, sh) {}
};

struct bg_rollout {
    bg::Rollout impl;
};

namespace {

thread_local std::string g_last_error;

}  // namespace

namespace bg {
// Message of a failed call that has no world (read by bg_last_error(NULL, ...)).
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace bg

namespace {

template <class F>
int guarded(bg_world* w, F&& f) {
    if (!w) {
        g_last_error = "null world";
        return BG_ERR_INVALID_ARGUMENT;
    }
    try {
        int st = f(w->impl);
        if (st != BG_OK) g_last_error = w->impl.last_error;
        return st;
    } catch (const std::bad_alloc&) {
        w->impl.last_error = g_last_error = "host out of memory";
        return BG_ERR_CAPACITY;
    } catch (const std::exception& e) {
        w->impl.last_error = g_last_error = e.what();
        return BG_ERR_PANIC;
    }
}

template <class F>
int rguard(bg_rollout* r, F&& f) {
    if (!r) {
        g_last_error = "null rollout";
        return BG_ERR_INVALID_ARGUMENT;
    }
    try {
        int st = f(r->impl);
        if (st != BG_OK) g_last_error = r->impl.last_error;
        return st;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return BG_ERR_PANIC;
    }
}

}  // namespace

extern "C" {

int bg_world_create(float gravity_x, float gravity_y, int use_spatial_hash, bg_world** out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    try {
        *out = new bg_world(gravity_x, gravity_y, use_spatial_hash != 0);
    } catch (...) {
        g_last_error = "allocation failed";
        return BG_ERR_CAPACITY;
    }
    return BG_OK;
}

int bg_world_destroy(bg_world* w) {
    delete w;
    return BG_OK;
}

int bg_world_reset(bg_world* w) {
    return guarded(w, [&](bg::World& x) { return x.reset(); });
}

int bg_get_params(bg_world* w, bg_params* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.get_params(out); });
}

int bg_set_params(bg_world* w, const bg_params* p) {
    if (!p) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.set_params(p); });
}

int bg_insert_rbd(bg_world* w, const bg_rigid_body* body, bg_handle* out) {
    if (!body) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.insert_rbd(*body, out); });
}

int bg_insert_collider_with_parent(bg_world* w, const bg_collider* col, bg_handle parent, bg_handle* out) {
    if (!col) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.insert_collider_with_parent(*col, parent, out); });
}

int bg_insert_rbds(bg_world* w, uint32_t n, const bg_body_batch* b, bg_handle* out) {
    if (!b) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.insert_rbds(n, *b, out); });
}

int bg_insert_colliders_with_parents(bg_world* w, uint32_t n, const bg_collider_batch* c, const bg_handle* parents,
                                     bg_handle* out) {
    if (!c) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.insert_colliders(n, *c, parents, out); });
}

int bg_create_fixed_joint(bg_world* w, bg_handle a, bg_handle b, const float anchor_a[2], const float anchor_b[2],
                          bg_handle* out) {
    if (!anchor_a || !anchor_b) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.create_fixed_joint(a, b, anchor_a, anchor_b, false, 0.0f, out); });
}

int bg_create_fixed_joint_with_distance(bg_world* w, bg_handle a, bg_handle b, const float anchor_a[2],
                                        const float anchor_b[2], float distance, bg_handle* out) {
    if (!anchor_a || !anchor_b) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w,
                   [&](bg::World& x) { return x.create_fixed_joint(a, b, anchor_a, anchor_b, true, distance, out); });
}

int bg_insert_spring(bg_world* w, const bg_spring* s, bg_handle* out) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.insert_spring(*s, out); });
}

int bg_push_constraint(bg_world* w, float x_, float y_, float radius) {
    return guarded(w, [&](bg::World& x) { return x.push_constraint(x_, y_, radius); });
}

int bg_remove_rbd(bg_world* w, bg_handle h) {
    return guarded(w, [&](bg::World& x) { return x.remove_rbd(h); });
}

int bg_remove_col(bg_world* w, bg_handle h) {
    return guarded(w, [&](bg::World& x) { return x.remove_col(h); });
}

int bg_step(bg_world* w, double delta, uint32_t n_steps) {
    return guarded(w, [&](bg::World& x) { return x.step(delta, n_steps); });
}

int bg_fixed_step(bg_world* w, double frame_time) {
    return guarded(w, [&](bg::World& x) { return x.fixed_step(frame_time); });
}

int bg_synchronize(bg_world* w) {
    return guarded(w, [&](bg::World& x) { return x.synchronize(); });
}

int bg_batch_run(bg_world* w, double delta, uint32_t chunk_steps, uint32_t n_chunks) {
    return guarded(w, [&](bg::World& x) { return x.step_chunks(delta, chunk_steps, n_chunks); });
}

int bg_batch_create(bg_world* const* worlds, uint32_t n_worlds, bg_world** out) {
    if (!out || !worlds || n_worlds == 0 || !worlds[0]) {
        g_last_error = "bg_batch_create: worlds, n_worlds > 0 and out are required";
        return BG_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    bg_params p;
    int st = bg_get_params(worlds[0], &p);
    if (st != BG_OK) return st;
    bg_world* b = nullptr;
    try {
        b = new bg_world(p.gravity[0], p.gravity[1], p.use_spatial_hash != 0);
    } catch (...) {
        g_last_error = "allocation failed";
        return BG_ERR_CAPACITY;
    }
    std::vector<bg::World*> src(n_worlds);
    for (uint32_t i = 0; i < n_worlds; ++i) src[i] = worlds[i] ? &worlds[i]->impl : nullptr;
    st = guarded(b, [&](bg::World& x) { return x.init_batch(src.data(), n_worlds); });
    if (st != BG_OK) {
        delete b;
        return st;
    }
    *out = b;
    return BG_OK;
}

int bg_batch_writeback(bg_world* batch, bg_world* const* worlds, uint32_t n_worlds) {
    if (!worlds) return BG_ERR_INVALID_ARGUMENT;
    std::vector<bg::World*> dst(n_worlds);
    for (uint32_t i = 0; i < n_worlds; ++i) dst[i] = worlds[i] ? &worlds[i]->impl : nullptr;
    return guarded(batch, [&](bg::World& x) { return x.batch_writeback(dst.data(), n_worlds); });
}

int bg_batch_size(bg_world* w, uint32_t* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) {
        *out = x.batch_size();
        return BG_OK;
    });
}

int bg_batch_set_persistent(bg_world* w, int enable) {
    return guarded(w, [&](bg::World& x) { return x.set_persistent(enable != 0); });
}

int bg_batch_persistent(bg_world* w, int* active) {
    if (!active) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) {
        *active = x.persistent_active() ? 1 : 0;
        return BG_OK;
    });
}

int bg_autotune(bg_world* w, double delta, uint32_t steps_per_trial, float* cell_scale, float* skin,
                uint32_t* rebuild_interval) {
    return guarded(w, [&](bg::World& x) { return x.autotune(delta, steps_per_trial, cell_scale, skin, rebuild_interval); });
}

int bg_rbd_count(bg_world* w, uint32_t* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) {
        *out = x.rbd_count();
        return BG_OK;
    });
}

int bg_get_rbd(bg_world* w, bg_handle h, bg_rigid_body* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.get_rbd(h, out); });
}

int bg_set_rbd(bg_world* w, bg_handle h, const bg_rigid_body* in) {
    if (!in) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.set_rbd(h, in); });
}

int bg_get_col(bg_world* w, bg_handle h, bg_collider* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.get_col(h, out); });
}

int bg_get_joint(bg_world* w, bg_handle h, bg_fixed_joint* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.get_joint(h, out); });
}

int bg_get_spring(bg_world* w, bg_handle h, bg_spring* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.get_spring(h, out); });
}

int bg_get_bodies(bg_world* w, float* position, float* position_old, float* velocity, float* rotation,
                  float* angular_velocity, bg_handle* handles) {
    return guarded(w, [&](bg::World& x) {
        return x.get_bodies(position, position_old, velocity, rotation, angular_velocity, handles);
    });
}

int bg_get_collider_positions(bg_world* w, float* absolute_translation, uint32_t* count) {
    return guarded(w, [&](bg::World& x) { return x.get_collider_positions(absolute_translation, count); });
}

int bg_contact_count(bg_world* w, uint64_t* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.contact_count(out); });
}

int bg_last_error(bg_world* w, char* buf, size_t len) {
    if (!buf || len == 0) return BG_ERR_INVALID_ARGUMENT;
    const std::string& s = w ? w->impl.last_error : g_last_error;
    size_t n = s.size() < len - 1 ? s.size() : len - 1;
    memcpy(buf, s.data(), n);
    buf[n] = 0;
    return BG_OK;
}

int bg_set_event_capacity(bg_world* w, uint32_t capacity) {
    return guarded(w, [&](bg::World& x) { return x.set_event_capacity(capacity); });
}

int bg_drain_events(bg_world* w, bg_collision_event* out, uint32_t cap, uint32_t* n_out, uint64_t* n_dropped) {
    if (cap > 0 && !out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.drain_events(out, cap, n_out, n_dropped); });
}

int bg_debug_data(bg_world* w, bg_affine2* bodies, bg_affine2* colliders, float* collider_radius, float* joint_bodies,
                  float* spring_bodies, uint32_t counts[4]) {
    return guarded(w, [&](bg::World& x) {
        return x.debug_data(bodies, colliders, collider_radius, joint_bodies, spring_bodies, counts);
    });
}

int bg_set_broadphase(bg_world* w, float cell_scale, float skin, uint32_t rebuild_interval) {
    return guarded(w, [&](bg::World& x) { return x.set_broadphase(cell_scale, skin, rebuild_interval); });
}

int bg_step_path(bg_world* w, uint32_t* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) {
        *out = (uint32_t)x.path();
        return BG_OK;
    });
}

// Internal diagnostic hook (bgx_: not part of the Physics ABI): the world's
// step record {n_outliers, dmax bits, push_outs, error_flags}.
int bgx_world_accum(bg_world* w, uint32_t* out4) {
    return guarded(w, [&](bg::World& x) { return x.debug_accum(out4); });
}

// Internal diagnostic hook: histogram of partner-list lengths (fused sphere
// path): out[0..cap] by length, out[cap + 1] bodies that walk the grid; the
// caller passes cap + 2 entries and learns cap from bgx_list_capacity.
int bgx_list_lengths(bg_world* w, uint32_t* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return guarded(w, [&](bg::World& x) { return x.debug_list_lengths(out); });
}

uint32_t bgx_list_capacity(void) { return bg::kNbrCap; }

int bg_rollout_create(bg_world* const* worlds, uint32_t n_worlds, const float* targets_xy, uint32_t checkpoint_every,
                      bg_rollout** out) {
    if (!out || !worlds || !targets_xy || n_worlds == 0) {
        g_last_error = "bg_rollout_create: worlds, n_worlds > 0, targets and out are required";
        return BG_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    bg_rollout* r = nullptr;
    try {
        r = new bg_rollout();
        std::vector<bg::World*> src(n_worlds);
        for (uint32_t i = 0; i < n_worlds; ++i) src[i] = worlds[i] ? &worlds[i]->impl : nullptr;
        int st = r->impl.create(src.data(), n_worlds, targets_xy, checkpoint_every);
        if (st != BG_OK) {
            g_last_error = r->impl.last_error;
            delete r;
            return st;
        }
    } catch (const std::bad_alloc&) {
        delete r;
        g_last_error = "host out of memory";
        return BG_ERR_CAPACITY;
    } catch (const std::exception& e) {
        delete r;
        g_last_error = e.what();
        return BG_ERR_PANIC;
    }
    *out = r;
    return BG_OK;
}

int bg_rollout_destroy(bg_rollout* r) {
    delete r;
    return BG_OK;
}

int bg_rollout_sizes(bg_rollout* r, uint32_t* n_worlds, uint32_t* n_bodies, uint32_t* n_springs) {
    return rguard(r, [&](bg::Rollout& x) {
        if (n_worlds) *n_worlds = x.n_worlds();
        if (n_bodies) *n_bodies = x.n_bodies();
        if (n_springs) *n_springs = x.n_springs();
        return BG_OK;
    });
}

int bg_rollout_set_params(bg_rollout* r, const float* velocity_request, const float* stiffness) {
    return rguard(r, [&](bg::Rollout& x) { return x.set_params(velocity_request, stiffness); });
}

int bg_rollout_forward(bg_rollout* r, double delta, uint32_t n_steps, float* loss_per_world) {
    return rguard(r, [&](bg::Rollout& x) { return x.forward(delta, n_steps, loss_per_world); });
}

int bg_rollout_backward(bg_rollout* r, const float* dloss_dfinal_position, float* grad_velocity_request,
                        float* grad_stiffness) {
    return rguard(r, [&](bg::Rollout& x) {
        return x.backward(dloss_dfinal_position, grad_velocity_request, grad_stiffness);
    });
}

int bg_rollout_nonfinite(bg_rollout* r, uint8_t* per_world) {
    return rguard(r, [&](bg::Rollout& x) { return x.nonfinite(per_world); });
}

int bg_rollout_final_positions(bg_rollout* r, float* out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    return rguard(r, [&](bg::Rollout& x) { return x.final_positions(out); });
}

int bg_set_deferred_checks(bg_world* w, int enable) {
    return guarded(w, [&](bg::World& x) { return x.set_deferred_checks(enable != 0); });
}

int bg_set_profiling(bg_world* w, int enable) {
    return guarded(w, [&](bg::World& x) { return x.set_profiling(enable != 0); });
}

int bg_stage_stats(bg_world* w, uint32_t stage, uint64_t* launches, double* total_ms) {
    return guarded(w, [&](bg::World& x) { return x.stage_stats(stage, launches, total_ms); });
}

}  // extern "C"
