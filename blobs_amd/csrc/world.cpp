// world.cpp -- host side of the `Physics` drop-in (see world.hpp).
#include "world.hpp"

#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>

namespace bg {

namespace {

constexpr uint32_t kAll = 0xFFFFFFFFu;

bool same_handle(bg_handle a, bg_handle b) { return a == b; }

float col_mass(const bg_collider& c) { return c.has_mass_override ? c.mass_override : c.radius * 2.0f; }

// Collider::inertia, collider.rs:44-50.
float col_inertia(const bg_collider& c) {
    float m = col_mass(c);
    float ox = c.offset_translation[0], oy = c.offset_translation[1];
    float d = sqrtf(ox * ox + oy * oy);
    float inertia = 0.5f * m * (c.radius * c.radius);
    return inertia + m * (d * d);
}

// Steps a single-sphere world stays on the general path after a push-out.
constexpr uint32_t kPushOutHold = 4;

template <class T>
hipError_t dmalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <class T>
hipError_t h2d(T* dst, const std::vector<T>& src, hipStream_t s) {
    if (src.empty()) return hipSuccess;
    return hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

template <class T>
hipError_t d2h(std::vector<T>& dst, const T* src, size_t n, hipStream_t s) {
    dst.resize(n);
    if (n == 0) return hipSuccess;
    return hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s);
}

uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) return fail(BG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

}  // namespace

World::World(float gx, float gy, bool use_spatial_hash) {
    // Physics::new defaults, physics.rs:42-68
    params_.gravity[0] = gx;
    params_.gravity[1] = gy;
    params_.substeps = 8;
    params_.joint_iterations = 4;
    params_.collisions_enabled = 1;
    params_.use_spatial_hash = use_spatial_hash ? 1u : 0u;
    params_.accumulator = 0.0;
    params_.time = 0.0;
    params_.old_dt = 1.0f;
}

World::~World() {
    if (d_.stream) (void)hipStreamSynchronize(d_.stream);
    if (check_event_) (void)hipEventDestroy(check_event_);
    if (host_acc_) (void)hipHostFree(host_acc_);
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    for (auto& v : pending_)
        for (auto& pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    free_device();
}

int World::set_profiling(bool enable) {
    if (!enable && profiling_) {
        int st = collect_stage_times();
        if (st) return st;
    }
    profiling_ = enable;
    for (int k = 0; k < BG_STAGE_COUNT; ++k) {
        stage_launches_[k] = 0;
        stage_ms_[k] = 0.0;
    }
    return BG_OK;
}

int World::debug_accum(uint32_t* out4) {
    out4[0] = out4[1] = out4[2] = out4[3] = 0;
    if (!d_.accum) return BG_OK;
    StepAccum acc;
    HIP_TRY(hipMemcpyAsync(&acc, d_.accum, sizeof(acc), hipMemcpyDeviceToHost, d_.stream));
    HIP_TRY(hipStreamSynchronize(d_.stream));
    out4[0] = acc.n_outliers;
    out4[1] = acc.dmax;
    out4[2] = acc.push_outs;
    out4[3] = acc.error_flags;
    return BG_OK;
}

// Histogram of the partner-list lengths now on the device: [0..kNbrCap]
// bodies with that many entries, [kNbrCap + 1] bodies without a list (they
// walk the grid).
int World::debug_list_lengths(uint32_t* out) {
    for (uint32_t i = 0; i < kNbrCap + 2; ++i) out[i] = 0;
    if (!d_.nbr_len || d_.nb == 0) return BG_OK;
    std::vector<uint8_t> len(d_.nb);
    HIP_TRY(hipMemcpyAsync(len.data(), d_.nbr_len, d_.nb, hipMemcpyDeviceToHost, d_.stream));
    HIP_TRY(hipStreamSynchronize(d_.stream));
    for (uint8_t v : len) ++out[v <= kNbrCap ? v : kNbrCap + 1];
    return BG_OK;
}

int World::stage_stats(uint32_t stage, uint64_t* launches, double* total_ms) {
    if (stage >= BG_STAGE_COUNT) return fail(BG_ERR_INVALID_ARGUMENT, "unknown stage");
    int st = collect_stage_times();
    if (st) return st;
    if (launches) *launches = stage_launches_[stage];
    if (total_ms) *total_ms = stage_ms_[stage];
    return BG_OK;
}

hipError_t World::pooled_event(hipEvent_t* out) {
    if (!event_pool_.empty()) {
        *out = event_pool_.back();
        event_pool_.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(out);
}

int World::stage_begin(uint32_t stage) {
    if (!profiling_) return BG_OK;
    hipEvent_t e;
    HIP_TRY(pooled_event(&e));
    HIP_TRY(hipEventRecord(e, d_.stream));
    open_[stage] = e;
    return BG_OK;
}

int World::stage_end(uint32_t stage) {
    if (!profiling_ || !open_[stage]) return BG_OK;
    hipEvent_t e;
    HIP_TRY(pooled_event(&e));
    HIP_TRY(hipEventRecord(e, d_.stream));
    pending_[stage].push_back({open_[stage], e});
    open_[stage] = nullptr;
    return BG_OK;
}

int World::collect_stage_times() {
    for (int k = 0; k < BG_STAGE_COUNT; ++k) {
        for (auto& pr : pending_[k]) {
            HIP_TRY(hipEventSynchronize(pr.second));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
            stage_ms_[k] += ms;
            stage_launches_[k] += 1;
            event_pool_.push_back(pr.first);
            event_pool_.push_back(pr.second);
        }
        pending_[k].clear();
    }
    return BG_OK;
}

// RigidBody::update_mass_and_inertia, rigid_body.rs:96-128.
void World::update_mass_and_inertia(HostBody& b) {
    float m = 0.0f, I = 0.0f, wx = 0.0f, wy = 0.0f;
    for (bg_handle h : b.colliders) {
        const HostCollider* c = colliders_.get(h);
        if (!c) continue;  // the reference prints "Collider .. not found" and skips
        if (c->f.is_sensor) continue;
        float cm = col_mass(c->f);
        m = m + cm;
        I = I + col_inertia(c->f);
        wx = wx + c->f.offset_translation[0] * cm;
        wy = wy + c->f.offset_translation[1] * cm;
    }
    if (m == 0.0f) m = 1.0f;
    if (I == 0.0f) I = 1.0f;
    b.f.calculated_mass = m;
    b.f.inertia = I;
    b.f.center_of_mass[0] = wx / m;
    b.f.center_of_mass[1] = wy / m;
}

int World::insert_rbd(const bg_rigid_body& b, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    HostBody hb;
    hb.f = b;
    bg_handle h = bodies_.insert(hb);
    if (out) *out = h;
    mark_dirty();
    return BG_OK;
}

int World::insert_collider_with_parent(const bg_collider& c, bg_handle parent, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    HostCollider hc;
    hc.f = c;
    hc.f.parent = parent;  // collider.rs:172
    hc.f.has_parent = 1;
    bg_handle h = colliders_.insert(hc);
    if (out) *out = h;
    mark_dirty();
    HostBody* b = bodies_.get(parent);
    if (!b) return fail(BG_ERR_PANIC, "parent rigid body must exist when inserting collider");  // physics.rs:142
    b->colliders.push_back(h);  // collider.rs:179-181
    b->colliders.push_back(h);  // physics.rs:144
    update_mass_and_inertia(*b);  // physics.rs:146
    return BG_OK;
}

int World::insert_rbds(uint32_t n, const bg_body_batch& bb, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    if (!bb.position) return fail(BG_ERR_INVALID_ARGUMENT, "position array is required");
    bodies_.reserve(bodies_.capacity_slots() + n);
    for (uint32_t i = 0; i < n; ++i) {
        // RigidBodyBuilder::new() defaults (rigid_body.rs:303-318), then .position(p) (:320-324)
        HostBody hb;
        bg_rigid_body& f = hb.f;
        memset(&f, 0, sizeof(f));
        f.position[0] = f.position_old[0] = bb.position[2 * i];
        f.position[1] = f.position_old[1] = bb.position[2 * i + 1];
        f.gravity_mod = bb.gravity_mod ? bb.gravity_mod[i] : 1.0f;
        f.rotation = bb.rotation ? bb.rotation[i] : 0.0f;
        f.scale[0] = f.scale[1] = 1.0f;
        if (bb.velocity_request && (!bb.has_velocity_request || bb.has_velocity_request[i])) {
            f.velocity_request[0] = bb.velocity_request[2 * i];
            f.velocity_request[1] = bb.velocity_request[2 * i + 1];
            f.has_velocity_request = 1;
        }
        f.body_type = bb.body_type ? bb.body_type[i] : BG_DYNAMIC;
        f.calculated_mass = 1.0f;  // build(), rigid_body.rs:385-388
        f.inertia = 1.0f;
        bg_handle h = bodies_.insert(hb);
        if (out) out[i] = h;
    }
    mark_dirty();
    return BG_OK;
}

int World::insert_colliders(uint32_t n, const bg_collider_batch& cb, const bg_handle* parents, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    if (!cb.radius || !parents) return fail(BG_ERR_INVALID_ARGUMENT, "radius and parents are required");
    colliders_.reserve(colliders_.capacity_slots() + n);
    for (uint32_t i = 0; i < n; ++i) {
        // ColliderBuilder::new() defaults (collider.rs:212-224)
        bg_collider c;
        memset(&c, 0, sizeof(c));
        c.offset_matrix[0] = c.offset_matrix[3] = 1.0f;
        c.absolute_matrix[0] = c.absolute_matrix[3] = 1.0f;
        if (cb.offset_translation) {
            c.offset_translation[0] = cb.offset_translation[2 * i];
            c.offset_translation[1] = cb.offset_translation[2 * i + 1];
        }
        if (cb.absolute_translation) {
            c.absolute_translation[0] = cb.absolute_translation[2 * i];
            c.absolute_translation[1] = cb.absolute_translation[2 * i + 1];
        }
        c.radius = cb.radius[i];
        if (cb.has_mass_override && cb.has_mass_override[i]) {
            c.has_mass_override = 1;
            c.mass_override = cb.mass_override ? cb.mass_override[i] : 0.0f;
        }
        c.is_sensor = cb.is_sensor ? cb.is_sensor[i] : 0;
        c.memberships = cb.memberships ? cb.memberships[i] : kAll;
        c.filter = cb.filter ? cb.filter[i] : kAll;
        int st = insert_collider_with_parent(c, parents[i], out ? &out[i] : nullptr);
        if (st != BG_OK) return st;
    }
    return BG_OK;
}

int World::create_fixed_joint(bg_handle a, bg_handle b, const float* aa, const float* ab, bool with_distance,
                              float distance, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    HostBody* ba = bodies_.get(a);
    HostBody* bb = bodies_.get(b);
    // get2_mut(..).zip().unwrap() panics on a missing body or a == b (physics.rs:191-196)
    if (!ba || !bb || same_handle(a, b)) return fail(BG_ERR_PANIC, "create_fixed_joint: bodies must exist and differ");
    if (!with_distance) {
        // (rbd_a.position + anchor_a - rbd_b.position - anchor_b).length(), physics.rs:198
        float x = ((ba->f.position[0] + aa[0]) - bb->f.position[0]) - ab[0];
        float y = ((ba->f.position[1] + aa[1]) - bb->f.position[1]) - ab[1];
        distance = sqrtf(x * x + y * y);
    }
    bg_fixed_joint j;
    j.rigid_body_a = a;
    j.rigid_body_b = b;
    j.anchor_a[0] = aa[0];
    j.anchor_a[1] = aa[1];
    j.anchor_b[0] = ab[0];
    j.anchor_b[1] = ab[1];
    j.distance = distance;
    j.target_angle = bb->f.rotation - ba->f.rotation;  // physics.rs:230
    bg_handle h = joints_.insert(j);
    ba->joints.push_back(h);
    bb->joints.push_back(h);
    if (out) *out = h;
    mark_dirty();
    return BG_OK;
}

int World::insert_spring(const bg_spring& s, bg_handle* out) {
    if (sync_host()) return BG_ERR_HIP;
    bg_handle h = springs_.insert(s);
    if (out) *out = h;
    mark_dirty();
    return BG_OK;
}

int World::push_constraint(float x, float y, float r) {
    if (sync_host()) return BG_ERR_HIP;
    constraints_.push_back({x, y, r});
    mark_dirty();
    return BG_OK;
}

// Physics::remove_rbd, physics.rs:163-172.
int World::remove_rbd(bg_handle h) {
    if (sync_host()) return BG_ERR_HIP;
    HostBody* b = bodies_.get(h);
    if (b) {
        for (bg_handle c : b->colliders) colliders_.remove(c);  // remove_ignoring_parent
    }
    bodies_.remove(h);
    mark_dirty();
    return BG_OK;
}

// ColliderSet::remove, collider.rs:134-164.
int World::remove_col(bg_handle h) {
    if (sync_host()) return BG_ERR_HIP;
    HostCollider* c = colliders_.get(h);
    if (c && c->f.has_parent) {
        bg_handle parent = c->f.parent;
        HostBody* b = bodies_.get(parent);
        bool remove_body = false;
        if (b) {
            b->colliders.erase(std::remove(b->colliders.begin(), b->colliders.end(), h), b->colliders.end());
            update_mass_and_inertia(*b);
            if (b->colliders.empty()) remove_body = true;
        }
        if (remove_body) bodies_.remove(parent);
    }
    colliders_.remove(h);
    mark_dirty();
    return BG_OK;
}

int World::reset() {
    if (sync_host()) return BG_ERR_HIP;
    bodies_.clear();
    colliders_.clear();
    joints_.clear();
    springs_.clear();
    mark_dirty();
    return BG_OK;
}

int World::get_params(bg_params* p) {
    if (host_stale_) {
        // old_dt lives on the device between steps
        if (d_.accum) {
            StepAccum acc;
            HIP_TRY(hipMemcpyAsync(&acc, d_.accum, sizeof(acc), hipMemcpyDeviceToHost, d_.stream));
            HIP_TRY(hipStreamSynchronize(d_.stream));
            params_.old_dt = acc.old_dt;
        }
    }
    *p = params_;
    return BG_OK;
}

int World::set_params(const bg_params* p) {
    float old = params_.old_dt;
    params_ = *p;
    if (d_.accum && !dev_dirty_ && __builtin_memcmp(&old, &p->old_dt, 4) != 0) {
        HIP_TRY(hipMemcpyAsync(&d_.accum->old_dt, &params_.old_dt, sizeof(float), hipMemcpyHostToDevice, d_.stream));
        HIP_TRY(hipStreamSynchronize(d_.stream));
    }
    return BG_OK;
}

int World::get_rbd(bg_handle h, bg_rigid_body* out) {
    if (sync_host()) return BG_ERR_HIP;
    HostBody* b = bodies_.get(h);
    if (!b) return fail(BG_ERR_NOT_FOUND, "rigid body not found");
    *out = b->f;
    return BG_OK;
}

int World::set_rbd(bg_handle h, const bg_rigid_body* in) {
    if (sync_host()) return BG_ERR_HIP;
    HostBody* b = bodies_.get(h);
    if (!b) return fail(BG_ERR_NOT_FOUND, "rigid body not found");
    b->f = *in;
    mark_dirty();
    return BG_OK;
}

int World::get_joint(bg_handle h, bg_fixed_joint* out) {
    const bg_fixed_joint* j = joints_.get(h);
    if (!j) return fail(BG_ERR_NOT_FOUND, "joint not found");
    *out = *j;
    return BG_OK;
}

int World::get_spring(bg_handle h, bg_spring* out) {
    const bg_spring* s = springs_.get(h);
    if (!s) return fail(BG_ERR_NOT_FOUND, "spring not found");
    *out = *s;
    return BG_OK;
}

int World::get_col(bg_handle h, bg_collider* out) {
    if (sync_host()) return BG_ERR_HIP;
    HostCollider* c = colliders_.get(h);
    if (!c) return fail(BG_ERR_NOT_FOUND, "collider not found");
    *out = c->f;
    return BG_OK;
}

int World::get_bodies(float* pos, float* pos_old, float* vel, float* rot, float* omega, bg_handle* handles) {
    if (sync_host()) return BG_ERR_HIP;
    uint32_t k = 0;
    for (uint32_t s = 0; s < bodies_.capacity_slots(); ++s) {
        if (!bodies_.occupied(s)) continue;
        const bg_rigid_body& f = bodies_.at_slot(s).f;
        if (pos) {
            pos[2 * k] = f.position[0];
            pos[2 * k + 1] = f.position[1];
        }
        if (pos_old) {
            pos_old[2 * k] = f.position_old[0];
            pos_old[2 * k + 1] = f.position_old[1];
        }
        if (vel) {
            vel[2 * k] = f.calculated_velocity[0];
            vel[2 * k + 1] = f.calculated_velocity[1];
        }
        if (rot) rot[k] = f.rotation;
        if (omega) omega[k] = f.angular_velocity;
        if (handles) handles[k] = bodies_.handle_at(s);
        ++k;
    }
    return BG_OK;
}

int World::get_collider_positions(float* abs, uint32_t* count) {
    if (sync_host()) return BG_ERR_HIP;
    uint32_t k = 0;
    for (uint32_t s = 0; s < colliders_.capacity_slots(); ++s) {
        if (!colliders_.occupied(s)) continue;
        if (abs) {
            abs[2 * k] = colliders_.at_slot(s).f.absolute_translation[0];
            abs[2 * k + 1] = colliders_.at_slot(s).f.absolute_translation[1];
        }
        ++k;
    }
    if (count) *count = k;
    return BG_OK;
}

int World::device_contacts(uint64_t* out) {
    *out = 0;
    if (!d_.block_contacts || path_ == StepPath::None) return BG_OK;
    std::vector<unsigned long long> slots;
    HIP_TRY(d2h(slots, d_.block_contacts, d_.n_block_slots, d_.stream));
    HIP_TRY(hipStreamSynchronize(d_.stream));
    uint64_t s = 0;
    for (unsigned long long v : slots) s += v;
    *out = s;
    return BG_OK;
}

int World::contact_count(uint64_t* out) {
    uint64_t dev = 0;
    int st = device_contacts(&dev);
    if (st) return st;
    *out = contact_base_ + dev;
    return BG_OK;
}

int World::set_event_capacity(uint32_t cap) {
    if (sync_host()) return BG_ERR_HIP;
    if (int st = collect_events()) return st;
    event_cap_ = cap;
    mark_dirty();  // buffers are sized at upload
    return BG_OK;
}

// Device events -> host queue, in the reference's order.
int World::collect_events() {
    DeviceWorld& d = d_;
    if (!d.allocated || !d.events) return BG_OK;
    unsigned long long n = 0;
    HIP_TRY(hipMemcpyAsync(&n, &d.accum->event_count, sizeof(n), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    if (n == 0) return BG_OK;
    const uint32_t kept = n < d.event_cap ? (uint32_t)n : d.event_cap;
    std::vector<DevEvent> ev;
    HIP_TRY(d2h(ev, d.events, kept, d.stream));
    HIP_TRY(hipMemsetAsync(&d.accum->event_count, 0, sizeof(n), d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    // (substep, A, B) is the reference's send order (physics.rs:248-249, 304-311); a
    // withdrawal record (pad 1, from the ordered contact pass) sorts right after the
    // record it cancels
    std::sort(ev.begin(), ev.end(), [](const DevEvent& x, const DevEvent& y) {
        if (x.sub != y.sub) return x.sub < y.sub;
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        return x.pad < y.pad;
    });
    // Records are reserved substep after substep, so a full buffer holds every event of
    // the substeps before the one that overflowed and an arbitrary part of that one:
    // that substep's records are dropped too, so the kept set does not depend on the
    // order of the device's atomics.
    uint64_t dropped = n - kept;
    size_t end = ev.size();
    if (n > kept && end > 0) {
        const uint32_t last = ev[end - 1].sub;
        while (end > 0 && ev[end - 1].sub == last) --end;
        for (size_t i = end; i < ev.size(); ++i) dropped += ev[i].pad ? 0u : 1u;
    }
    events_dropped_ += dropped;
    if (event_head_ > 0 && event_head_ == event_queue_.size()) {
        event_queue_.clear();
        event_head_ = 0;
    }
    for (size_t i = 0; i < end; ++i) {
        const DevEvent& e = ev[i];
        if (e.pad) continue;  // a withdrawal whose record is not here
        if (i + 1 < end && ev[i + 1].pad && ev[i + 1].sub == e.sub && ev[i + 1].a == e.a && ev[i + 1].b == e.b) {
            ++i;  // withdrawn by the ordered contact pass
            continue;
        }
        bg_collision_event o;
        o.col_handle_a = colliders_.handle_at(col_slot_of_dense_[e.a]);
        o.col_handle_b = colliders_.handle_at(col_slot_of_dense_[e.b]);
        o.impact_vel_a[0] = e.va.x;
        o.impact_vel_a[1] = e.va.y;
        o.impact_vel_b[0] = e.vb.x;
        o.impact_vel_b[1] = e.vb.y;
        event_queue_.push_back(o);
    }
    return BG_OK;
}

int World::drain_events(bg_collision_event* out, uint32_t cap, uint32_t* n_out, uint64_t* n_dropped) {
    if (int st = collect_events()) return st;
    const size_t avail = event_queue_.size() - event_head_;
    const size_t n = avail < cap ? avail : cap;
    if (n && out) memcpy(out, event_queue_.data() + event_head_, n * sizeof(bg_collision_event));
    event_head_ += n;
    if (event_head_ == event_queue_.size()) {
        event_queue_.clear();
        event_head_ = 0;
    }
    if (n_out) *n_out = (uint32_t)n;
    if (n_dropped) *n_dropped = events_dropped_;
    events_dropped_ = 0;
    return BG_OK;
}

// Physics::debug_data, debug.rs:34-91 (from the host mirror).
int World::debug_data(bg_affine2* bodies, bg_affine2* cols, float* radius, float* joints, float* springs,
                      uint32_t* counts) {
    if (sync_host()) return BG_ERR_HIP;
    uint32_t nb = 0, nc = 0, nj = 0, ns = 0;
    for (uint32_t s = 0; s < bodies_.capacity_slots(); ++s) {
        if (!bodies_.occupied(s)) continue;
        const bg_rigid_body& f = bodies_.at_slot(s).f;
        if (bodies) {
            // Affine2::from_angle_translation: Mat2::from_angle = cols (cos, sin), (-sin, cos)
            const float sn = sinf(f.rotation), cs = cosf(f.rotation);
            bg_affine2& t = bodies[nb];
            t.matrix[0] = cs;
            t.matrix[1] = sn;
            t.matrix[2] = -sn;
            t.matrix[3] = cs;
            t.translation[0] = f.position[0];
            t.translation[1] = f.position[1];
        }
        ++nb;
    }
    for (uint32_t s = 0; s < colliders_.capacity_slots(); ++s) {
        if (!colliders_.occupied(s)) continue;
        const bg_collider& c = colliders_.at_slot(s).f;
        if (cols) {
            memcpy(cols[nc].matrix, c.absolute_matrix, sizeof(cols[nc].matrix));
            memcpy(cols[nc].translation, c.absolute_translation, sizeof(cols[nc].translation));
        }
        if (radius) radius[nc] = c.radius;
        ++nc;
    }
    auto ends = [&](bg_handle a, bg_handle b, float* out4) -> bool {
        const HostBody* ba = bodies_.get(a);
        const HostBody* bb = bodies_.get(b);
        if (!ba || !bb) return false;  // the reference indexes the arena and panics
        out4[0] = ba->f.position[0];
        out4[1] = ba->f.position[1];
        out4[2] = bb->f.position[0];
        out4[3] = bb->f.position[1];
        return true;
    };
    float tmp[4];
    for (uint32_t s = 0; s < joints_.capacity_slots(); ++s) {
        if (!joints_.occupied(s)) continue;
        const bg_fixed_joint& j = joints_.at_slot(s);
        if (!ends(j.rigid_body_a, j.rigid_body_b, joints ? joints + 4 * nj : tmp))
            return fail(BG_ERR_PANIC, "joint body missing (debug.rs:48-49 index)");
        ++nj;
    }
    for (uint32_t s = 0; s < springs_.capacity_slots(); ++s) {
        if (!springs_.occupied(s)) continue;
        const bg_spring& sp = springs_.at_slot(s);
        if (!ends(sp.rigid_body_a, sp.rigid_body_b, springs ? springs + 4 * ns : tmp))
            return fail(BG_ERR_PANIC, "spring body missing (debug.rs:79-80 index)");
        ++ns;
    }
    if (counts) {
        counts[0] = nb;
        counts[1] = nc;
        counts[2] = nj;
        counts[3] = ns;
    }
    return BG_OK;
}

int World::set_broadphase(float cell_scale, float skin, uint32_t rebuild_interval) {
    if (!(cell_scale >= 0.25f) || !(skin > 0.0f && skin <= 4.0f) || rebuild_interval == 0)
        return fail(BG_ERR_INVALID_ARGUMENT,
                    "cell_scale must be >= 0.25, skin in (0, 4] and rebuild_interval >= 1");
    cell_scale_ = cell_scale;
    skin_scale_ = skin;
    rebuild_interval_ = rebuild_interval;
    need_rebuild_ = true;
    return BG_OK;
}

// ------------------------------------------------------------------ device

int World::ensure_device() {
    if (d_.stream) return BG_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(BG_ERR_NO_DEVICE, "no HIP device available: the GPU step has no CPU fallback");
    HIP_TRY(hipGetDevice(&d_.device));
    HIP_TRY(hipStreamCreateWithFlags(&d_.stream, hipStreamNonBlocking));
    return BG_OK;
}

void World::free_arrays() {
    DeviceWorld& d = d_;
    if (d.p_stream) (void)hipStreamSynchronize(d.p_stream);
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    {
        BodyArrays* b = &d.b;
        dfree(b->pos), dfree(b->pos_old), dfree(b->vel), dfree(b->acc), dfree(b->vreq);
        dfree(b->rot), dfree(b->omega), dfree(b->torque), dfree(b->inertia), dfree(b->mass), dfree(b->gmod);
        dfree(b->bflags), dfree(b->aid);
        ColliderArrays* c = &d.c;
        dfree(c->abs), dfree(c->off), dfree(c->radius), dfree(c->parent), dfree(c->groups), dfree(c->cflags);
    }
    for (SphereArrays* q : {&d.s, &d.s_alt}) {
        dfree(q->pp), dfree(q->snap), dfree(q->vel), dfree(q->rw), dfree(q->mg), dfree(q->bflags), dfree(q->off);
        dfree(q->groups), dfree(q->inertia), dfree(q->torque), dfree(q->acc), dfree(q->vreq);
    }
    for (int i = 0; i < 3; ++i) dfree(d.oh_head[i]), dfree(d.oh_next[i]), dfree(d.oh_bits[i]), dfree(d.oh_wave[i]);
    dfree(d.cell_count), dfree(d.perm);
    dfree(d.wid), dfree(d.cell_base), dfree(d.tile_seg), dfree(d.grids), dfree(d.wbbox), dfree(d.wold_dt), dfree(d.wstat), dfree(d.wvel);
    dfree(d.p_order), dfree(d.p_wstart), dfree(d.p_counter), dfree(d.p_skip);
    for (uint32_t c = 0; c < kPersistClasses; ++c) d.p_groups[c] = 0;
    d.p_grid_worlds = 0;
    dfree(d.snap_next), dfree(d.x0), dfree(d.bc_start), dfree(d.bc_list), dfree(d.nbr), dfree(d.nbr_len);
    dfree(d.sp_a), dfree(d.sp_b), dfree(d.sp_rest), dfree(d.sp_k), dfree(d.sp_c), dfree(d.bs_start),
        dfree(d.bs_list);
    dfree(d.j_a), dfree(d.j_b), dfree(d.j_anchor_a), dfree(d.j_anchor_b), dfree(d.j_dist), dfree(d.j_target);
    dfree(d.j_mass), dfree(d.jrec), dfree(d.bj_start), dfree(d.bj_list), dfree(d.bj_comp);
    dfree(d.jc_start), dfree(d.jc_list), dfree(d.jb_start), dfree(d.jb_list), dfree(d.j_local), dfree(d.jq_anchor), dfree(d.jq_mass), dfree(d.jq_len), dfree(d.con_pos),
        dfree(d.con_r);
    dfree(d.keys), dfree(d.vals), dfree(d.keys_tmp), dfree(d.vals_tmp), dfree(d.cell_start), dfree(d.crec);
    dfree(d.grid), dfree(d.accum), dfree(d.block_contacts), dfree(d.events), dfree(d.vel_next);
    dfree(d.ord_rows), dfree(d.ord_n), dfree(d.ord_mark), dfree(d.pos_pre), dfree(d.mv_key), dfree(d.mv_abs);
    dfree(d.mv_list), dfree(d.ord_heap), dfree(d.ord_log), dfree(d.aff_list), dfree(d.aff_mark), dfree(d.ord_cand);
    dfree(d.ord_terms), dfree(d.clist), dfree(d.ccnt), dfree(d.cfb_list), dfree(d.cfb_n);
    sort_temp_free(d.sort);
    hipStream_t s = d.stream, ps = d.p_stream;
    hipEvent_t pf = d.p_fork, pj = d.p_join;
    int dev = d.device;
    d = DeviceWorld{};
    d.stream = s;
    d.p_stream = ps;
    d.p_fork = pf;
    d.p_join = pj;
    d.device = dev;
}

void World::free_device() {
    free_arrays();
    if (d_.p_stream) (void)hipStreamSynchronize(d_.p_stream);
    if (d_.p_fork) (void)hipEventDestroy(d_.p_fork);
    if (d_.p_join) (void)hipEventDestroy(d_.p_join);
    if (d_.p_stream) (void)hipStreamDestroy(d_.p_stream);
    if (d_.stream) (void)hipStreamDestroy(d_.stream);
    d_.stream = d_.p_stream = nullptr;
    d_.p_fork = d_.p_join = nullptr;
}

// Whether two colliders of the snapshot pass the reference predicate with
// coincident centres (physics.rs:251-272: parents exist and differ, groups
// test, dist < r_a + r_b and dist < 1e-6), i.e. whether the next contact pass
// takes the push-out branch.  Same f32 arithmetic as the device; candidates
// from a sort by x (coincident centres are within 1e-6 in x).
static bool has_coincident_pair(const std::vector<float2>& abs, const std::vector<float>& radius,
                                const std::vector<int32_t>& parent, const std::vector<uint2>& groups,
                                bool use_groups) {
    const size_t n = abs.size();
    std::vector<uint32_t> idx;
    idx.reserve(n);
    for (uint32_t i = 0; i < n; ++i)
        if (std::isfinite(abs[i].x) && std::isfinite(abs[i].y) && parent[i] >= 0) idx.push_back(i);
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return abs[a].x < abs[b].x; });
    for (size_t k = 0; k < idx.size(); ++k) {
        const uint32_t a = idx[k];
        for (size_t q = k + 1; q < idx.size() && abs[idx[q]].x - abs[a].x < 1e-6f; ++q) {
            const uint32_t b = idx[q];
            if (parent[a] == parent[b]) continue;
            if (use_groups && !((groups[a].x & groups[b].y) != 0u && (groups[b].x & groups[a].y) != 0u)) continue;
            const float dx = abs[a].x - abs[b].x, dy = abs[a].y - abs[b].y;
            const float dist = sqrtf(dx * dx + dy * dy);
            if (dist < radius[a] + radius[b] && dist < 1e-6f) return true;
        }
    }
    return false;
}

// Flatten the arenas into dense SoA (slot order without holes = the
// reference's iteration order) and copy them to the device.
int World::upload() {
    if (d_.stream) HIP_TRY(hipStreamSynchronize(d_.stream));
    // keep the contact counter and pending collision events across re-uploads
    {
        uint64_t dev = 0;
        int st = device_contacts(&dev);
        if (st) return st;
        contact_base_ += dev;
        st = collect_events();
        if (st) return st;
    }
    free_arrays();

    structure_error_ = false;
    structure_msg_.clear();

    // dense maps
    const uint32_t nb = bodies_.len(), nc = colliders_.len();
    body_slot_of_dense_.clear();
    col_slot_of_dense_.clear();
    body_dense_of_slot_.assign(bodies_.capacity_slots(), -1);
    col_dense_of_slot_.assign(colliders_.capacity_slots(), -1);
    for (uint32_t s = 0; s < bodies_.capacity_slots(); ++s)
        if (bodies_.occupied(s)) {
            body_dense_of_slot_[s] = (int32_t)body_slot_of_dense_.size();
            body_slot_of_dense_.push_back(s);
        }
    for (uint32_t s = 0; s < colliders_.capacity_slots(); ++s)
        if (colliders_.occupied(s)) {
            col_dense_of_slot_[s] = (int32_t)col_slot_of_dense_.size();
            col_slot_of_dense_.push_back(s);
        }
    auto body_dense = [&](bg_handle h) -> int32_t {
        if (!bodies_.get(h)) return -1;
        return body_dense_of_slot_[Arena<HostBody>::slot_of(h)];
    };
    auto col_dense = [&](bg_handle h) -> int32_t {
        if (!colliders_.get(h)) return -1;
        return col_dense_of_slot_[Arena<HostCollider>::slot_of(h)];
    };

    // bodies
    std::vector<float2> pos(nb), pos_old(nb), vel(nb), acc(nb), vreq(nb);
    std::vector<float> rot(nb), omega(nb), torque(nb), inertia(nb), mass(nb), gmod(nb);
    std::vector<uint32_t> bflags(nb), aid(nb);
    std::vector<uint32_t> bc_start(nb + 1, 0), bc_list;
    bc_list.reserve(nc);
    first_dyn_ = -1;
    for (uint32_t i = 0; i < nb; ++i) {
        const HostBody& hb = bodies_.at_slot(body_slot_of_dense_[i]);
        const bg_rigid_body& f = hb.f;
        pos[i] = make_float2(f.position[0], f.position[1]);
        pos_old[i] = make_float2(f.position_old[0], f.position_old[1]);
        vel[i] = make_float2(f.calculated_velocity[0], f.calculated_velocity[1]);
        acc[i] = make_float2(f.acceleration[0], f.acceleration[1]);
        vreq[i] = make_float2(f.velocity_request[0], f.velocity_request[1]);
        rot[i] = f.rotation;
        omega[i] = f.angular_velocity;
        torque[i] = f.torque;
        inertia[i] = f.inertia;
        mass[i] = f.calculated_mass;
        gmod[i] = f.gravity_mod;
        uint32_t fl = 0;
        if (f.body_type == BG_STATIC) fl |= BF_STATIC;
        if (f.has_velocity_request) fl |= BF_HAS_VREQ;
        if (f.inertia > 0.0f && isfinite(f.inertia)) fl |= BF_INERTIA_POS;
        bflags[i] = fl;
        aid[i] = i;
        if (first_dyn_ < 0 && f.body_type != BG_STATIC) first_dyn_ = (int32_t)i;
        // colliders of this body, first occurrence order, duplicates dropped
        for (bg_handle ch : hb.colliders) {
            int32_t cd = col_dense(ch);
            if (cd < 0) continue;
            bool dup = false;
            for (uint32_t k = bc_start[i]; k < bc_list.size(); ++k)
                if (bc_list[k] == (uint32_t)cd) dup = true;
            if (!dup) bc_list.push_back((uint32_t)cd);
        }
        bc_start[i + 1] = (uint32_t)bc_list.size();
    }

    // colliders
    std::vector<float2> cabs(nc), coff(nc);
    std::vector<float> crad(nc);
    std::vector<int32_t> cpar(nc);
    std::vector<uint2> cgrp(nc);
    std::vector<uint32_t> cfl(nc);
    rmax_ = 0.0f;
    groups_ = false;
    sensors_ = false;
    zero_off_ = true;
    float minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
    for (uint32_t i = 0; i < nc; ++i) {
        const bg_collider& f = colliders_.at_slot(col_slot_of_dense_[i]).f;
        cabs[i] = make_float2(f.absolute_translation[0], f.absolute_translation[1]);
        coff[i] = make_float2(f.offset_translation[0], f.offset_translation[1]);
        uint32_t ox, oy;
        memcpy(&ox, &f.offset_translation[0], 4);
        memcpy(&oy, &f.offset_translation[1], 4);
        if (ox != 0u || oy != 0u) zero_off_ = false;
        crad[i] = f.radius;
        if (!(f.radius <= rmax_)) rmax_ = f.radius;  // NaN radius propagates to a full scan
        cpar[i] = f.has_parent ? body_dense(f.parent) : -1;
        cgrp[i] = make_uint2(f.memberships, f.filter);
        if (f.memberships != kAll || f.filter != kAll) groups_ = true;
        cfl[i] = f.is_sensor ? 1u : 0u;
        if (f.is_sensor) sensors_ = true;
        float x = cabs[i].x, y = cabs[i].y;
        if (isfinite(x) && isfinite(y)) {
            minx = fminf(minx, x);
            maxx = fmaxf(maxx, x);
            miny = fminf(miny, y);
            maxy = fmaxf(maxy, y);
        }
    }

    // springs (arena order) and body -> spring lists
    const uint32_t ns = springs_.len();
    std::vector<int32_t> sp_a, sp_b;
    std::vector<float> sp_rest, sp_k, sp_c;
    std::vector<uint32_t> bs_count(nb + 1, 0), bs_start(nb + 1, 0), bs_list;
    for (uint32_t s = 0; s < springs_.capacity_slots(); ++s) {
        if (!springs_.occupied(s)) continue;
        const bg_spring& sp = springs_.at_slot(s);
        int32_t a = body_dense(sp.rigid_body_a), b = body_dense(sp.rigid_body_b);
        if (a < 0 || b < 0 || a == b) {
            structure_error_ = true;
            structure_msg_ = "spring endpoints must be two existing, distinct bodies (springs.rs:26-29 unwrap)";
            a = b = 0;
        }
        sp_a.push_back(a);
        sp_b.push_back(b);
        sp_rest.push_back(sp.rest_length);
        sp_k.push_back(sp.stiffness);
        sp_c.push_back(sp.damping);
    }
    if (!structure_error_ && ns > 0) {
        for (uint32_t s = 0; s < ns; ++s) {
            bs_count[sp_a[s]]++;
            bs_count[sp_b[s]]++;
        }
        for (uint32_t i = 0; i < nb; ++i) bs_start[i + 1] = bs_start[i] + bs_count[i];
        bs_list.resize(bs_start[nb]);
        std::vector<uint32_t> fillp(bs_start.begin(), bs_start.end() - 1);
        for (uint32_t s = 0; s < ns; ++s) {
            bs_list[fillp[sp_a[s]]++] = s;
            bs_list[fillp[sp_b[s]]++] = s;
        }
    }

    // joints (arena order), grouped into components of the body graph
    const uint32_t nj = joints_.len();
    std::vector<int32_t> j_a, j_b;
    std::vector<float2> j_an_a, j_an_b;
    std::vector<float> j_dist, j_target;
    for (uint32_t s = 0; s < joints_.capacity_slots(); ++s) {
        if (!joints_.occupied(s)) continue;
        const bg_fixed_joint& j = joints_.at_slot(s);
        int32_t a = body_dense(j.rigid_body_a), b = body_dense(j.rigid_body_b);
        if (a < 0 || b < 0 || a == b) {
            structure_error_ = true;
            structure_msg_ = "joint bodies must exist and differ (physics.rs:427-432 unwrap)";
            a = b = 0;
        }
        j_a.push_back(a);
        j_b.push_back(b);
        j_an_a.push_back(make_float2(j.anchor_a[0], j.anchor_a[1]));
        j_an_b.push_back(make_float2(j.anchor_b[0], j.anchor_b[1]));
        j_dist.push_back(j.distance);
        j_target.push_back(j.target_angle);
    }
    // mass terms of each joint (physics.rs:447-460), from the masses of this upload
    std::vector<float4> j_mass(nj);
    for (uint32_t k = 0; k < nj; ++k) {
        const float ma = mass[j_a[k]], mb = mass[j_b[k]];
        const float inv_a = 1.0f / ma, inv_b = 1.0f / mb;
        const float ims = inv_a + inv_b;
        const float ratio = inv_a / ims;
        j_mass[k] = make_float4(ims, ratio, 1.0f - ratio, (ma > 0.0f && mb > 0.0f) ? 1.0f : 0.0f);
    }
    std::vector<uint32_t> jc_start, jc_list, jb_start, j_local;
    std::vector<int32_t> jb_list;
    uint32_t n_jcomp = 0;
    if (nj > 0 && !structure_error_) {
        std::vector<int32_t> uf(nb);
        std::iota(uf.begin(), uf.end(), 0);
        auto find = [&](int32_t x) {
            while (uf[x] != x) {
                uf[x] = uf[uf[x]];
                x = uf[x];
            }
            return x;
        };
        for (uint32_t k = 0; k < nj; ++k) {
            int32_t ra = find(j_a[k]), rb = find(j_b[k]);
            if (ra != rb) uf[ra] = rb;
        }
        std::vector<int32_t> comp_of_root(nb, -1);
        std::vector<uint32_t> jcomp(nj), counts;
        for (uint32_t k = 0; k < nj; ++k) {
            int32_t r = find(j_a[k]);
            if (comp_of_root[r] < 0) {
                comp_of_root[r] = (int32_t)counts.size();
                counts.push_back(0);
            }
            jcomp[k] = (uint32_t)comp_of_root[r];
            counts[jcomp[k]]++;
        }
        n_jcomp = (uint32_t)counts.size();
        jc_start.assign(n_jcomp + 1, 0);
        for (uint32_t c = 0; c < n_jcomp; ++c) jc_start[c + 1] = jc_start[c] + counts[c];
        jc_list.resize(nj);
        std::vector<uint32_t> fillp(jc_start.begin(), jc_start.end() - 1);
        for (uint32_t k = 0; k < nj; ++k) jc_list[fillp[jcomp[k]]++] = k;
        // local body indices of the components small enough to iterate in LDS
        jb_start.assign(n_jcomp + 1, 0);
        j_local.assign(nj, 0);
        std::vector<int32_t> local_of(nb, -1), bodies;
        for (uint32_t c = 0; c < n_jcomp; ++c) {
            bodies.clear();
            for (uint32_t kk = jc_start[c]; kk < jc_start[c + 1]; ++kk) {
                const uint32_t j = jc_list[kk];
                for (int32_t bi : {j_a[j], j_b[j]})
                    if (local_of[bi] < 0) {
                        local_of[bi] = (int32_t)bodies.size();
                        bodies.push_back(bi);
                    }
                j_local[kk] = (uint32_t)(local_of[j_a[j]] & 0xFF) | ((uint32_t)(local_of[j_b[j]] & 0xFF) << 8);
            }
            const bool fits = jc_start[c + 1] - jc_start[c] <= kJointJoints && bodies.size() <= kJointBodies;
            for (int32_t bi : bodies) {
                local_of[bi] = -1;
                if (fits) jb_list.push_back(bi);
            }
            jb_start[c + 1] = (uint32_t)jb_list.size();
        }
    }
    // the joints' constant records in component order (entry kk of jc_list), for the joint
    // stage's coalesced staging
    std::vector<float4> jq_anchor(jc_list.size()), jq_mass(jc_list.size());
    std::vector<float2> jq_len(jc_list.size());
    for (size_t kk = 0; kk < jc_list.size(); ++kk) {
        const uint32_t j = jc_list[kk];
        jq_anchor[kk] = make_float4(j_an_a[j].x, j_an_a[j].y, j_an_b[j].x, j_an_b[j].y);
        jq_mass[kk] = j_mass[j];
        jq_len[kk] = make_float2(j_dist[j], j_target[j]);
    }
    // per body: its joints in arena order and its component (deferred joint rotations)
    std::vector<uint32_t> bj_start(nb + 1, 0), bj_list;
    std::vector<int32_t> bj_comp(nb, -1);
    if (n_jcomp > 0) {
        for (uint32_t k = 0; k < nj; ++k) {
            bj_start[j_a[k] + 1]++;
            bj_start[j_b[k] + 1]++;
        }
        for (uint32_t i = 0; i < nb; ++i) bj_start[i + 1] += bj_start[i];
        bj_list.resize(bj_start[nb]);
        std::vector<uint32_t> fillb(bj_start.begin(), bj_start.end() - 1);
        for (uint32_t c = 0; c < n_jcomp; ++c)
            for (uint32_t kk = jc_start[c]; kk < jc_start[c + 1]; ++kk) {
                const uint32_t j = jc_list[kk];
                bj_list[fillb[j_a[j]]++] = kk << 1;
                bj_list[fillb[j_b[j]]++] = (kk << 1) | 1u;
                bj_comp[j_a[j]] = bj_comp[j_b[j]] = (int32_t)c;
            }
    }

    // constraints
    const uint32_t nk = (uint32_t)constraints_.size();
    std::vector<float2> con_pos(nk);
    std::vector<float> con_r(nk);
    for (uint32_t k = 0; k < nk; ++k) {
        con_pos[k] = make_float2(constraints_[k].x, constraints_[k].y);
        con_r[k] = constraints_[k].r;
    }

    // path: one collider per body in the same dense order, no springs/joints
    bool sphere = (nb == nc) && ns == 0 && nj == 0 && nb > 0;
    for (uint32_t i = 0; sphere && i < nb; ++i)
        if (bc_start[i + 1] - bc_start[i] != 1 || bc_list[bc_start[i]] != i || cpar[i] != (int32_t)i) sphere = false;
    // A push-out needs the ordered contact pass, which the general path carries: a
    // single-sphere world takes that path while a pair of its snapshot coincides (or
    // while push-outs kept happening in the last steps).
    if (sphere && batch_start_.empty() && params_.collisions_enabled &&
        (push_out_hold_ > 0 || has_coincident_pair(cabs, crad, cpar, cgrp, groups_))) {
        sphere = false;
        if (push_out_hold_ == 0) push_out_hold_ = kPushOutHold;
    }
    path_ = sphere ? StepPath::Sphere : StepPath::General;
    if (sphere) {
        for (uint32_t i = 0; i < nb; ++i) {
            if (cfl[i] & 1u) bflags[i] |= BF_SENSOR;
            uint32_t ob, rb;
            memcpy(&ob, &omega[i], 4);
            memcpy(&rb, &rot[i], 4);
            if ((bflags[i] & BF_STATIC) ||
                (ob == 0u && rb != 0x80000000u && std::isfinite(rot[i]) && (bflags[i] & BF_INERTIA_POS)))
                bflags[i] |= BF_ROT_STILL;
        }
    }
    // common drift for the broadphase query centre: worth it when every
    // moving body feels the same gravity
    {
        uint32_t dyn = 0;
        bool uniform = true;
        for (uint32_t i = 0; i < nb; ++i) {
            if (bflags[i] & BF_STATIC) continue;
            ++dyn;
            if (gmod[i] != 1.0f) uniform = false;
        }
        use_drift_ = uniform && dyn * 2 >= nb;
    }

    // allocate
    DeviceWorld& d = d_;
    d.nb = nb;
    d.nc = nc;
    d.ns = ns;
    d.nj = nj;
    d.nk = nk;
    d.n_jcomp = n_jcomp;
    d.cells_cap = next_pow2(std::max<uint32_t>(4u * nc, 1024u));
    d.key_bits = 0;
    while ((1u << d.key_bits) < d.cells_cap) ++d.key_bits;
    hipStream_t s = d.stream;
    uint32_t cell_rows = d.cells_cap;  // entries of the cell table (+1)
    if (!batch_start_.empty() && (!sphere || batch_start_.back() != nb)) {
        structure_error_ = true;
        structure_msg_ = "a batch holds single-sphere worlds only (one collider per body, no springs or joints)";
    }
    if (sphere) {
        // worlds: one for a plain world; a batch keeps its source worlds' body ranges
        std::vector<uint32_t> wstart = batch_start_.empty() ? std::vector<uint32_t>{0u, nb} : batch_start_;
        const uint32_t W = (uint32_t)wstart.size() - 1;
        std::vector<uint32_t> cell_base(W + 1, 0), wid;
        std::vector<float> wold(W, params_.old_dt);
        if (!batch_old_dt_.empty()) wold = batch_old_dt_;
        if (W > 1) wid.resize(nb);
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t n_w = wstart[w + 1] - wstart[w];
            // cells of the world's grid: worlds spread (nothing bounds them), and a grid that
            // outgrows its cells doubles its cell size, which crowds the partner walks; the
            // rebuild scans only the cells a grid uses, so room is cheap
            const uint64_t cap = std::max<uint64_t>(8ull * n_w, 8192ull);
            const uint64_t base = (uint64_t)cell_base[w] + ((cap + 63ull) & ~63ull);
            if (base >= 0xFFFFFFF0ull) return fail(BG_ERR_CAPACITY, "cell table of the batch exceeds 2^32 entries");
            cell_base[w + 1] = (uint32_t)base;
            bool first = true;
            for (uint32_t i = wstart[w]; i < wstart[w + 1]; ++i) {
                if (W > 1) wid[i] = w;
                if (first && !(bflags[i] & BF_STATIC)) {
                    bflags[i] |= BF_FIRST_DYN;
                    first = false;
                }
            }
        }
        d.n_worlds = W;
        d.total_cells = cell_base[W];
        cell_rows = d.total_cells;
        HIP_TRY(dmalloc(&d.cell_base, W + 1));
        HIP_TRY(dmalloc(&d.grids, W));
        d.stat_rep = 1;
        while (d.stat_rep < 256u && (size_t)d.stat_rep * 2u * W <= 32768u) d.stat_rep *= 2u;
        HIP_TRY(dmalloc(&d.wbbox, 4u * d.stat_rep * (size_t)W));
        HIP_TRY(dmalloc(&d.wold_dt, W));
        HIP_TRY(dmalloc(&d.wstat, 8u * d.stat_rep * (size_t)W));
        HIP_TRY(dmalloc(&d.wvel, W));
        HIP_TRY(hipMemsetAsync(d.wstat, 0, sizeof(float) * 8u * d.stat_rep * W, s));
        HIP_TRY(hipMemsetAsync(d.wvel, 0, sizeof(float2) * W, s));
        HIP_TRY(h2d(d.cell_base, cell_base, s));
        if (W > 1) {
            // scan tile -> the world holding its first cell (the last world past the table)
            const uint32_t ntiles = (uint32_t)(((uint64_t)d.total_cells + 1u + kScanTile - 1u) / kScanTile);
            std::vector<uint32_t> tile_seg(ntiles);
            uint32_t w = 0;
            for (uint32_t t = 0; t < ntiles; ++t) {
                const uint64_t c0 = (uint64_t)t * kScanTile;
                while (w + 1u < W && cell_base[w + 1] <= c0) ++w;
                tile_seg[t] = w;
            }
            HIP_TRY(dmalloc(&d.tile_seg, ntiles));
            HIP_TRY(h2d(d.tile_seg, tile_seg, s));
            HIP_TRY(hipStreamSynchronize(s));  // the host vector dies here
        }
        {
            std::vector<uint32_t> wbbox(4u * d.stat_rep * (size_t)W);  // empty boxes, filled by each rebuild
            for (size_t i = 0; i < wbbox.size(); ++i) wbbox[i] = (i & 2u) ? 0u : 0xFFFFFFFFu;
            HIP_TRY(h2d(d.wbbox, wbbox, s));
            HIP_TRY(hipStreamSynchronize(s));  // the host vector dies here
        }
        HIP_TRY(h2d(d.wold_dt, wold, s));
        if (W > 1) {
            HIP_TRY(dmalloc(&d.wid, nb));
            HIP_TRY(h2d(d.wid, wid, s));
        }
        // packed records in slot order; slot = dense index until the first rebuild
        for (int copy = 0; copy < 2; ++copy) {
            SphereArrays& q = copy ? d.s_alt : d.s;
            HIP_TRY(dmalloc(&q.pp, nb));
            HIP_TRY(dmalloc(&q.snap, nb));
            HIP_TRY(dmalloc(&q.vel, nb));
            HIP_TRY(dmalloc(&q.rw, nb));
            HIP_TRY(dmalloc(&q.mg, nb));
            HIP_TRY(dmalloc(&q.bflags, nb));
            HIP_TRY(dmalloc(&q.off, nb));
            HIP_TRY(dmalloc(&q.groups, nb));
            HIP_TRY(dmalloc(&q.inertia, nb));
            HIP_TRY(dmalloc(&q.torque, nb));
            HIP_TRY(dmalloc(&q.acc, nb));
            HIP_TRY(dmalloc(&q.vreq, nb));
        }
        HIP_TRY(dmalloc(&d.snap_next, nb));
        // the kernels index the entry-major lists with 32-bit offsets
        if ((uint64_t)(kNbrCap / 2) * nb > 0xFFFFFFFFull)
            return fail(BG_ERR_CAPACITY, "too many bodies for one fused sphere world (partner-list offsets are 32-bit)");
        HIP_TRY(dmalloc(&d.nbr, (size_t)(kNbrCap / 2) * nb));
        HIP_TRY(dmalloc(&d.nbr_len, nb));
        HIP_TRY(hipMemsetAsync(d.nbr_len, kNbrNone, nb, s));
        d.oh_size = next_pow2(std::max<uint32_t>(1024u, nb / 4u));
        for (int i = 0; i < 3; ++i) {
            HIP_TRY(dmalloc(&d.oh_head[i], d.oh_size));
            HIP_TRY(dmalloc(&d.oh_next[i], nb));
            HIP_TRY(dmalloc(&d.oh_bits[i], d.oh_size / 32u));
            d.n_waves = (nb + 255u) / 256u * 4u;  // waves of the 256-thread substep launch
            HIP_TRY(dmalloc(&d.oh_wave[i], d.n_waves));
            HIP_TRY(hipMemsetAsync(d.oh_wave[i], 0, sizeof(uint32_t) * d.n_waves, s));
        }
        HIP_TRY(dmalloc(&d.cell_count, (size_t)d.total_cells + 1));
        HIP_TRY(hipMemsetAsync(d.cell_count, 0, sizeof(uint32_t) * ((size_t)d.total_cells + 1), s));
        HIP_TRY(dmalloc(&d.perm, nb));
        d.gather_pending = false;
        std::vector<float4> pp(nb), snap(nb);
        std::vector<float2> rw(nb), mg(nb);
        snap_flags_ = W == 1 && nb <= kSnapIdMask;
        for (uint32_t i = 0; i < nb; ++i) {
            pp[i] = make_float4(pos[i].x, pos[i].y, pos_old[i].x, pos_old[i].y);
            uint32_t idbits = i | (snap_flags_ ? (bflags[i] & kSnapFlagMask) << kSnapFlagShift : 0u);
            float w;
            memcpy(&w, &idbits, 4);
            snap[i] = make_float4(cabs[i].x, cabs[i].y, crad[i], w);
            rw[i] = make_float2(rot[i], omega[i]);
            mg[i] = make_float2(mass[i], gmod[i]);
        }
        // A single-sphere body's calculated_mass is 2 r per push of its collider (4 r with
        // the double push).  When every body's mass is exactly k r for one power of two k
        // (normal values only, so the product is exact on host and device alike), the
        // contact pass takes a partner's mass from the radius in its snapshot record
        // instead of loading the partner's mass record.
        mass_per_r_ = 0.0f;
        for (const float kk : {4.0f, 2.0f}) {
            bool all = nb > 0;
            for (uint32_t i = 0; i < nb && all; ++i) {
                const float m = kk * crad[i];
                uint32_t mb, wb;
                memcpy(&mb, &m, 4);
                memcpy(&wb, &mass[i], 4);
                all = std::isnormal(crad[i]) && std::isnormal(m) && mb == wb;
            }
            if (all) {
                mass_per_r_ = kk;
                break;
            }
        }
        gmod_one_ = std::all_of(gmod.begin(), gmod.begin() + nb, [](float g) {
            uint32_t gb;
            memcpy(&gb, &g, 4);
            return gb == 0x3F800000u;
        });
        HIP_TRY(h2d(d.s.pp, pp, s));
        HIP_TRY(h2d(d.s.snap, snap, s));
        HIP_TRY(h2d(d.s.vel, vel, s));
        HIP_TRY(h2d(d.s.rw, rw, s));
        HIP_TRY(h2d(d.s.mg, mg, s));
        HIP_TRY(h2d(d.s.bflags, bflags, s));
        HIP_TRY(h2d(d.s.off, coff, s));
        HIP_TRY(h2d(d.s.groups, cgrp, s));
        HIP_TRY(h2d(d.s.inertia, inertia, s));
        HIP_TRY(h2d(d.s.torque, torque, s));
        HIP_TRY(h2d(d.s.acc, acc, s));
        HIP_TRY(h2d(d.s.vreq, vreq, s));
        HIP_TRY(hipStreamSynchronize(s));
    } else {
        BodyArrays& b = d.b;
        HIP_TRY(dmalloc(&b.pos, nb));
        HIP_TRY(dmalloc(&b.pos_old, nb));
        HIP_TRY(dmalloc(&b.vel, nb));
        HIP_TRY(dmalloc(&b.acc, nb));
        HIP_TRY(dmalloc(&b.vreq, nb));
        HIP_TRY(dmalloc(&b.rot, nb));
        HIP_TRY(dmalloc(&b.omega, nb));
        HIP_TRY(dmalloc(&b.torque, nb));
        HIP_TRY(dmalloc(&b.inertia, nb));
        HIP_TRY(dmalloc(&b.mass, nb));
        HIP_TRY(dmalloc(&b.gmod, nb));
        HIP_TRY(dmalloc(&b.bflags, nb));
        HIP_TRY(dmalloc(&b.aid, nb));
        ColliderArrays& c = d.c;
        HIP_TRY(dmalloc(&c.abs, nc));
        HIP_TRY(dmalloc(&c.off, nc));
        HIP_TRY(dmalloc(&c.radius, nc));
        HIP_TRY(dmalloc(&c.parent, nc));
        HIP_TRY(dmalloc(&c.groups, nc));
        HIP_TRY(dmalloc(&c.cflags, nc));
        HIP_TRY(h2d(d.b.pos, pos, s));
        HIP_TRY(h2d(d.b.pos_old, pos_old, s));
        HIP_TRY(h2d(d.b.vel, vel, s));
        HIP_TRY(h2d(d.b.acc, acc, s));
        HIP_TRY(h2d(d.b.vreq, vreq, s));
        HIP_TRY(h2d(d.b.rot, rot, s));
        HIP_TRY(h2d(d.b.omega, omega, s));
        HIP_TRY(h2d(d.b.torque, torque, s));
        HIP_TRY(h2d(d.b.inertia, inertia, s));
        HIP_TRY(h2d(d.b.mass, mass, s));
        HIP_TRY(h2d(d.b.gmod, gmod, s));
        HIP_TRY(h2d(d.b.bflags, bflags, s));
        HIP_TRY(h2d(d.b.aid, aid, s));
        HIP_TRY(h2d(d.c.abs, cabs, s));
        HIP_TRY(h2d(d.c.off, coff, s));
        HIP_TRY(h2d(d.c.radius, crad, s));
        HIP_TRY(h2d(d.c.parent, cpar, s));
        HIP_TRY(h2d(d.c.groups, cgrp, s));
        HIP_TRY(h2d(d.c.cflags, cfl, s));
    }
    HIP_TRY(dmalloc(&d.x0, nc));
    HIP_TRY(dmalloc(&d.bc_start, nb + 1));
    HIP_TRY(dmalloc(&d.bc_list, bc_list.size()));
    HIP_TRY(dmalloc(&d.sp_a, ns));
    HIP_TRY(dmalloc(&d.sp_b, ns));
    HIP_TRY(dmalloc(&d.sp_rest, ns));
    HIP_TRY(dmalloc(&d.sp_k, ns));
    HIP_TRY(dmalloc(&d.sp_c, ns));
    HIP_TRY(dmalloc(&d.bs_start, nb + 1));
    HIP_TRY(dmalloc(&d.bs_list, bs_list.size()));
    HIP_TRY(dmalloc(&d.j_a, nj));
    HIP_TRY(dmalloc(&d.j_b, nj));
    HIP_TRY(dmalloc(&d.j_anchor_a, nj));
    HIP_TRY(dmalloc(&d.j_anchor_b, nj));
    HIP_TRY(dmalloc(&d.j_dist, nj));
    HIP_TRY(dmalloc(&d.j_target, nj));
    HIP_TRY(dmalloc(&d.j_mass, nj));
    if (n_jcomp > 0 && params_.joint_iterations > 0) {
        d.j_iters = params_.joint_iterations;
        HIP_TRY(dmalloc(&d.jrec, (size_t)d.j_iters * nj));
        HIP_TRY(dmalloc(&d.bj_start, nb + 1));
        HIP_TRY(dmalloc(&d.bj_list, bj_list.size()));
        HIP_TRY(dmalloc(&d.bj_comp, nb));
    }
    HIP_TRY(dmalloc(&d.jc_start, n_jcomp + 1));
    HIP_TRY(dmalloc(&d.jc_list, nj));
    HIP_TRY(dmalloc(&d.jb_start, n_jcomp + 1));
    HIP_TRY(dmalloc(&d.jb_list, jb_list.size()));
    HIP_TRY(dmalloc(&d.j_local, nj));
    HIP_TRY(dmalloc(&d.jq_anchor, jq_anchor.size()));
    HIP_TRY(dmalloc(&d.jq_mass, jq_mass.size()));
    HIP_TRY(dmalloc(&d.jq_len, jq_len.size()));
    HIP_TRY(dmalloc(&d.con_pos, nk));
    HIP_TRY(dmalloc(&d.con_r, nk));
    HIP_TRY(dmalloc(&d.keys, nc));
    HIP_TRY(dmalloc(&d.vals, nc));
    HIP_TRY(dmalloc(&d.keys_tmp, nc));
    HIP_TRY(dmalloc(&d.vals_tmp, nc));
    HIP_TRY(dmalloc(&d.cell_start, (size_t)cell_rows + 1));
    if (!sphere) {  // the general path's cell counts (counting sort by cell)
        HIP_TRY(dmalloc(&d.cell_count, (size_t)d.cells_cap + 1));
        HIP_TRY(hipMemsetAsync(d.cell_count, 0, sizeof(uint32_t) * ((size_t)d.cells_cap + 1), s));
        HIP_TRY(dmalloc(&d.crec, nc));
        // ordered contact pass (push-outs): idle state is no rows, no marks, no rewrites
        const uint32_t ncw = std::max<uint32_t>(nc, 1), nbw = std::max<uint32_t>(nb, 1);
        HIP_TRY(dmalloc(&d.ord_rows, ncw));
        HIP_TRY(dmalloc(&d.ord_n, 1));
        HIP_TRY(dmalloc(&d.ord_mark, (ncw + 31) / 32));
        HIP_TRY(dmalloc(&d.pos_pre, nbw));
        HIP_TRY(dmalloc(&d.mv_key, ncw));
        HIP_TRY(dmalloc(&d.mv_abs, ncw));
        HIP_TRY(dmalloc(&d.mv_list, ncw));
        HIP_TRY(dmalloc(&d.ord_heap, ncw));
        d.ord_log_cap = std::min<uint32_t>(std::max<uint32_t>(2u * ncw, 16384u), 1u << 21);
        HIP_TRY(dmalloc(&d.ord_log, d.ord_log_cap));
        HIP_TRY(dmalloc(&d.aff_list, nbw));
        HIP_TRY(dmalloc(&d.aff_mark, (nbw + 31) / 32));
        HIP_TRY(dmalloc(&d.ord_cand, kOrdCand));
        HIP_TRY(dmalloc(&d.ord_terms, kOrdTerms));
        HIP_TRY(hipMemsetAsync(d.ord_n, 0, sizeof(uint32_t), s));
        HIP_TRY(hipMemsetAsync(d.ord_mark, 0, sizeof(uint32_t) * ((ncw + 31) / 32), s));
        HIP_TRY(hipMemsetAsync(d.aff_mark, 0, sizeof(uint32_t) * ((nbw + 31) / 32), s));
        HIP_TRY(hipMemsetAsync(d.mv_key, 0xFF, sizeof(unsigned long long) * ncw, s));
        // two-kernel contacts
        HIP_TRY(dmalloc(&d.clist, (size_t)16 * ncw));
        HIP_TRY(dmalloc(&d.ccnt, ncw));
        HIP_TRY(dmalloc(&d.cfb_list, nbw));
        HIP_TRY(dmalloc(&d.cfb_n, 1));
        HIP_TRY(hipMemsetAsync(d.cfb_n, 0, sizeof(uint32_t), s));
    }
    HIP_TRY(dmalloc(&d.grid, 1));
    HIP_TRY(dmalloc(&d.accum, 1));
    if (event_cap_ > 0) {
        d.event_cap = event_cap_;
        HIP_TRY(dmalloc(&d.events, event_cap_));
        if (sphere) HIP_TRY(dmalloc(&d.vel_next, nb));
    }
    // one contact slot per block of any per-body launch (the smallest blocks have 64 threads)
    d.n_block_slots = (std::max<uint32_t>(std::max(nb, nc), 1) + 63) / 64;  // per-body and per-collider launches
    HIP_TRY(dmalloc(&d.block_contacts, d.n_block_slots));
    HIP_TRY(hipMemsetAsync(d.block_contacts, 0, sizeof(unsigned long long) * d.n_block_slots, d.stream));
    // persistent batched step: a batch of centred single spheres with default groups, no
    // sensors, constraints or events; its worlds of at most kLdsBodies bodies run there,
    // the rest on the grid path
    persist_ = false;
    if (sphere && !batch_start_.empty() && persist_enabled_ && zero_off_ && !groups_ && !sensors_ && nk == 0 &&
        event_cap_ == 0) {
        const uint32_t W = batch_size();
        std::vector<uint32_t> cls[kPersistClasses];
        std::vector<uint8_t> skip(W, 0);
        uint32_t grid_worlds = 0;
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t n_w = batch_start_[w + 1] - batch_start_[w];
            uint32_t c = kPersistClasses;
            for (uint32_t q = 0; q < kPersistClasses; ++q)
                if (n_w <= kPersistClassThreads[q] * kPersistPer) {
                    c = q;
                    break;
                }
            if (c == kPersistClasses || n_w < kPersistMinBodies) {
                ++grid_worlds;
                continue;
            }
            cls[c].push_back(w);
            skip[w] = 1;
        }
        if (grid_worlds < W) {
            // each class's queue: largest world first (longest processing time first)
            std::vector<uint32_t> order;
            int cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.device) != hipSuccess || cus < 1)
                cus = 1;
            // resident workgroups per CU of each class (LDS: about 105 KB a workgroup); beside the
            // grid path the persistent launch takes a quarter of the CUs (measured on the
            // 4,096-world batch against the grid path alone: 5% faster with a quarter, 3% with
            // half, 2% with all of them)
            const uint32_t per_cu[kPersistClasses] = {1u};
            const uint32_t cu_share = grid_worlds > 0 ? 4u : 1u;
            d.p_class_start[0] = 0;
            for (uint32_t c = 0; c < kPersistClasses; ++c) {
                std::stable_sort(cls[c].begin(), cls[c].end(), [&](uint32_t a, uint32_t b) {
                    return batch_start_[a + 1] - batch_start_[a] > batch_start_[b + 1] - batch_start_[b];
                });
                order.insert(order.end(), cls[c].begin(), cls[c].end());
                d.p_class_start[c + 1] = (uint32_t)order.size();
                d.p_groups[c] = std::min<uint32_t>((uint32_t)cls[c].size(),
                                                   std::max<uint32_t>(1u, per_cu[c] * (uint32_t)cus / cu_share));
            }
            d.p_grid_worlds = grid_worlds;
            HIP_TRY(dmalloc(&d.p_order, order.size()));
            HIP_TRY(dmalloc(&d.p_wstart, W + 1));
            HIP_TRY(dmalloc(&d.p_counter, kPersistClasses));
            HIP_TRY(dmalloc(&d.p_skip, W));
            HIP_TRY(h2d(d.p_order, order, s));
            HIP_TRY(h2d(d.p_wstart, batch_start_, s));
            HIP_TRY(h2d(d.p_skip, skip, s));
            HIP_TRY(hipStreamSynchronize(s));  // the host vectors die here
            p_src_ = d.s;  // the persistent worlds' records: where the upload put them
            persist_ = true;
        }
    }
    HIP_TRY(sort_temp_reserve(d.sort, std::max<uint32_t>(nc, cell_rows + 1)));
    HIP_TRY(hipMemsetAsync(d.cell_start, 0, sizeof(uint32_t) * ((size_t)cell_rows + 1), d.stream));
    d.allocated = true;

    HIP_TRY(h2d(d.bc_start, bc_start, s));
    HIP_TRY(h2d(d.bc_list, bc_list, s));
    HIP_TRY(h2d(d.sp_a, sp_a, s));
    HIP_TRY(h2d(d.sp_b, sp_b, s));
    HIP_TRY(h2d(d.sp_rest, sp_rest, s));
    HIP_TRY(h2d(d.sp_k, sp_k, s));
    HIP_TRY(h2d(d.sp_c, sp_c, s));
    HIP_TRY(h2d(d.bs_start, bs_start, s));
    HIP_TRY(h2d(d.bs_list, bs_list, s));
    HIP_TRY(h2d(d.j_a, j_a, s));
    HIP_TRY(h2d(d.j_b, j_b, s));
    HIP_TRY(h2d(d.j_anchor_a, j_an_a, s));
    HIP_TRY(h2d(d.j_anchor_b, j_an_b, s));
    HIP_TRY(h2d(d.j_dist, j_dist, s));
    HIP_TRY(h2d(d.j_target, j_target, s));
    HIP_TRY(h2d(d.j_mass, j_mass, s));
    if (d.jrec) {
        HIP_TRY(h2d(d.bj_start, bj_start, s));
        HIP_TRY(h2d(d.bj_list, bj_list, s));
        HIP_TRY(h2d(d.bj_comp, bj_comp, s));
    }
    HIP_TRY(h2d(d.jc_start, jc_start, s));
    HIP_TRY(h2d(d.jc_list, jc_list, s));
    HIP_TRY(h2d(d.jb_start, jb_start, s));
    HIP_TRY(h2d(d.jb_list, jb_list, s));
    HIP_TRY(h2d(d.j_local, j_local, s));
    HIP_TRY(h2d(d.jq_anchor, jq_anchor, s));
    HIP_TRY(h2d(d.jq_mass, jq_mass, s));
    HIP_TRY(h2d(d.jq_len, jq_len, s));
    HIP_TRY(h2d(d.con_pos, con_pos, s));
    HIP_TRY(h2d(d.con_r, con_r, s));
    StepAccum acc0;
    memset(&acc0, 0, sizeof(acc0));
    bbox_parity_ = 0;
    acc0.bbox[0][0] = f2ord_host(minx);
    acc0.bbox[0][1] = f2ord_host(miny);
    acc0.bbox[0][2] = f2ord_host(maxx);
    acc0.bbox[0][3] = f2ord_host(maxy);
    acc0.bbox[1][0] = acc0.bbox[1][1] = 0xFFFFFFFFu;
    acc0.old_dt = params_.old_dt;
    HIP_TRY(hipMemcpyAsync(d.accum, &acc0, sizeof(acc0), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // host vectors die at return

    dev_dirty_ = false;
    host_stale_ = false;
    stepped_since_upload_ = false;
    fresh_ = true;
    need_rebuild_ = true;
    keys_pending_ = false;
    steps_since_rebuild_ = 0;
    return BG_OK;
}

// Device -> host mirror, in arena order.
int World::download() {
    DeviceWorld& d = d_;
    if (!d.allocated) {
        host_stale_ = false;
        return BG_OK;
    }
    hipStream_t s = d.stream;
    const uint32_t nb = d.nb, nc = d.nc;
    std::vector<float2> pos(nb), pos_old(nb), vel, cabs(nc);
    std::vector<float> rot(nb), omega(nb);
    std::vector<uint32_t> bflags, aid(nb);
    if (path_ == StepPath::Sphere) {
        std::vector<float4> pp, snap;
        std::vector<float2> rw;
        HIP_TRY(d2h(pp, d.s.pp, nb, s));
        HIP_TRY(d2h(snap, d.s.snap, nb, s));
        HIP_TRY(d2h(vel, d.s.vel, nb, s));
        HIP_TRY(d2h(rw, d.s.rw, nb, s));
        HIP_TRY(d2h(bflags, d.s.bflags, nb, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (uint32_t k = 0; k < nb; ++k) {
            pos[k] = make_float2(pp[k].x, pp[k].y);
            pos_old[k] = make_float2(pp[k].z, pp[k].w);
            rot[k] = rw[k].x;
            omega[k] = rw[k].y;
            cabs[k] = make_float2(snap[k].x, snap[k].y);
            uint32_t w;
            memcpy(&w, &snap[k].w, 4);
            aid[k] = w & (snap_flags_ ? kSnapIdMask : ~kSnapOutlier);
        }
    } else {
        HIP_TRY(d2h(pos, d.b.pos, nb, s));
        HIP_TRY(d2h(pos_old, d.b.pos_old, nb, s));
        HIP_TRY(d2h(vel, d.b.vel, nb, s));
        HIP_TRY(d2h(rot, d.b.rot, nb, s));
        HIP_TRY(d2h(omega, d.b.omega, nb, s));
        HIP_TRY(d2h(bflags, d.b.bflags, nb, s));
        HIP_TRY(d2h(cabs, d.c.abs, nc, s));
        for (uint32_t k = 0; k < nb; ++k) aid[k] = k;
    }
    if (path_ == StepPath::Sphere) {
        std::vector<float> wold;
        HIP_TRY(d2h(wold, d.wold_dt, d.n_worlds, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (!batch_start_.empty())
            batch_old_dt_ = wold;
        else
            params_.old_dt = wold[0];
    } else {
        StepAccum acc;
        HIP_TRY(hipMemcpyAsync(&acc, d.accum, sizeof(acc), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        params_.old_dt = acc.old_dt;
    }
    // in the sphere path slot k holds dense body aid[k] (and its collider)
    for (uint32_t k = 0; k < nb; ++k) {
        uint32_t dense = (path_ == StepPath::Sphere) ? aid[k] : k;
        bg_rigid_body& f = bodies_.at_slot(body_slot_of_dense_[dense]).f;
        f.position[0] = pos[k].x;
        f.position[1] = pos[k].y;
        f.position_old[0] = pos_old[k].x;
        f.position_old[1] = pos_old[k].y;
        f.calculated_velocity[0] = vel[k].x;
        f.calculated_velocity[1] = vel[k].y;
        f.rotation = rot[k];
        f.angular_velocity = omega[k];
        f.has_velocity_request = (bflags[k] & BF_HAS_VREQ) ? 1u : 0u;
        if (stepped_since_upload_) {
            f.acceleration[0] = f.acceleration[1] = 0.0f;  // physics.rs:329,354
            if (f.body_type != BG_STATIC) f.torque = 0.0f;  // physics.rs:350,355
        }
    }
    for (uint32_t k = 0; k < nc; ++k) {
        uint32_t dense = (path_ == StepPath::Sphere) ? aid[k] : k;
        bg_collider& f = colliders_.at_slot(col_slot_of_dense_[dense]).f;
        f.absolute_translation[0] = cabs[k].x;
        f.absolute_translation[1] = cabs[k].y;
        // matrix part of body.transform() * offset (physics.rs:363); unused by the step
        if (f.has_parent) {
            const HostBody* hb = bodies_.get(f.parent);
            if (hb) {
                float sn = sinf(hb->f.rotation), cs = cosf(hb->f.rotation);
                for (int col = 0; col < 2; ++col) {
                    float x = f.offset_matrix[2 * col], y = f.offset_matrix[2 * col + 1];
                    f.absolute_matrix[2 * col] = cs * x + (-sn) * y;
                    f.absolute_matrix[2 * col + 1] = sn * x + cs * y;
                }
            }
        }
    }
    host_stale_ = false;
    return BG_OK;
}

int World::rebuild() {
    DeviceWorld& d = d_;
    float cell = (rmax_ > 0.0f) ? cell_scale_ * 2.0f * rmax_ : 1.0f;
    if (!isfinite(cell)) cell = 1.0f;
    const bool sphere = path_ == StepPath::Sphere;
    if (int st = stage_begin(BG_STAGE_REBUILD)) return st;
    if (sphere) {
        HIP_TRY(launch_sphere_rebuild(d, cell, bbox_parity_));
    } else {
        // launched by the substep, beside its forces
        keys_ = KeysArgs{d.c.abs, d.nc, d.accum, bbox_parity_, cell, d.cells_cap, d.grid, d.keys, d.vals_tmp,
                         d.cell_count, d.x0};
        keys_pending_ = true;
    }
    if (int st = stage_end(BG_STAGE_REBUILD)) return st;
    bbox_parity_ ^= 1u;
    drift_x_ = drift_y_ = vdrift_x_ = vdrift_y_ = 0.0;
    h_prev_ = 0.0;
    sub_since_rebuild_ = 0;
    skin_ = skin_scale_ * cell;
    need_rebuild_ = false;
    steps_since_rebuild_ = 0;
    return BG_OK;
}

int World::run_substep(float h, bool fresh, bool last) {
    DeviceWorld& d = d_;
    SubstepArgs a{};
    a.h = h;
    a.ratio_rest = h / h;
    a.gx = params_.gravity[0];
    a.gy = params_.gravity[1];
    a.rmax = rmax_;
    a.mass_per_r = path_ == StepPath::Sphere ? mass_per_r_ : 0.0f;
    a.gmod_one = (path_ == StepPath::Sphere && gmod_one_) ? 1u : 0u;
    a.snap_flags = (path_ == StepPath::Sphere && snap_flags_) ? 1u : 0u;
    a.drift_prev_x = (float)drift_x_;
    a.drift_prev_y = (float)drift_y_;
    if (use_drift_) {
        double hh = (double)h;
        double scale = h_prev_ > 0.0 ? hh / h_prev_ : 0.0;
        vdrift_x_ = vdrift_x_ * scale + (double)params_.gravity[0] * hh * hh;
        vdrift_y_ = vdrift_y_ * scale + (double)params_.gravity[1] * hh * hh;
        drift_x_ += vdrift_x_;
        drift_y_ += vdrift_y_;
        h_prev_ = hh;
    }
    a.drift_x = (float)drift_x_;
    a.drift_y = (float)drift_y_;
    a.first_dyn = first_dyn_;
    a.fresh = fresh ? 1u : 0u;
    a.collisions = params_.collisions_enabled ? 1u : 0u;
    a.parity = bbox_parity_;
    a.has_springs = d.ns > 0 ? 1u : 0u;
    a.skin = skin_;
    a.sub = ++sub_since_rebuild_;
    a.sub_global = substep_seq_++;
    // calculated_velocity is read only after a step (getters, the next rebuild's
    // statistics) or by collision events, which need every substep's
    a.write_vel = (last || d.events) ? 1u : 0u;
    a.t_prev = (float)((double)(a.sub - 1) * (double)h);
    a.t_cur = (float)((double)a.sub * (double)h);
    if (path_ == StepPath::Sphere) {
        // the fused substep is one kernel: its launch records the stage events at the
        // kernel's own start and end, so the stage time is the kernel's duration
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (profiling_) {
            HIP_TRY(pooled_event(&e0));
            HIP_TRY(pooled_event(&e1));
        }
        d.prof_start = e0;
        d.prof_stop = e1;
        const hipError_t e = launch_sphere_substep(d, a, groups_, sensors_, zero_off_, d.nk > 0);
        d.prof_start = d.prof_stop = nullptr;
        HIP_TRY(e);
        if (profiling_) pending_[BG_STAGE_SUBSTEP].push_back({e0, e1});
        return BG_OK;
    }
    if (int st = stage_begin(BG_STAGE_SUBSTEP)) return st;
    const KeysArgs* keys = keys_pending_ ? &keys_ : nullptr;
    keys_pending_ = false;
    HIP_TRY(launch_general_substep(d, a, params_.joint_iterations, groups_, sensors_, keys));
    return stage_end(BG_STAGE_SUBSTEP);
}

// Panics and routing from the step record's error flags (physics.rs:293,
// 447-448, 471-474).
int World::process_flags(uint32_t flags) {
    const uint32_t panic_bits = flags & (ERR_NAN_NORMAL | ERR_JOINT_MASS | ERR_JOINT_ROT | ERR_ORDERED_CAP);
    route_after_step(flags);
    if (panic_bits & ERR_NAN_NORMAL) return fail(BG_ERR_PANIC, "assertion failed: !n.is_nan() (physics.rs:293)");
    if (panic_bits & ERR_JOINT_MASS)
        return fail(BG_ERR_PANIC, "assertion failed: body calculated_mass > 0.0 (physics.rs:447-448)");
    if (panic_bits & ERR_JOINT_ROT)
        return fail(BG_ERR_PANIC, "assertion failed: body rotation is finite (physics.rs:471-474)");
    if (panic_bits & ERR_ORDERED_CAP)
        return fail(BG_ERR_CAPACITY, "a push-out cluster (physics.rs:272-286) outgrew the ordered contact pass's buffers");
    return BG_OK;
}

// Deferred checks: the step record's flags come back through pinned memory
// behind the step's kernels, and are read by a later call once they arrived.
int World::set_deferred_checks(bool on) {
    if (!on) {
        int st = finish_deferred_check(true);
        deferred_checks_ = false;
        return st;
    }
    if (!host_acc_) HIP_TRY(hipHostMalloc((void**)&host_acc_, sizeof(StepAccum), hipHostMallocDefault));
    if (!check_event_) HIP_TRY(hipEventCreateWithFlags(&check_event_, hipEventDisableTiming));
    deferred_checks_ = true;
    return BG_OK;
}

int World::finish_deferred_check(bool wait) {
    if (!check_pending_) return BG_OK;
    if (wait) {
        HIP_TRY(hipEventSynchronize(check_event_));
    } else {
        const hipError_t q = hipEventQuery(check_event_);
        if (q == hipErrorNotReady) return BG_OK;
        HIP_TRY(q);
    }
    check_pending_ = false;
    return process_flags(host_acc_->error_flags);
}

int World::check_errors(uint32_t* flags_out) {
    DeviceWorld& d = d_;
    StepAccum acc;
    HIP_TRY(hipMemcpyAsync(&acc, d.accum, sizeof(acc), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    if (flags_out) *flags_out = acc.error_flags;
    if (acc.error_flags) {
        uint32_t zero = 0;
        HIP_TRY(hipMemcpyAsync(&d.accum->error_flags, &zero, 4, hipMemcpyHostToDevice, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
    }
    return BG_OK;
}

int World::prepare_step() {
    if (params_.collisions_enabled && params_.use_spatial_hash)
        return fail(BG_ERR_UNSUPPORTED, "spatial collisions not supported right now (physics.rs:411-412)");
    int st = ensure_device();
    if (st) return st;
    if (dev_dirty_) {
        st = upload();
        if (st) return st;
    }
    if (structure_error_) return fail(BG_ERR_PANIC, structure_msg_);
    return BG_OK;
}

// Enqueues n steps (no host synchronisation; safe under stream capture).
// The persistent batched step: every substep of the n steps for the worlds
// it runs, in one launch per size class (batch_kernels.hip).  `fresh`: the
// first of these substeps consumes the one-off user inputs.
int World::launch_persistent(double delta, uint32_t n, bool fresh, hipStream_t stream) {
    DeviceWorld& d = d_;
    const float df = (float)delta;                 // physics.rs:80
    const float h = df / (float)params_.substeps;  // physics.rs:399
    PersistArgs k{};
    k.src = p_src_;
    k.dst = d.s;
    k.wstart = d.p_wstart;
    k.wold_dt = d.wold_dt;
    k.order = d.p_order;
    for (uint32_t c = 0; c <= kPersistClasses; ++c) k.class_start[c] = d.p_class_start[c];
    for (uint32_t c = 0; c < kPersistClasses; ++c) k.groups[c] = d.p_groups[c];
    k.counter = d.p_counter;
    k.accum = d.accum;
    k.block_contacts = d.block_contacts;
    k.h = h;
    k.ratio_rest = h / h;
    k.gx = params_.gravity[0];
    k.gy = params_.gravity[1];
    k.nsub = n * params_.substeps;
    k.fresh = fresh ? 1u : 0u;
    k.collisions = params_.collisions_enabled ? 1u : 0u;
    HIP_TRY(launch_batch_persistent(k, stream));
    p_src_ = d.s;  // the persistent worlds' records are current again
    return BG_OK;
}

int World::launch_steps(double delta, uint32_t n) {
    if (persist_ && path_ == StepPath::Sphere && n > 0) {
        DeviceWorld& d = d_;
        const bool fresh = fresh_;
        if (d.p_grid_worlds == 0) {  // every world on the persistent kernel
            if (int st = stage_begin(BG_STAGE_SUBSTEP)) return st;
            if (int st = launch_persistent(delta, n, fresh, d.stream)) return st;
            if (int st = stage_end(BG_STAGE_SUBSTEP)) return st;
            substep_seq_ += n * params_.substeps;
        } else {
            // The persistent launch runs beside the grid path's substeps: it touches only its
            // own worlds' slots and the step's atomic counters, and writes its records into
            // the buffers the grid path leaves current, which the host knows once it has
            // queued those substeps.
            if (!d.p_stream) {
                HIP_TRY(hipStreamCreateWithFlags(&d.p_stream, hipStreamNonBlocking));
                HIP_TRY(hipEventCreateWithFlags(&d.p_fork, hipEventDisableTiming));
                HIP_TRY(hipEventCreateWithFlags(&d.p_join, hipEventDisableTiming));
            }
            HIP_TRY(hipEventRecord(d.p_fork, d.stream));
            HIP_TRY(hipStreamWaitEvent(d.p_stream, d.p_fork, 0));
            if (int st = launch_grid_steps(delta, n)) return st;
            if (int st = launch_persistent(delta, n, fresh, d.p_stream)) return st;
            HIP_TRY(hipEventRecord(d.p_join, d.p_stream));
            HIP_TRY(hipStreamWaitEvent(d.stream, d.p_join, 0));
        }
        fresh_ = false;
        stepped_since_upload_ = true;
        return BG_OK;
    }
    return launch_grid_steps(delta, n);
}

int World::launch_grid_steps(double delta, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) {
        const float df = (float)delta;                             // physics.rs:80
        const float h = df / (float)params_.substeps;              // physics.rs:399
        for (uint32_t sub = 0; sub < params_.substeps; ++sub) {
            if (need_rebuild_ && params_.collisions_enabled) {
                int st = rebuild();
                if (st) return st;
            }
            int st = run_substep(h, fresh_, sub + 1 == params_.substeps);
            if (st) return st;
            fresh_ = false;
            stepped_since_upload_ = true;
            // the general path widens every query by the largest displacement since its
            // rebuild (one fast collider would slow every query): it re-sorts every substep
            if (path_ == StepPath::General) need_rebuild_ = true;
        }
        const uint32_t interval = rebuild_interval_ ? rebuild_interval_ : (d_.nb < 65536u ? 1u : 3u);
        if (++steps_since_rebuild_ >= interval) need_rebuild_ = true;
    }
    return BG_OK;
}

// time += delta per step (physics.rs:81), for every world of a batch
void World::advance_time(double delta, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) {
        params_.time += delta;
        for (double& t : batch_time_) t += delta;
    }
}

int World::step(double delta, uint32_t n) {
    if (n == 0) return BG_OK;
    int st = finish_deferred_check(false);  // a panic of an earlier step surfaces here
    if (st) return st;
    st = prepare_step();
    if (st) return st;
    st = launch_steps(delta, n);
    if (st) return st;
    advance_time(delta, n);
    host_stale_ = true;
    if (deferred_checks_) {
        // wait for the previous record before its pinned buffer is reused
        st = finish_deferred_check(true);
        if (st) return st;
        DeviceWorld& d = d_;
        HIP_TRY(hipMemcpyAsync(host_acc_, d.accum, sizeof(StepAccum), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(launch_fill_u32(&d.accum->error_flags, 0u, 1, d.stream));
        HIP_TRY(hipEventRecord(check_event_, d.stream));
        check_pending_ = true;
        return BG_OK;
    }
    uint32_t flags = 0;
    st = check_errors(&flags);
    if (st) return st;
    return process_flags(flags);
}

// Path of the next steps of a single-sphere world (see push_out_hold_).  A
// switch re-uploads from the host mirror, which carries the whole state.
void World::route_after_step(uint32_t flags) {
    if (!batch_start_.empty()) return;
    bool to_general = false, to_sphere = false;
    if (flags & INFO_PUSH_OUT) {
        push_out_hold_ = kPushOutHold;
        if (path_ == StepPath::Sphere) {
            ++pairwise_push_out_steps_;
            to_general = true;
        }
    } else if (push_out_hold_ > 0 && --push_out_hold_ == 0 && path_ == StepPath::General) {
        to_sphere = sphere_eligible();
    }
    if ((to_general || to_sphere) && sync_host() == BG_OK) mark_dirty();
}

// Physics::fixed_step, physics.rs:84-99.
int World::fixed_step(double frame_time) {
    params_.accumulator += frame_time;
    const double delta = 1.0 / 60.0;
    int max_steps = 3;
    while (params_.accumulator >= delta && max_steps > 0) {
        int st = step(delta, 1);  // integrate(substeps, delta as f32); time += delta
        if (st) return st;
        params_.accumulator -= delta;
        max_steps -= 1;
    }
    return BG_OK;
}

int World::synchronize() {
    if (d_.stream) HIP_TRY(hipStreamSynchronize(d_.stream));
    return finish_deferred_check(true);
}

}  // namespace bg
