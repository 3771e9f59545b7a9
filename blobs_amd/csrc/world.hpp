// world.hpp -- host side of one `Physics` world: the arenas the reference
// keeps (physics.rs:11-16), mass setup, and the orchestration of the device
// step.  The host mirror is authoritative between mutations; after a step
// the device is, and getters download lazily.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/blobs_gpu.h"
#include "arena.hpp"
#include "device_state.hpp"

namespace bg {

struct HostBody {
    bg_rigid_body f{};
    std::vector<bg_handle> colliders;  // rbd.colliders, each handle twice (Q2)
    std::vector<bg_handle> joints;     // rbd.connected_joints
};

struct HostCollider {
    bg_collider f{};
};

struct HostConstraint {
    float x, y, r;
};

enum class StepPath : uint32_t { None = 0, General = 1, Sphere = 2 };

// One world flattened for a differentiable rollout (dense iteration order).
struct DiffHostWorld {
    std::vector<float2> pos, pos_old, vel, vreq, acc;
    std::vector<float> mass, radius, gmod;
    std::vector<uint32_t> flags;
    std::vector<int32_t> sa, sb;  // spring endpoints, dense body index
    std::vector<float> rest, k, c;
    std::vector<int32_t> ja, jb;  // fixed joints in arena order, dense body index
    std::vector<float2> jan_a, jan_b;
    std::vector<float> jdist;
    float old_dt = 1.0f, gx = 0.0f, gy = 0.0f;
    uint32_t substeps = 8, joint_iterations = 4;
    bool collisions = true;
};

struct Status {
    int code = BG_OK;
    std::string msg;
};

class World {
public:
    World(float gx, float gy, bool use_spatial_hash);
    ~World();

    // construction (physics.rs:121-239)
    int insert_rbd(const bg_rigid_body& b, bg_handle* out);
    int insert_collider_with_parent(const bg_collider& c, bg_handle parent, bg_handle* out);
    int insert_rbds(uint32_t n, const bg_body_batch& b, bg_handle* out);
    int insert_colliders(uint32_t n, const bg_collider_batch& c, const bg_handle* parents, bg_handle* out);
    int create_fixed_joint(bg_handle a, bg_handle b, const float* aa, const float* ab, bool with_distance,
                           float distance, bg_handle* out);
    int insert_spring(const bg_spring& s, bg_handle* out);
    int push_constraint(float x, float y, float r);
    int remove_rbd(bg_handle h);
    int remove_col(bg_handle h);
    int reset();

    // stepping
    int step(double delta, uint32_t n);
    int fixed_step(double frame_time);
    int synchronize();
    // n_chunks x chunk_steps steps, each chunk replayed from a captured hipGraph
    int step_chunks(double delta, uint32_t chunk_steps, uint32_t n_chunks);
    // batches: step with the persistent kernel when the batch allows it (default on)
    int set_persistent(bool enable);
    bool persistent_active() const { return persist_; }
    // time candidate broadphase settings over real steps; keep the fastest
    int autotune(double delta, uint32_t steps_per_trial, float* cell_scale, float* skin, uint32_t* interval);

    // batches of single-sphere worlds (batch.cpp)
    int init_batch(World* const* src, uint32_t n);
    int batch_writeback(World* const* dst, uint32_t n);
    uint32_t batch_size() const { return batch_start_.empty() ? 0u : (uint32_t)batch_start_.size() - 1; }
    // differentiable rollouts (diff.cpp)
    int export_diff(DiffHostWorld& out);

    // access
    int get_params(bg_params* p);
    int set_params(const bg_params* p);
    int get_rbd(bg_handle h, bg_rigid_body* out);
    int set_rbd(bg_handle h, const bg_rigid_body* in);
    int get_col(bg_handle h, bg_collider* out);
    int get_joint(bg_handle h, bg_fixed_joint* out);
    int get_spring(bg_handle h, bg_spring* out);
    int get_bodies(float* pos, float* pos_old, float* vel, float* rot, float* omega, bg_handle* handles);
    int get_collider_positions(float* abs, uint32_t* count);
    int contact_count(uint64_t* out);
    int set_broadphase(float cell_scale, float skin, uint32_t rebuild_interval);
    int set_event_capacity(uint32_t cap);
    int drain_events(bg_collision_event* out, uint32_t cap, uint32_t* n_out, uint64_t* n_dropped);
    int debug_data(bg_affine2* bodies, bg_affine2* cols, float* radius, float* joints, float* springs,
                   uint32_t* counts);
    int set_profiling(bool enable);
    int set_deferred_checks(bool on);
    int debug_accum(uint32_t* out4);
    int debug_list_lengths(uint32_t* out);
    int stage_stats(uint32_t stage, uint64_t* launches, double* total_ms);
    uint32_t rbd_count() const { return bodies_.len(); }
    StepPath path() const { return path_; }

    int fail(int code, const std::string& msg) {
        last_error = msg;
        return code;
    }
    std::string last_error;

private:
    void update_mass_and_inertia(HostBody& b);
    int ensure_device();
    int upload();
    int download();
    int sync_host() { return host_stale_ ? download() : BG_OK; }
    void mark_dirty() { dev_dirty_ = true; }
    int rebuild();
    int run_substep(float h, bool fresh, bool last);
    int prepare_step();
    int launch_steps(double delta, uint32_t n);
    int launch_grid_steps(double delta, uint32_t n);
    int launch_persistent(double delta, uint32_t n, bool fresh, hipStream_t stream);
    void advance_time(double delta, uint32_t n);
    bool sphere_eligible() const;
    void dense_slots(std::vector<uint32_t>& bodies, std::vector<uint32_t>& cols) const;
    int check_errors(uint32_t* flags_out = nullptr);
    int process_flags(uint32_t flags);
    int check_and_process() {
        uint32_t f = 0;
        int st = check_errors(&f);
        return st ? st : process_flags(f);
    }
    int finish_deferred_check(bool wait);
    void route_after_step(uint32_t flags);
    int device_contacts(uint64_t* out);
    void free_device();
    void free_arrays();
    hipError_t pooled_event(hipEvent_t* out);
    int stage_begin(uint32_t stage);
    int stage_end(uint32_t stage);
    int collect_stage_times();
    int collect_events();

    Arena<HostBody> bodies_;
    Arena<HostCollider> colliders_;
    Arena<bg_fixed_joint> joints_;
    Arena<bg_spring> springs_;
    std::vector<HostConstraint> constraints_;
    bg_params params_{};

    // device mirror
    DeviceWorld d_{};
    StepPath path_ = StepPath::None;
    bool dev_dirty_ = true;
    bool host_stale_ = false;
    bool stepped_since_upload_ = false;
    bool fresh_ = true;
    bool need_rebuild_ = true;
    // general path: the rebuild's key pass waits for the substep's first launch
    bool keys_pending_ = false;
    KeysArgs keys_{};
    uint32_t steps_since_rebuild_ = 0;
    uint32_t bbox_parity_ = 0;
    // broadphase settings until bg_set_broadphase: measured defaults (the small plain scene
    // steps 2.6x faster than with a skin of half a cell; large worlds gain from rebuilding
    // every third step)
    float cell_scale_ = 1.25f;
    float skin_scale_ = 0.25f;  // outlier skin as a fraction of the cell size
    float skin_ = 0.0f;
    uint32_t sub_since_rebuild_ = 0;
    uint32_t rebuild_interval_ = 0;  // 0: every step below 65,536 bodies, every third step from there
    // world properties fixed at upload
    float rmax_ = 0.0f;
    float mass_per_r_ = 0.0f;  // sphere path: every body's mass is exactly this times its radius (0: not so)
    bool gmod_one_ = false;    // sphere path: every body's gravity_mod is exactly 1
    bool snap_flags_ = false;  // sphere path: snapshot words carry the body flags (one world, < 2^24 bodies)
    int32_t first_dyn_ = -1;
    bool groups_ = false, sensors_ = false, zero_off_ = false, use_drift_ = false;
    bool structure_error_ = false;
    std::string structure_msg_;
    // push-outs (physics.rs:272-286) are replayed in order on the general path only: a
    // single-sphere world steps there while its uploaded snapshot has coincident pairs
    // and for push_out_hold_ more steps after a step that had a push-out
    uint32_t push_out_hold_ = 0;
    // deferred checks: the step record's flags arrive in pinned host memory
    bool deferred_checks_ = false, check_pending_ = false;
    StepAccum* host_acc_ = nullptr;
    hipEvent_t check_event_ = nullptr;
    uint64_t pairwise_push_out_steps_ = 0;  // fused-path steps that resolved push-outs pair by pair
    // dense <-> slot maps of the last upload
    std::vector<uint32_t> body_slot_of_dense_, col_slot_of_dense_;
    std::vector<int32_t> body_dense_of_slot_, col_dense_of_slot_;
    // broadphase drift of the current rebuild period
    double drift_x_ = 0, drift_y_ = 0, vdrift_x_ = 0, vdrift_y_ = 0, h_prev_ = 0;
    uint64_t contact_base_ = 0;
    // collision events: device capacity, host queue (reference order), losses
    uint32_t event_cap_ = 0;
    uint32_t substep_seq_ = 0;
    std::vector<bg_collision_event> event_queue_;
    size_t event_head_ = 0;
    uint64_t events_dropped_ = 0;
    // batch: dense body range, old_dt and time of each source world
    std::vector<uint32_t> batch_start_;
    // batch stepped by the persistent kernel (one workgroup per world): allowed, and in
    // effect since the last upload (single centred spheres, default groups, no sensors,
    // constraints or events, worlds of at most kPersistMaxBodies bodies).  Off by default:
    // over long runs of spreading worlds the grid path alone is faster
    bool persist_enabled_ = false, persist_ = false;
    SphereArrays p_src_{};  // where the last persistent launch (or the upload) left its worlds' records
    std::vector<float> batch_old_dt_;
    std::vector<double> batch_time_;
    // profiling: event pairs recorded since the last collection, per stage
    bool profiling_ = false;
    std::vector<hipEvent_t> event_pool_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_[BG_STAGE_COUNT];
    hipEvent_t open_[BG_STAGE_COUNT] = {nullptr, nullptr};
    uint64_t stage_launches_[BG_STAGE_COUNT] = {0, 0};
    double stage_ms_[BG_STAGE_COUNT] = {0.0, 0.0};
};

}  // namespace bg
