// diff.hpp -- differentiable rollouts of batches of single-sphere worlds with
// springs (damped or not) and fixed joints: forward with checkpoints, adjoint
// by segment recomputation.
//
// Each world runs in one 1024-thread workgroup for the whole rollout (two
// bodies per thread, all shared state in LDS), so a world holds at most
// kDiffMaxBodies bodies.  The forward is the reference substep
// (physics.rs:401-421: gravity, springs, contacts, joint positions, Verlet,
// collider transform) with the same f32 operation order as the general
// path; the adjoint differentiates it with the contact pair set of every
// substep held fixed (recomputed from the recorded snapshot, not re-decided).
// Joint rotations are not part of the rollout state: with centred colliders
// and unrotated anchors no position reads a rotation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace bg {

constexpr uint32_t kDiffThreads = 1024;
constexpr uint32_t kDiffPer = 2;                                // bodies per thread
constexpr uint32_t kDiffMaxBodies = kDiffThreads * kDiffPer;    // per world
constexpr uint32_t kDiffCells = 4096;                           // LDS grid cells per world
constexpr uint32_t kDiffTapeNbr = 8;                            // u16 per tape record: count + 7 partners

struct DiffWorld {
    uint32_t b0, nb;    // bodies [b0, b0 + nb)
    uint32_t s0, ns;    // springs [s0, s0 + ns)
    float old_dt0;      // Physics::old_dt when the rollout starts (Q1, physics.rs:338-340)
    int32_t first_dyn;  // local index of the first non-static body, -1 if none
    float rmax;         // largest radius of the world
    float tx, ty;       // target point of the centroid loss
    uint32_t jc0, njc;  // joint components [jc0, jc0 + njc) (their joints are contiguous)
};

struct DiffArgs {
    const DiffWorld* worlds;
    uint32_t n_worlds, nb_total;
    // bodies, global index
    const float* mass;
    const float* radius;
    const float* gmod;
    const uint32_t* flags;  // BF_STATIC, BF_HAS_VREQ
    const float2* vreq;     // initial velocity requests (a parameter)
    const float2* acc0;     // user accelerations, consumed by the first substep
    // springs, global index; endpoints are global body indices
    const int32_t* sp_a;
    const int32_t* sp_b;
    const float* sp_rest;
    const float* sp_k;  // stiffness (a parameter)
    const float* sp_c;
    const uint32_t* inc_start;  // per body: its springs in ascending index
    const uint32_t* inc_list;
    // fixed joints, grouped by connected component (arena order inside one)
    const uint32_t* jc_start;  // per component: first joint, n_components + 1 entries
    const uint32_t* j_ab;      // local body a | local body b << 16
    const float4* j_anchor;    // anchor a (x, y), anchor b (z, w)
    const float* j_dist;
    const float4* j_mass;  // 1/m_a + 1/m_b, ratio, 1 - ratio, unused
    uint32_t n_joints, j_iters;
    // checkpoints [n_ck][nb_total] of position, position_old, calculated_velocity
    float2* ck_p;
    float2* ck_q;
    float2* ck_v;
    float2* final_p;
    float* loss;  // per world: |centroid(final position) - target|^2
    // tape of one segment, [substep in segment][nb_total]
    float2* tape_s;
    uint16_t* tape_nbr;  // [..][nb_total][kDiffTapeNbr]
    float2* tape_v;      // calculated_velocity at the substep start (damped springs), else null
    float2* tape_pj;     // positions entering the joint stage (joints), else null
    float2* j_axis;      // [j_iters][n_joints]: anchor axes of the reverse pass's joint sweep
    // adjoint
    float2* lam_p;  // d loss / d position at a segment boundary
    float2* lam_q;  // d loss / d position_old
    float2* lam_v;  // d loss / d calculated_velocity
    const float2* dl_dfinal;  // d loss / d final position, or null for the centroid loss
    float2* grad_vreq;
    float* grad_k;
    float gx, gy, h;
    uint32_t substeps, collisions;
};

hipError_t launch_diff_forward(const DiffArgs& a, uint32_t n_steps, uint32_t ck_every, hipStream_t s);
hipError_t launch_diff_backward_segment(const DiffArgs& a, uint32_t seg, uint32_t seg_first_step,
                                        uint32_t seg_steps, bool last_segment, hipStream_t s);

class World;

// A batch of worlds prepared for forward/adjoint rollouts (diff.cpp).  The
// parameters are the initial velocity requests and the spring stiffnesses;
// the loss of world w is |centroid of its final positions - target_w|^2
// unless the caller supplies d loss / d final positions.
class Rollout {
public:
    ~Rollout();
    int create(World* const* worlds, uint32_t n, const float* targets, uint32_t checkpoint_every);
    int set_params(const float* vreq, const float* k);
    int forward(double delta, uint32_t n_steps, float* loss_per_world);
    int backward(const float* dl_dfinal, float* grad_vreq, float* grad_k);
    int final_positions(float* out);
    int nonfinite(uint8_t* per_world) const;
    uint32_t n_bodies() const { return nb_; }
    uint32_t n_springs() const { return ns_; }
    uint32_t n_worlds() const { return n_worlds_; }
    int fail(int code, const std::string& msg) {
        last_error = msg;
        return code;
    }
    std::string last_error;

private:
    void release();
    int ensure_buffers(uint32_t n_steps);
    DiffArgs args() const;

    hipStream_t stream_ = nullptr;
    uint32_t n_worlds_ = 0, nb_ = 0, ns_ = 0, substeps_ = 8, ck_every_ = 20, n_steps_ = 0;
    uint32_t n_ck_alloc_ = 0, tape_sub_alloc_ = 0, nj_ = 0, j_iters_ = 0;
    float gx_ = 0.0f, gy_ = 0.0f, h_ = 0.0f;
    bool collisions_ = true, damped_ = false, have_forward_ = false;
    std::vector<uint8_t> nonfinite_;  // per world, after the last backward: a gradient entry is inf or NaN
    std::vector<uint32_t> body0_;     // first body of each world (n_worlds + 1 entries)
    std::vector<uint32_t> spring_world_;  // world of each spring
    std::vector<float2> p0_, q0_, v0_, vreq_, acc0_;
    std::vector<float> mass_, radius_, gmod_, rest_, k_, c_;
    std::vector<uint32_t> flags_;
    std::vector<int32_t> sa_, sb_;
    DiffWorld* d_worlds_ = nullptr;
    float *d_mass_ = nullptr, *d_radius_ = nullptr, *d_gmod_ = nullptr, *d_rest_ = nullptr, *d_k_ = nullptr,
          *d_c_ = nullptr, *d_loss_ = nullptr, *d_gk_ = nullptr;
    uint32_t *d_flags_ = nullptr, *d_inc_start_ = nullptr, *d_inc_list_ = nullptr;
    int32_t *d_sa_ = nullptr, *d_sb_ = nullptr;
    float2 *d_vreq_ = nullptr, *d_acc0_ = nullptr, *d_ck_p_ = nullptr, *d_ck_q_ = nullptr, *d_ck_v_ = nullptr,
           *d_final_ = nullptr, *d_tape_s_ = nullptr, *d_lam_p_ = nullptr, *d_lam_q_ = nullptr, *d_dl_ = nullptr,
           *d_gv_ = nullptr;
    uint16_t* d_tape_nbr_ = nullptr;
    float2 *d_tape_v_ = nullptr, *d_tape_pj_ = nullptr, *d_j_axis_ = nullptr, *d_lam_v_ = nullptr;
    uint32_t *d_jc_start_ = nullptr, *d_j_ab_ = nullptr;
    float4 *d_j_anchor_ = nullptr, *d_j_mass_ = nullptr;
    float* d_j_dist_ = nullptr;
};

}  // namespace bg
