// device_state.hpp -- layout of one world's state in HBM and the launchers
// of the per-substep kernels.
//
// Structure of arrays, f32, indexed by "dense" index (arena slot order with
// holes removed, i.e. the reference's iteration order).  The fused sphere
// path additionally keeps bodies physically ordered by broadphase cell
// (`aid[slot]` maps back to the dense index); the general path keeps dense
// order and reaches colliders through the sorted index array.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "sort.hpp"

namespace bg {

struct BodyArrays {
    float2 *pos, *pos_old, *vel, *acc, *vreq;
    float *rot, *omega, *torque, *inertia, *mass, *gmod;
    uint32_t *bflags;
    uint32_t *aid;
};

struct ColliderArrays {
    float2 *abs;     // translation of absolute_transform (snapshot)
    float2 *off;     // translation of offset
    float *radius;
    int32_t *parent; // dense body index, -1 = None
    uint2 *groups;   // (memberships, filter)
    uint32_t *cflags;  // bit0 = sensor
};

struct DeviceWorld {
    uint32_t nb = 0, nc = 0, ns = 0, nj = 0, nk = 0;
    uint32_t cap_b = 0, cap_c = 0, cap_s = 0, cap_j = 0, cap_k = 0, cap_bc = 0, cap_bs = 0, cap_jc = 0;
    uint32_t cells_cap = 0;
    int key_bits = 0;

    BodyArrays b{}, b_alt{};
    ColliderArrays c{}, c_alt{};
    float2 *abs_next = nullptr;  // sphere path ping-pong snapshot
    float2 *x0 = nullptr;        // collider snapshot at the last rebuild
    // body -> colliders (general path)
    uint32_t *bc_start = nullptr, *bc_list = nullptr;
    // springs
    int32_t *sp_a = nullptr, *sp_b = nullptr;
    float *sp_rest = nullptr, *sp_k = nullptr, *sp_c = nullptr;
    uint32_t *bs_start = nullptr, *bs_list = nullptr;
    // joints, grouped into independent components (disjoint body sets)
    int32_t *j_a = nullptr, *j_b = nullptr;
    float2 *j_anchor_a = nullptr, *j_anchor_b = nullptr;
    float *j_dist = nullptr, *j_target = nullptr;
    uint32_t *jc_start = nullptr, *jc_list = nullptr;
    uint32_t n_jcomp = 0;
    // constraints
    float2 *con_pos = nullptr;
    float *con_r = nullptr;
    // broadphase
    uint32_t *keys = nullptr, *vals = nullptr, *keys_tmp = nullptr, *vals_tmp = nullptr;
    uint32_t *cell_start = nullptr;
    uint32_t *slot_of = nullptr;  // sphere path: dense id -> slot
    GridParams *grid = nullptr;
    StepAccum *accum = nullptr;
    unsigned long long *block_contacts = nullptr;  // contact count per block of the per-body kernels
    uint32_t n_block_slots = 0;
    SortTemp sort{};

    hipStream_t stream = nullptr;
    int device = -1;
    bool allocated = false;
};

// Scalars of one substep launch.
struct SubstepArgs {
    float h;            // step_delta = delta / substeps (physics.rs:399)
    float ratio_rest;   // h / h: displacement ratio of every non-static body but the first (physics.rs:338)
    float gx, gy;       // Physics::gravity
    float rmax;         // largest collider radius
    float drift_prev_x, drift_prev_y;  // common drift at the previous snapshot (broadphase query centre)
    float drift_x, drift_y;            // common drift after this substep
    int32_t first_dyn;  // dense index of the first non-static body (Q1), -1 if none
    uint32_t fresh;     // first substep after an upload: user acc/torque are live
    uint32_t collisions;
    uint32_t parity;    // which bbox accumulator this substep writes
    uint32_t has_springs;
    uint32_t inertia_all_pos;
};

// Launchers (step_kernels.hip).
hipError_t launch_rebuild_keys(DeviceWorld& d, float cell, uint32_t parity, bool sphere_path);
hipError_t launch_rebuild_sort_and_cells(DeviceWorld& d);
hipError_t launch_sphere_permute(DeviceWorld& d, bool carry_user_inputs);
hipError_t launch_general_substep(DeviceWorld& d, const SubstepArgs& a, uint32_t joint_iterations, bool groups,
                                  bool sensors);
hipError_t launch_sphere_substep(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors, bool zero_off,
                                 float2* abs_in, float2* abs_out);
hipError_t launch_sphere_unpermute(DeviceWorld& d, float2* abs_cur);

}  // namespace bg
