// device_state.hpp -- layout of one world's state in HBM and the launchers
// of the per-substep kernels.
//
// Structure of arrays, f32, indexed by "dense" index (arena slot order with
// holes removed, i.e. the reference's iteration order).  The general path
// keeps dense order and reaches colliders through the sorted index array.
// The fused sphere path keeps bodies physically ordered by broadphase cell
// (`aid` in the snapshot maps a slot back to its dense index) in packed
// 16-byte records.
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

// Fused sphere path: body and its only collider share a slot.
//   snap.w bits: dense id (31 bits) | outlier flag for the next substep (bit 31)
struct SphereArrays {
    float4 *pp;    // {position.x, position.y, position_old.x, position_old.y}
    float4 *snap;  // {absolute translation x, y, radius, bits(id | outlier << 31)}
    float2 *vel;   // calculated_velocity
    float2 *rw;    // {rotation, angular_velocity}
    float2 *mg;    // {calculated_mass, gravity_mod}
    uint32_t *bflags;
    float2 *off;   // collider offset translation
    uint2 *groups;
    float *inertia, *torque;
    float2 *acc, *vreq;
};

constexpr uint32_t kSnapOutlier = 0x80000000u;
// A single world below 2^24 bodies also carries its body flags (bits 0-5) in bits 24-29 of
// the snapshot word, so a steady substep reads them with the snapshot; its ids are then the
// low 24 bits.
constexpr uint32_t kSnapFlagShift = 24;
constexpr uint32_t kSnapFlagMask = 0x3Fu;
constexpr uint32_t kSnapIdMask = 0x00FFFFFFu;
constexpr uint32_t kNil = 0xFFFFFFFFu;
constexpr uint32_t kNbrCap = 32;      // partner-list entries per body (sphere path)
constexpr uint8_t kNbrNone = 0xFF;    // no list: the body walks the grid

constexpr uint32_t kJointBodies = 16, kJointJoints = 16;

// Ordered contact pass (general path).  A push-out (physics.rs:272-286)
// rewrites the two colliders' snapshots from their bodies' live positions,
// and every later pair of the (i, j<i) loop reads the rewritten values, so
// the pass stops being order-independent once a pair's snapshots coincide.
// k_contacts evaluates every pair on the substep's snapshot and records the
// rows (A colliders) of coincident pairs; k_contacts_ordered then replays
// the pairs whose outcome can differ ("dirty" pairs: a coincident pair, or a
// pair one of whose colliders was rewritten at an earlier key) in reference
// order and re-sums the positions of the bodies they touch.
struct OrderedPair {
    unsigned long long key;  // (A << 32) | B, A the higher collider index
    int32_t pa, pb;          // parent bodies of A and B
    uint32_t flags;          // OP_* bits
    uint32_t pad;
    float2 ca, cb;           // contact corrections of pa and pb (the push-out is separate)
};
enum : uint32_t {
    OP_CONTACT = 1u,  // a contact when replayed in order
    OP_PUSHED = 2u,   // took the push-out branch
    OP_ORIG = 4u,     // a contact on the substep's snapshot (counted and sent by k_contacts)
    OP_SENSOR = 8u,   // a sensor is involved: no correction
};
// One addend of a body's ordered position sum.
struct OrderedTerm {
    unsigned long long key;
    float2 push;  // push-out, added first (OT_PUSH)
    float2 c;     // contact correction (OT_CORR)
    uint32_t flags;
    uint32_t pad;
};
enum : uint32_t { OT_PUSH = 1u, OT_CORR = 2u };
constexpr uint32_t kOrdCand = 8192;   // candidate colliders of one replayed row
constexpr uint32_t kOrdTerms = 8192;  // addends of one body's position sum

// Persistent batched step (batch_kernels.hip): the worlds of a batch with
// kPersistMinBodies to kLdsBodies bodies, each run by ONE workgroup of
// kPersistClassThreads threads for every substep of a launch, bodies in
// registers (kPersistPer per thread), snapshot and grid in LDS.  Smaller
// worlds are cheaper on the grid path: a world of a few hundred bodies is
// too little parallel work for a workgroup's serial chain of grid-build
// phases per substep (measured on the 4,096-world batch: 0.45 vs 0.20 ms per
// step for its worlds up to 256 bodies, 0.59 vs 0.41 up to 1,024, 0.73 vs
// 1.12 from 1,025 to 4,096).
constexpr uint32_t kPersistPer = 4;
constexpr uint32_t kPersistClasses = 1;
constexpr uint32_t kPersistClassThreads[kPersistClasses] = {1024};
constexpr uint32_t kLdsBodies = kPersistPer * 1024;
constexpr uint32_t kPersistMinBodies = 1025;

struct DeviceWorld {
    uint32_t nb = 0, nc = 0, ns = 0, nj = 0, nk = 0;
    uint32_t cells_cap = 0;
    int key_bits = 0;

    // general path
    BodyArrays b{};
    ColliderArrays c{};
    float2 *x0 = nullptr;  // collider snapshot at the last rebuild (sphere path: per slot)
    uint32_t *bc_start = nullptr, *bc_list = nullptr;  // body -> colliders
    // sphere path
    SphereArrays s{}, s_alt{};
    float4 *snap_next = nullptr;
    // partner lists, written by the first substep after a rebuild: 16-bit
    // slot deltas of the bodies whose rebuild snapshots lie within r_a + r_b
    // + 2 skin, two a word, [kNbrCap/2][nb]; nbr_len[s] == kNbrNone: walk
    // the grid instead
    uint32_t *nbr = nullptr;
    uint8_t *nbr_len = nullptr;
    // outlier cells: hashed linked lists (plus an occupancy bitmap) of the
    // bodies that moved more than the skin since the last rebuild, refilled
    // every substep (3 rotating tables)
    uint32_t *oh_head[3] = {nullptr, nullptr, nullptr};
    uint32_t *oh_next[3] = {nullptr, nullptr, nullptr};
    uint32_t *oh_bits[3] = {nullptr, nullptr, nullptr};
    // per wave of the substep launch, rotating with the tables: 1 if an outlier of the
    // previous substep may touch one of the wave's bodies
    uint32_t *oh_wave[3] = {nullptr, nullptr, nullptr};
    uint32_t n_waves = 0;
    uint32_t oh_size = 0;  // power of two, >= 32
    // rebuild: bodies per cell, and the slot -> previous slot permutation the
    // first substep after a rebuild gathers its records through
    uint32_t *cell_count = nullptr;
    uint32_t *perm = nullptr;
    bool gather_pending = false;
    // worlds of the sphere path (one for a plain world, many for a batch):
    // bodies of world w keep the slot range [body_start[w], body_start[w+1])
    // through every sort because their cell keys start at cell_base[w]
    uint32_t n_worlds = 0;
    uint32_t *wid = nullptr;        // world of each slot (batch only)
    uint32_t *cell_base = nullptr;  // [n_worlds + 1]
    GridParams *grids = nullptr;    // [n_worlds]
    uint32_t *wbbox = nullptr;      // [n_worlds][4] f2ord(min x, min y, max x, max y) of the snapshot being sorted
    float *wold_dt = nullptr;       // [n_worlds] Physics::old_dt of each world (Q1)
    float *wstat = nullptr;         // [n_worlds][stat_rep][8] rebuild statistics (sums; see k_world_stats)
    uint32_t stat_rep = 1;          // power of two, stat_rep * n_worlds <= 4096
    float2 *wvel = nullptr;         // [n_worlds] mean velocity of the dynamic bodies at the last rebuild
    uint32_t total_cells = 0;       // cell_base[n_worlds]
    uint32_t *tile_seg = nullptr;   // batch: per scan tile of the cell table, the world holding its first cell
    // springs
    int32_t *sp_a = nullptr, *sp_b = nullptr;
    float *sp_rest = nullptr, *sp_k = nullptr, *sp_c = nullptr;
    uint32_t *bs_start = nullptr, *bs_list = nullptr;
    // joints, grouped into independent components (disjoint body sets)
    int32_t *j_a = nullptr, *j_b = nullptr;
    float2 *j_anchor_a = nullptr, *j_anchor_b = nullptr;
    float *j_dist = nullptr, *j_target = nullptr;
    // per joint: 1/m_a + 1/m_b, (1/m_a)/ims, 1 - that, and 1 if both masses are positive
    // (physics.rs:447-460; the same f32 operations, evaluated at upload)
    float4 *j_mass = nullptr;
    // deferred joint rotations: per solve (component-major, then iteration, then joint),
    // the axis record (dx, dy, target angle, 1; 0 for a skipped solve); per body, its joints
    // in arena order (entry = position in jc_list << 1 | 1 if the body is the joint's b) and
    // component
    uint32_t j_iters = 0;  // joint_iterations the records are sized for
    float4 *jrec = nullptr;
    uint32_t *bj_start = nullptr, *bj_list = nullptr;
    int32_t *bj_comp = nullptr;
    uint32_t *jc_start = nullptr, *jc_list = nullptr;
    uint32_t n_jcomp = 0;
    // components of at most kJointBodies bodies and kJointJoints joints iterate in LDS:
    // their bodies (jb_start/jb_list, empty for a larger component) and, per entry of
    // jc_list, the joint's two bodies as local indices (a | b << 8)
    uint32_t *jb_start = nullptr, *j_local = nullptr;
    int32_t *jb_list = nullptr;
    // per entry of jc_list: the joint's anchors (a.x, a.y, b.x, b.y), mass terms and
    // (distance, target angle), so that a block stages its components' joints coalesced
    float4 *jq_anchor = nullptr, *jq_mass = nullptr;
    float2 *jq_len = nullptr;
    // constraints
    float2 *con_pos = nullptr;
    float *con_r = nullptr;
    // broadphase
    uint32_t *keys = nullptr, *vals = nullptr, *keys_tmp = nullptr, *vals_tmp = nullptr;
    uint32_t *cell_start = nullptr;
    float4 *crec = nullptr;  // general path: (x, y, radius, parent bits) of each collider, in cell order
    GridParams *grid = nullptr;
    // contacts in two kernels (general path): per collider, its kColList smallest-key
    // corrections [kColList][nc] as (key hi, key lo, cx, cy) and its contact count; the
    // bodies the per-body merge leaves to the full walk (count in cfb_n[0])
    float4 *clist = nullptr;
    uint32_t *ccnt = nullptr;
    uint32_t *cfb_list = nullptr, *cfb_n = nullptr;
    // ordered contact pass (general path; see OrderedPair)
    uint32_t *ord_rows = nullptr;            // [nc] rows of coincident pairs, count in ord_n[0]
    uint32_t *ord_n = nullptr;               // [1]
    uint32_t *ord_mark = nullptr;            // [ceil(nc / 32)] row queued (bitmap)
    float2 *pos_pre = nullptr;               // [nb] body positions before the contact pass
    unsigned long long *mv_key = nullptr;    // [nc] key of a collider's first snapshot rewrite (~0: none)
    float2 *mv_abs = nullptr;                // [nc] its current rewritten snapshot
    uint32_t *mv_list = nullptr;             // [nc] rewritten colliders
    uint32_t *ord_heap = nullptr;            // [nc] rows to replay (min-heap)
    OrderedPair *ord_log = nullptr;          // dirty pairs in key order
    uint32_t ord_log_cap = 0;
    uint32_t *aff_list = nullptr;            // [nb] bodies of dirty pairs
    uint32_t *aff_mark = nullptr;            // [ceil(nb / 32)]
    uint32_t *ord_cand = nullptr;            // [kOrdCand]
    OrderedTerm *ord_terms = nullptr;        // [kOrdTerms]
    StepAccum *accum = nullptr;
    unsigned long long *block_contacts = nullptr;  // contact count per block of the per-body kernels
                                                   // (ceil(nb / 64) slots: blocks of 64 or more threads)
    uint32_t n_block_slots = 0;
    SortTemp sort{};

    // persistent batched step (batch_kernels.hip): the persistent worlds by size class
    // (largest first in each), the queue counters, and per world whether the grid path
    // skips it (it does when the persistent step runs the world)
    uint32_t *p_order = nullptr, *p_wstart = nullptr, *p_counter = nullptr;
    uint8_t *p_skip = nullptr;
    uint32_t p_class_start[kPersistClasses + 1] = {};
    uint32_t p_groups[kPersistClasses] = {};
    uint32_t p_grid_worlds = 0;  // worlds left to the grid path
    // the persistent launch runs on its own stream beside the grid path's substeps
    // (default priority; forked from and joined to `stream` by events)
    hipStream_t p_stream = nullptr;
    hipEvent_t p_fork = nullptr, p_join = nullptr;
    // collision events (optional): buffer of event_cap records, count in accum
    DevEvent *events = nullptr;
    uint32_t event_cap = 0;
    float2 *vel_next = nullptr;  // sphere path with events: new velocities (the old ones stay readable)

    hipStream_t stream = nullptr;
    int device = -1;
    bool allocated = false;
    // stage timing of the fused substep: when set, its launch records these events
    // at the kernel's own start and end (hipExtLaunchKernelGGL)
    hipEvent_t prof_start = nullptr, prof_stop = nullptr;
};

// Arguments of one persistent launch.
struct PersistArgs {
    // sphere-path records, slot = dense index (the grid path never moves the slots of a
    // persistent world): read from `src`, where the previous persistent launch left them,
    // written to `dst`, the records current after the grid path's substeps of this launch
    SphereArrays src, dst;
    const uint32_t* wstart;  // [n_worlds + 1] slot range of each world
    float* wold_dt;          // [n_worlds] Physics::old_dt (Q1)
    const uint32_t* order;   // persistent worlds by class, largest first within a class
    uint32_t class_start[kPersistClasses + 1];  // class c: order[class_start[c], class_start[c+1])
    uint32_t groups[kPersistClasses];           // workgroups launched per class
    uint32_t* counter;                          // [kPersistClasses] queue positions
    StepAccum* accum;
    unsigned long long* block_contacts;
    float h, ratio_rest, gx, gy;
    uint32_t nsub;        // substeps of this launch
    uint32_t fresh;       // its first substep consumes user accelerations and torques
    uint32_t collisions;
};

hipError_t launch_batch_persistent(const PersistArgs& k, hipStream_t s);

// Scalars of one substep launch.
struct SubstepArgs {
    float h;            // step_delta = delta / substeps (physics.rs:399)
    float ratio_rest;   // h / h: displacement ratio of every non-static body but the first (physics.rs:338)
    float gx, gy;       // Physics::gravity
    float rmax;         // largest collider radius
    float mass_per_r;   // sphere path: every body's mass is this times its radius (0: load the masses)
    uint32_t gmod_one;  // sphere path: every body's gravity_mod is exactly 1
    uint32_t snap_flags;  // sphere path: snapshot words carry the body flags (ids in the low 24 bits)
    float drift_prev_x, drift_prev_y;  // common drift at the previous snapshot (broadphase query centre)
    float drift_x, drift_y;            // common drift after this substep
    float skin;         // sphere path: displacement since the rebuild beyond which a body is an outlier
    float t_prev, t_cur;  // sphere path: time since the rebuild at the previous / this snapshot
    int32_t first_dyn;  // dense index of the first non-static body (Q1), -1 if none
    uint32_t fresh;     // first substep after an upload: user acc/torque are live
    uint32_t collisions;
    uint32_t parity;    // which bbox accumulator this substep writes
    uint32_t has_springs;
    uint32_t sub;       // sphere path: substeps since the last rebuild (1-based)
    uint32_t sub_global;  // substeps since the world was created (orders collision events)
    uint32_t write_vel;   // sphere path: store calculated_velocity (last substep of a step, or
                          // every substep when collision events read it)
    uint32_t joint_rot;   // general path: k_integrate applies the joint stage's rotation corrections
    uint32_t joint_iters;
};

// A general-path rebuild's key pass (step_kernels.hip): the grid from the
// bounding box accumulated under `parity`, then every collider's cell key
// and rank.
struct KeysArgs {
    const float2* xy;
    uint32_t nc;
    StepAccum* accum;
    uint32_t parity;
    float cell;
    uint32_t cells_cap;
    GridParams* grid_out;
    uint32_t* keys;
    uint32_t* rank;
    uint32_t* cnt;
    float2* x0;
};

// Launchers (step_kernels.hip: general path + broadphase build).  A substep
// that rebuilds passes its key pass (launched with the substep's forces).
hipError_t launch_rebuild_sort_and_cells(DeviceWorld& d);
hipError_t launch_general_substep(DeviceWorld& d, const SubstepArgs& a, uint32_t joint_iterations, bool groups,
                                  bool sensors, const KeysArgs* rebuild);
// Launchers (sphere_kernels.hip).
hipError_t launch_sphere_rebuild(DeviceWorld& d, float cell, uint32_t parity);
hipError_t launch_sphere_substep(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors, bool zero_off,
                                 bool constraints);
// Launcher (ordered_contacts.hip): the ordered replay after k_contacts (exits at once
// when no pair of the substep's snapshot coincides).
hipError_t launch_contacts_ordered(DeviceWorld& d, const SubstepArgs& a, bool groups, bool sensors);
// fills n copies of v (graph-safe replacement of hipMemsetAsync)
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

}  // namespace bg
