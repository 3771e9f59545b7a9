/*
 * blobs_gpu.h -- C ABI of the MI355X-native `Physics::step` drop-in.
 *
 * The reference (darthdeus/blobs) is a Rust crate with no FFI; its boundary is
 * the `Physics` value and its methods (blobs/src/physics.rs:3-239).  Every
 * entry point below replaces one of those methods (cited per function) so a
 * Rust `-sys` crate can bind them 1:1 and keep the reference API on top.
 *
 * Conventions
 *  - Every call returns a bg_status (0 = OK).  No exception or panic crosses
 *    the ABI.  Where the reference panics or asserts (unwrap on a missing
 *    body, assert! in the solver), the call returns BG_ERR_PANIC and
 *    bg_last_error() holds the message; a Rust facade maps that to panic!.
 *  - Handles are thunderdome `Index::to_bits()` values:
 *    (generation << 32) | slot.  Iteration order is slot order, as in the
 *    reference arenas.
 *  - A world owns its device memory and one HIP stream.  Calls on one world
 *    are not thread-safe (the reference `Physics` is !Send); distinct worlds
 *    may be driven from distinct threads.
 *  - Host-side construction (insert, create joint, mass setup) needs no GPU.
 *    bg_step and friends need one and return BG_ERR_NO_DEVICE otherwise;
 *    there is no CPU fallback.
 */
#ifndef BLOBS_GPU_H
#define BLOBS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BG_ABI_VERSION 1

typedef struct bg_world bg_world;
typedef uint64_t bg_handle;

typedef enum {
    BG_OK = 0,
    BG_ERR_INVALID_ARGUMENT = 1,
    BG_ERR_NO_DEVICE = 2,   /* no HIP device: GPU entry points refuse to run */
    BG_ERR_HIP = 3,         /* a HIP runtime call failed */
    BG_ERR_NOT_FOUND = 4,   /* stale or unknown handle (reference: None) */
    BG_ERR_PANIC = 5,       /* the reference would panic/assert here */
    BG_ERR_CAPACITY = 6,    /* a fixed-size device buffer overflowed */
    BG_ERR_UNSUPPORTED = 7, /* reference path that itself panics */
} bg_status;

/* RigidBodyType, rigid_body.rs:221-242 (same discriminants). */
enum { BG_DYNAMIC = 0, BG_STATIC = 1, BG_KINEMATIC_POSITION_BASED = 2, BG_KINEMATIC_VELOCITY_BASED = 3 };

/* RigidBody fields, rigid_body.rs:42-74.  Used by insert (as the builder
 * output, RigidBodyBuilder::build rigid_body.rs:376-400) and by get/set. */
typedef struct {
    float position[2];
    float position_old[2];
    float center_of_mass[2];
    float calculated_mass;
    float gravity_mod;
    float rotation;
    float angular_velocity;
    float torque;
    float inertia;
    float scale[2];
    float acceleration[2];
    float velocity_request[2];
    uint32_t has_velocity_request; /* Option<Vec2> */
    float calculated_velocity[2];
    uint32_t body_type;
    uint64_t user_data[2]; /* u128, little-endian halves */
} bg_rigid_body;

/* Collider fields, collider.rs:4-20 (shape omitted: physics reads radius,
 * collider.rs:11).  Affine2 = Mat2 columns (x_axis, y_axis) + translation. */
typedef struct {
    float offset_matrix[4];
    float offset_translation[2];
    float absolute_matrix[4];
    float absolute_translation[2];
    uint64_t user_data[2];
    bg_handle parent; /* 0 with has_parent = 0 means None */
    uint32_t has_parent;
    float radius;
    float mass_override;
    uint32_t has_mass_override;
    uint32_t is_sensor;
    uint32_t memberships; /* InteractionGroups, groups.rs:7-12 */
    uint32_t filter;
} bg_collider;

/* Spring, springs.rs:16-22. */
typedef struct {
    bg_handle rigid_body_a;
    bg_handle rigid_body_b;
    float rest_length;
    float stiffness;
    float damping;
} bg_spring;

/* FixedJoint, joints.rs:7-19. */
typedef struct {
    bg_handle rigid_body_a;
    bg_handle rigid_body_b;
    float anchor_a[2];
    float anchor_b[2];
    float distance;
    float target_angle;
} bg_fixed_joint;

/* Physics pub fields that steer the step, physics.rs:6-9,20,27,30-33. */
typedef struct {
    float gravity[2];
    uint32_t substeps;         /* default 8, physics.rs:46 */
    uint32_t joint_iterations; /* default 4, physics.rs:47 */
    uint32_t collisions_enabled;
    uint32_t use_spatial_hash; /* true makes the reference panic, physics.rs:411-412 */
    double accumulator;
    double time;
    float old_dt;
} bg_params;

/* Bulk body/collider description, structure-of-arrays.  Inserting n entries
 * with one call is defined as n sequential single inserts. */
typedef struct {
    const float *position;         /* [2n] */
    const float *velocity_request; /* [2n] or NULL */
    const uint8_t *has_velocity_request; /* [n] or NULL (all true if velocity_request given) */
    const float *gravity_mod;      /* [n] or NULL (1.0) */
    const float *rotation;         /* [n] or NULL (0.0) */
    const uint8_t *body_type;      /* [n] or NULL (Dynamic) */
} bg_body_batch;

typedef struct {
    const float *offset_translation;   /* [2n] or NULL (identity offset) */
    const float *absolute_translation; /* [2n] or NULL (identity, quirk of collider.rs:214-215) */
    const float *radius;               /* [n] */
    const float *mass_override;        /* [n] or NULL */
    const uint8_t *has_mass_override;  /* [n] or NULL */
    const uint8_t *is_sensor;          /* [n] or NULL */
    const uint32_t *memberships;       /* [n] or NULL (ALL) */
    const uint32_t *filter;            /* [n] or NULL (ALL) */
} bg_collider_batch;

/* ---- world lifetime ------------------------------------------------- */

/* Physics::new(gravity, use_spatial_hash), physics.rs:37-69. */
int bg_world_create(float gravity_x, float gravity_y, int use_spatial_hash, bg_world **out);
/* Drop for Physics. */
int bg_world_destroy(bg_world *w);
/* Physics::reset, physics.rs:71-76 (clears bodies, colliders, joints, springs). */
int bg_world_reset(bg_world *w);

/* Read/write the pub fields listed in bg_params (physics.rs:3-34). */
int bg_get_params(bg_world *w, bg_params *out);
int bg_set_params(bg_world *w, const bg_params *p);

/* ---- construction (host side, no GPU needed) ------------------------ */

/* Physics::insert_rbd, physics.rs:121-128. */
int bg_insert_rbd(bg_world *w, const bg_rigid_body *body, bg_handle *out);
/* Physics::insert_collider_with_parent, physics.rs:130-149 (the collider is
 * pushed twice into rbd.colliders and mass/inertia recomputed, exactly as
 * collider.rs:179-181 + physics.rs:144-146 do). */
int bg_insert_collider_with_parent(bg_world *w, const bg_collider *col, bg_handle parent, bg_handle *out);
/* n x (RigidBodyBuilder::new().position(p).velocity_request(v)...build() +
 * insert_rbd), the demo recipe of demo/src/simulation.rs:87-97. */
int bg_insert_rbds(bg_world *w, uint32_t n, const bg_body_batch *b, bg_handle *out);
/* n x insert_collider_with_parent (physics.rs:130-149). */
int bg_insert_colliders_with_parents(bg_world *w, uint32_t n, const bg_collider_batch *c, const bg_handle *parents,
                                     bg_handle *out);
/* Physics::create_fixed_joint, physics.rs:184-207. */
int bg_create_fixed_joint(bg_world *w, bg_handle a, bg_handle b, const float anchor_a[2], const float anchor_b[2],
                          bg_handle *out);
/* Physics::create_fixed_joint_with_distance, physics.rs:209-239. */
int bg_create_fixed_joint_with_distance(bg_world *w, bg_handle a, bg_handle b, const float anchor_a[2],
                                        const float anchor_b[2], float distance, bg_handle *out);
/* physics.springs.insert(Spring { .. }) (pub arena, physics.rs:14). */
int bg_insert_spring(bg_world *w, const bg_spring *s, bg_handle *out);
/* physics.constraints.push(Constraint { position, radius }) (physics.rs:15, lib.rs:189-193). */
int bg_push_constraint(bg_world *w, float x, float y, float radius);
/* Physics::remove_rbd, physics.rs:163-172. */
int bg_remove_rbd(bg_world *w, bg_handle h);
/* Physics::remove_col, physics.rs:159-161 (ColliderSet::remove, collider.rs:134-164). */
int bg_remove_col(bg_world *w, bg_handle h);

/* ---- stepping (GPU) -------------------------------------------------- */

/* Physics::step(delta), physics.rs:78-82, repeated n_steps times. */
int bg_step(bg_world *w, double delta, uint32_t n_steps);
/* Physics::fixed_step(frame_time), physics.rs:84-99. */
int bg_fixed_step(bg_world *w, double frame_time);
/* Wait for queued work on the world's stream (and report a deferred panic). */
int bg_synchronize(bg_world *w);
/* Deferred checks (enable != 0): bg_step returns once its kernels are queued,
 * without waiting for the GPU; the step's panic flags (physics.rs:293,
 * 447-448, 471-474) travel back behind the kernels and are reported by a
 * later bg_step or by bg_synchronize.  Default off: every bg_step waits and
 * reports its own panic, as the reference's step panics in place.  No
 * counterpart in the reference (its step runs on the calling thread). */
int bg_set_deferred_checks(bg_world *w, int enable);
/* n_chunks x chunk_steps calls of Physics::step(delta) (physics.rs:78-82),
 * each chunk replayed from a hipGraph captured once (launch overhead paid
 * once per capture, not per kernel).  Same results as bg_step with
 * n_steps = chunk_steps * n_chunks.  The reference has no counterpart: it
 * is the multi-step driver the batched benchmark loop needs. */
int bg_batch_run(bg_world *w, double delta, uint32_t chunk_steps, uint32_t n_chunks);

/* ---- batches of worlds (GPU-native; no counterpart in the reference) --
 * A batch steps many independent `Physics` worlds (physics.rs:11-34 each)
 * with one launch per substep.  Every source world must be a single-sphere
 * world (each body has exactly one collider, inserted in the same order;
 * no springs, joints or circle constraints) and all must share substeps,
 * gravity and collisions_enabled.  The batch copies their state; each
 * world keeps its own old_dt (Q1, physics.rs:338-340), time and grid, and
 * contacts never cross worlds, so every world evolves bit-identically to
 * stepping it alone.  The batch is itself a bg_world: bg_step,
 * bg_batch_run, bg_autotune and the getters work on it (bodies in world
 * order).  Destroy it with bg_world_destroy. */
int bg_batch_create(bg_world *const *worlds, uint32_t n_worlds, bg_world **out);
/* Copy the batch's state back into its source worlds (same order). */
int bg_batch_writeback(bg_world *batch, bg_world *const *worlds, uint32_t n_worlds);
/* Number of worlds in a batch (0 for a plain world). */
int bg_batch_size(bg_world *w, uint32_t *out);
/* Step the batch's worlds of 1,025 to 4,096 bodies with the persistent
 * kernel (default off): one launch per bg_step / chunk, one workgroup per
 * world taken from a largest-first queue, every substep of the launch inside
 * the workgroup with the world's snapshot, grid and state on chip; the other
 * worlds step on the grid path.  In effect from the next step when the
 * batch's colliders are centred with default groups and no sensors, and no
 * collision events are recorded; otherwise (or with enable = 0) every world
 * steps on the grid path.  Results are the same bits either way.  With it
 * on, bg_autotune times both schedules and keeps the faster.  No counterpart
 * in the reference (it has no batches). */
int bg_batch_set_persistent(bg_world *w, int enable);
/* *active = 1 if the batch's last upload gave worlds to the persistent kernel. */
int bg_batch_persistent(bg_world *w, int *active);

/* ---- state access ---------------------------------------------------- */

/* Physics::rbd_count, physics.rs:113-115. */
int bg_rbd_count(bg_world *w, uint32_t *out);
/* Physics::get_rbd, physics.rs:105-107 (BG_ERR_NOT_FOUND for None). */
int bg_get_rbd(bg_world *w, bg_handle h, bg_rigid_body *out);
/* Physics::get_mut_rbd + write-back of every pub field, physics.rs:109-111. */
int bg_set_rbd(bg_world *w, bg_handle h, const bg_rigid_body *in);
/* Physics::get_col, physics.rs:117-119. */
int bg_get_col(bg_world *w, bg_handle h, bg_collider *out);
/* physics.joints.get(h) (pub arena, physics.rs:13). */
int bg_get_joint(bg_world *w, bg_handle h, bg_fixed_joint *out);
/* physics.springs.get(h) (pub arena, physics.rs:14). */
int bg_get_spring(bg_world *w, bg_handle h, bg_spring *out);
/* Bulk read of body state in arena (iteration) order: each pointer may be
 * NULL; arrays hold rbd_count entries (x,y interleaved for Vec2 fields). */
int bg_get_bodies(bg_world *w, float *position, float *position_old, float *velocity, float *rotation,
                  float *angular_velocity, bg_handle *handles);
/* Bulk read of collider absolute translations in arena order. */
int bg_get_collider_positions(bg_world *w, float *absolute_translation, uint32_t *count);
/* perf counter "collisions" (physics.rs:316): contacts counted since creation. */
int bg_contact_count(bg_world *w, uint64_t *out);
/* Text of the last error on this world (or the global one if w is NULL). */
int bg_last_error(bg_world *w, char *buf, size_t len);

/* ---- collision events and debug data ---------------------------------- */

/* One CollisionEvent (lib.rs:146-153): collider A is the later one in
 * iteration order; velocities are the parents' calculated_velocity before
 * the contact pass (physics.rs:288-289). */
typedef struct {
    bg_handle col_handle_a, col_handle_b;
    float impact_vel_a[2], impact_vel_b[2];
} bg_collision_event;

/* The reference sends every contact into the world's mpsc channel
 * (physics.rs:38, 304-311).  Recording is off by default; with capacity
 * > 0 the device keeps up to `capacity` events between drains (more are
 * counted as dropped).  Re-uploads keep pending events. */
int bg_set_event_capacity(bg_world *w, uint32_t capacity);
/* Physics::collision_recv drained: up to `cap` events in the reference's
 * order (substep, then A, then B -- the (i, j<i) loop of physics.rs:248-249);
 * the rest stay queued.  *n_dropped: events lost to a full buffer since the
 * last drain. */
int bg_drain_events(bg_world *w, bg_collision_event *out, uint32_t cap, uint32_t *n_out, uint64_t *n_dropped);

typedef struct {
    float matrix[4];     /* column-major Mat2 */
    float translation[2];
} bg_affine2;

/* Physics::debug_data (physics.rs:479-481, debug.rs:34-91): bodies
 * (Affine2::from_angle_translation(rotation, position)), colliders
 * (absolute_transform, radius), joints and springs (positions of their two
 * bodies, 4 floats each), all in arena order.  Any output may be NULL;
 * counts receives {bodies, colliders, joints, springs}. */
int bg_debug_data(bg_world *w, bg_affine2 *bodies, bg_affine2 *colliders, float *collider_radius,
                  float *joint_bodies, float *spring_bodies, uint32_t counts[4]);

/* ---- tuning ---------------------------------------------------------- */

/* Broadphase tuning: grid cell size as a multiple of 2*max collider radius
 * (>= 0.25), the skin (fraction of a cell a body may drift from its
 * rebuild position before it is tracked as an outlier; fused sphere path)
 * and the number of steps between grid rebuilds (>= 1).  Results do not
 * depend on any of them; only speed does.  Until it is called: cell scale
 * 1.25 (1.5 for a batch), skin 0.25, a rebuild every step below 65,536
 * bodies and every third step from there. */
int bg_set_broadphase(bg_world *w, float cell_scale, float skin, uint32_t rebuild_interval);
/* Which device path the last step ran: 0 none, 1 general, 2 fused spheres. */
int bg_step_path(bg_world *w, uint32_t *out);
/* Broadphase autotuner: runs steps_per_trial real steps of delta under each
 * of a fixed list of (cell_scale, skin, rebuild_interval) candidates, timed
 * with HIP events, and keeps the fastest (reported through the non-NULL
 * out pointers).  The steps advance the simulation as bg_step would; the
 * choice changes speed only. */
int bg_autotune(bg_world *w, double delta, uint32_t steps_per_trial, float *cell_scale, float *skin,
                uint32_t *rebuild_interval);

/* ---- differentiable rollouts (GPU-native; no counterpart in the reference)
 * A rollout batches single-sphere worlds with springs (damped or not) and
 * fixed joints (one centred, non-sensor collider per body inserted in body
 * order; no circle constraints; at most 2,048 bodies per world; worlds
 * share gravity, substeps, joint_iterations and collisions_enabled) and
 * runs n_steps calls of Physics::step(delta) (physics.rs:78-82) on copies
 * of their state, keeping a checkpoint every `checkpoint_every` steps.
 * Parameters: the initial velocity requests (one Vec2 per body,
 * rigid_body.rs:182-184) and the spring stiffnesses (springs.rs:16-22).
 * The loss of world w is |centroid(final positions) - target_w|^2.  The
 * adjoint differentiates every stage of the step (gravity, springs with
 * their damping, contacts, joint positions, Verlet) with every substep's
 * contact pair set held fixed, recomputing each segment from its
 * checkpoint.  Body rotations are not tracked: with centred colliders and
 * unrotated joint anchors (physics.rs:434-435) no position depends on them. */
typedef struct bg_rollout bg_rollout;
int bg_rollout_create(bg_world *const *worlds, uint32_t n_worlds, const float *targets_xy, uint32_t checkpoint_every,
                      bg_rollout **out);
int bg_rollout_destroy(bg_rollout *r);
int bg_rollout_sizes(bg_rollout *r, uint32_t *n_worlds, uint32_t *n_bodies, uint32_t *n_springs);
/* Replace the parameters (NULL keeps the current ones); bodies and springs
 * in world order, then iteration order. */
int bg_rollout_set_params(bg_rollout *r, const float *velocity_request, const float *stiffness);
/* Forward rollout from the worlds' initial state; loss per world (may be NULL). */
int bg_rollout_forward(bg_rollout *r, double delta, uint32_t n_steps, float *loss_per_world);
/* Gradients of the last forward: of the centroid loss summed over worlds,
 * or of sum(dloss_dfinal_position . final position) when that is given. */
int bg_rollout_backward(bg_rollout *r, const float *dloss_dfinal_position, float *grad_velocity_request,
                        float *grad_stiffness);
int bg_rollout_final_positions(bg_rollout *r, float *out);
/* After a backward: per_world[w] = 1 if world w's gradients hold inf or NaN
 * (the fixed-pair derivative of the contact projection grows with the
 * horizon and can leave the f32 range).  No counterpart in the reference,
 * which has no adjoint. */
int bg_rollout_nonfinite(bg_rollout *r, uint8_t *per_world);

/* ---- SpatialHash queries (GPU) ---------------------------------------- */

/* One cell membership of SpatialHash (spatial.rs:30-41): a point (id,
 * position, radius, spatial.rs:10-15) filed under `cell`.  Normally one per
 * point with cell = get_cell_coords(position) (spatial.rs:55-60); after
 * insert_with_id of an existing id the old cell keeps its entry
 * (spatial.rs:69-87), i.e. a second record of the id under the old cell. */
typedef struct {
    uint64_t id;
    float position[2];
    float radius;
    int32_t cell[2];
    uint32_t pad;
} bg_spatial_point;

/* One SpatialHash::query(position, query_radius) (spatial.rs:155-195). */
typedef struct {
    float position[2];
    float radius;
    uint32_t pad;
} bg_spatial_query;

/* One CellPoint of a query's result; `point` is the record's index. */
typedef struct {
    uint64_t id;
    float position[2];
    float radius;
    uint32_t point;
} bg_spatial_hit;

/* SpatialHash::query (spatial.rs:155-195) for n_queries queries over
 * n_points memberships, SpatialHash::new(cell_size) (spatial.rs:43-53).
 * All arrays are device memory.  offsets (n_queries + 1 entries) receives
 * the exclusive scan of the per-query hit counts; the hits of query q are
 * hits[offsets[q] .. offsets[q+1]): points of the 3x3 cells around the
 * query's cell (get_neighbor_cells order, spatial.rs:140-153) with
 * |p - q|^2 <= (query_radius + p.radius)^2, in record order inside a cell
 * (the reference's HashSet order is unspecified).  *total_hits = offsets[n].
 * Hits past hit_cap are not written and the call returns BG_ERR_CAPACITY
 * (offsets stay complete).  Scratch of about 40 bytes per point is
 * allocated and freed by the call; runs on the null stream and waits. */
int bg_spatial_hash_query(float cell_size, const bg_spatial_point *points, uint32_t n_points,
                     const bg_spatial_query *queries, uint32_t n_queries, uint32_t *offsets, bg_spatial_hit *hits,
                     uint32_t hit_cap, uint64_t *total_hits);

/* bg_spatial_hash_query on host arrays: copies them to the device, runs the
 * query, copies offsets and hits back (for callers whose SpatialHash lives
 * in host memory, as the reference's does). */
int bg_spatial_hash_query_host(float cell_size, const bg_spatial_point *points, uint32_t n_points,
                               const bg_spatial_query *queries, uint32_t n_queries, uint32_t *offsets,
                               bg_spatial_hit *hits, uint32_t hit_cap, uint64_t *total_hits);

/* bg_spatial_hash_query without any allocation: the same arguments, results
 * and kernels, with the scratch in a device array the library holds on each
 * GPU until the process ends (36 MiB).  For callers that cannot allocate
 * device memory (code under stream capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Limits: at most
 * 1,048,576 points and 1,048,576 queries (BG_ERR_CAPACITY beyond, nothing
 * launched).  Calls that overlap on one GPU share that array and race. */
int bg_spatial_hash_query_noalloc(float cell_size, const bg_spatial_point *points, uint32_t n_points,
                                  const bg_spatial_query *queries, uint32_t n_queries, uint32_t *offsets,
                                  bg_spatial_hit *hits, uint32_t hit_cap, uint64_t *total_hits);

/* ---- profiling ------------------------------------------------------- */

/* Per-stage timing with HIP events on the world's stream, the analogue of
 * the reference's tracy spans around `substep` and the contact pass
 * (lib.rs:195-209, physics.rs:402,242).  Stages: BG_STAGE_SUBSTEP (all
 * kernels of one substep), BG_STAGE_REBUILD (broadphase rebuild). */
enum { BG_STAGE_SUBSTEP = 0, BG_STAGE_REBUILD = 1, BG_STAGE_COUNT = 2 };
int bg_set_profiling(bg_world *w, int enable);
/* Launch count and summed event time (ms) of a stage since profiling was
 * enabled (or last reset). */
int bg_stage_stats(bg_world *w, uint32_t stage, uint64_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
