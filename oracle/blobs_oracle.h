/*
 * blobs_oracle.h -- CPU restatement of blobs' per-step simulation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in blobs_amd/ links or calls it.
 *
 * It restates, operation for operation in IEEE f32 (no contraction, correctly
 * rounded divide and sqrt, glibc sinf/cosf/atan2f), the reference crate's
 * `Physics::step` (blobs/src/physics.rs:78-82) and everything it calls:
 *   integrate                physics.rs:397-422
 *   apply_gravity            physics.rs:369-375, rigid_body.rs:207-209
 *   Spring::apply_force      springs.rs:25-47, rigid_body.rs:155-160
 *   brute_force_collisions   physics.rs:241-317, groups.rs:52-57, physics.rs:319-321
 *   solve_fixed_joints       physics.rs:424-477
 *   update_objects           physics.rs:323-367, rigid_body.rs:88-90
 *   apply_constraints        physics.rs:377-395
 * plus the construction-time mass setup (rigid_body.rs:96-128,
 * collider.rs:40-50, the double push at collider.rs:179-181 + physics.rs:144)
 * and create_fixed_joint (physics.rs:184-239).
 *
 * The same source is compiled twice: an f32 restatement (suffix _f32, the
 * parity oracle) and an FP64 twin (suffix _f64, the error oracle).
 *
 * Parity pinning: the reference is a Rust crate that cannot be built here
 * (no cargo/rustc, crates.io dependencies unavailable offline) and ships no
 * trajectory fixtures.  Each stage is pinned by closed forms, invariants and
 * brute force in tests/test_oracle.py; the one reference test on the path
 * (collider.rs:340-389, body transform * offset) is a golden vector there.
 * glibc sinf/cosf/atan2f bits and thunderdome slot reuse after removals are
 * parity unpinned.
 */
#ifndef BLOBS_ORACLE_H
#define BLOBS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Body type codes, RigidBodyType (rigid_body.rs:221-242). */
enum { ORC_DYNAMIC = 0, ORC_STATIC = 1, ORC_KINEMATIC_POS = 2, ORC_KINEMATIC_VEL = 3 };

/* Error bits raised where the reference panics or asserts. */
enum {
    ORC_ERR_NAN_NORMAL = 1u << 0,   /* physics.rs:293 assert!(!n.is_nan()) */
    ORC_ERR_JOINT_MASS = 1u << 1,   /* physics.rs:447-448 */
    ORC_ERR_JOINT_ROT = 1u << 2,    /* physics.rs:471-474 */
    ORC_ERR_EVENT_OVERFLOW = 1u << 3,
    ORC_INFO_PUSH_OUT = 1u << 8,    /* physics.rs:272-286 happened */
};

#define ORACLE_WORLD_FIELDS(R)                                                  \
    int32_t n_bodies, n_colliders, n_springs, n_joints, n_constraints;          \
    /* bodies, arena order */                                                   \
    R *pos, *pos_old, *vel, *acc, *vreq;         /* [2n] (x,y) interleaved */   \
    R *rot, *omega, *torque, *inertia, *mass, *gravity_mod; /* [n] */           \
    uint8_t *has_vreq, *body_type;                /* [n] */                     \
    /* colliders of each body (rbd.colliders, duplicates removed), CSR */       \
    int32_t *body_col_start;                      /* [n+1] */                   \
    int32_t *body_col;                                                          \
    /* colliders, arena order */                                                \
    int32_t *parent;                              /* [c] */                     \
    R *off_t, *abs_t;                             /* [2c] */                    \
    R *off_m;                                     /* [4c] Mat2 columns x0,y0,x1,y1 */ \
    R *radius;                                    /* [c] */                     \
    uint32_t *memberships, *filter;               /* [c] */                     \
    uint8_t *sensor;                              /* [c] */                     \
    /* springs, arena order */                                                  \
    int32_t *spring_a, *spring_b;                                               \
    R *spring_rest, *spring_k, *spring_c;                                       \
    /* fixed joints, arena order */                                             \
    int32_t *joint_a, *joint_b;                                                 \
    R *joint_anchor_a, *joint_anchor_b;           /* [2j] */                    \
    R *joint_dist, *joint_target;                                               \
    /* circle constraints */                                                    \
    R *con_pos;                                   /* [2k] */                    \
    R *con_radius;                                                              \
    /* parameters (Physics pub fields) */                                       \
    R gravity[2];                                                               \
    int32_t substeps, joint_iterations, collisions_enabled;                     \
    R old_dt;                                                                   \
    double time;                                                                \
    /* outputs */                                                               \
    uint64_t contact_count;                       /* perf counter "collisions" */ \
    uint32_t error_flags;                                                       \
    uint32_t push_outs;                                                         \
    /* optional collision event sink (physics.rs:304-311) */                    \
    int64_t event_cap, n_events;                                                \
    int32_t *event_cols;                          /* [2*cap] (col_a, col_b) */  \
    R *event_vels;                                /* [4*cap] vel_a, vel_b */    \
    /* 0: reference O(C^2) loop; 1: grid-enumerated candidates, same order */   \
    int32_t use_grid;

typedef struct { ORACLE_WORLD_FIELDS(float) } oracle_world_f32;
typedef struct { ORACLE_WORLD_FIELDS(double) } oracle_world_f64;

/* Physics::step(delta) repeated n_steps times (physics.rs:78-82). */
void oracle_step_f32(oracle_world_f32 *w, double delta, int32_t n_steps);
void oracle_step_f64(oracle_world_f64 *w, double delta, int32_t n_steps);

/* Physics::fixed_step(frame_time) (physics.rs:84-99); accumulator in/out. */
void oracle_fixed_step_f32(oracle_world_f32 *w, double frame_time, double *accumulator);
void oracle_fixed_step_f64(oracle_world_f64 *w, double frame_time, double *accumulator);

/* One contact pass only (physics.rs:241-317), for pair-set checks.
 * pairs_out receives (i, j) of every contact in visit order (cap entries). */
int64_t oracle_contact_pairs_f32(oracle_world_f32 *w, int32_t *pairs_out, int64_t cap);
int64_t oracle_contact_pairs_f64(oracle_world_f64 *w, int32_t *pairs_out, int64_t cap);

/* Rows [i0, i1) of the O(C^2) contact loop; returns the number of pair tests. */
int64_t oracle_bruteforce_rows_f32(oracle_world_f32 *w, int32_t i0, int32_t i1);

/*
 * Mass setup: RigidBody::update_mass_and_inertia run on the list
 * rbd.colliders as insert_collider_with_parent leaves it (each handle twice).
 * col_of_body lists, per body, its colliders in insertion order (CSR).
 * Outputs mass/inertia/center_of_mass (com [2n]).
 */
void oracle_mass_setup_f32(int32_t n_bodies, const int32_t *body_col_start, const int32_t *body_col,
                           const float *radius, const float *mass_override, const uint8_t *has_override,
                           const uint8_t *sensor, const float *off_t, float *mass, float *inertia, float *com);

/* create_fixed_joint: distance = |pa + anchor_a - pb - anchor_b| (physics.rs:198),
 * target_angle = rot_b - rot_a (physics.rs:230). */
void oracle_joint_setup_f32(const float *pa, const float *pb, const float *anchor_a, const float *anchor_b,
                            float rot_a, float rot_b, float *distance, float *target);

/* Affine2::from_angle_translation(rot, pos) * offset, translation and matrix
 * (rigid_body.rs:88-90, physics.rs:363).  out_m: Mat2 columns. */
void oracle_body_transform_mul_f32(float rot, const float *pos, const float *off_m, const float *off_t,
                                   float *out_m, float *out_t);

#ifdef __cplusplus
}
#endif
#endif
