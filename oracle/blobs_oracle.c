/*
 * blobs_oracle.c -- CPU restatement of blobs `Physics::step`, f32 and FP64.
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation, loaded
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include "blobs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- f32 restatement (the parity oracle) ----
 * Rust's f32::sin_cos / f32::atan2 call the platform libm (glibc's sinf,
 * cosf, atan2f on Linux), whose last bit is not pinned by any standard.
 * Built with -DORACLE_LIBM_CR (liboracle_cr.so) they are instead evaluated
 * in double and rounded to f32 -- the correctly rounded result except for
 * rare double-rounding cases -- which isolates every other operation of the
 * step from the choice of libm. */
#ifdef ORACLE_LIBM_CR
static float cr_sinf(float x) { return (float)sin((double)x); }
static float cr_cosf(float x) { return (float)cos((double)x); }
static float cr_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
#define SIN cr_sinf
#define COS cr_cosf
#define ATAN2 cr_atan2f
#else
#define SIN sinf
#define COS cosf
#define ATAN2 atan2f
#endif
#define REAL float
#define SFX(x) x##_f32
#define SQRT sqrtf
#define ISINF(x) isinf(x)
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef ATAN2
#undef ISINF

/* ---- FP64 twin (the error oracle) ---- */
#define REAL double
#define SFX(x) x##_f64
#define SQRT sqrt
#define SIN sin
#define COS cos
#define ATAN2 atan2
#define ISINF(x) isinf(x)
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef ATAN2
#undef ISINF

int64_t oracle_contact_pairs_f32(oracle_world_f32 *w, int32_t *pairs_out, int64_t cap) {
    return contacts_f32(w, pairs_out, cap);
}

int64_t oracle_contact_pairs_f64(oracle_world_f64 *w, int32_t *pairs_out, int64_t cap) {
    return contacts_f64(w, pairs_out, cap);
}

/* Rows i in [i0, i1) of the reference O(C^2) contact loop (physics.rs:248-249),
 * for timing a bounded sample of a large scene.  Returns the pair tests run. */
int64_t oracle_bruteforce_rows_f32(oracle_world_f32 *w, int32_t i0, int32_t i1) {
    uint64_t count = 0;
    int64_t np = 0, tests = 0;
    for (int32_t i = i0; i < i1; ++i) {
        for (int32_t j = 0; j < i; ++j) contact_pair_f32(w, i, j, &count, NULL, 0, &np);
        tests += i;
    }
    w->contact_count += count;
    return tests;
}

/* Collider::mass (collider.rs:40-42) and Collider::inertia (collider.rs:44-50). */
static float col_mass(float r, float mo, uint8_t has_mo) { return has_mo ? mo : r * 2.0f; }

static float col_inertia(float r, float mo, uint8_t has_mo, float ox, float oy) {
    float m = col_mass(r, mo, has_mo);
    float d = sqrtf(ox * ox + oy * oy);
    float inertia = 0.5f * m * (r * r); /* 0.5 * mass * radius.powi(2) */
    return inertia + m * (d * d);      /* + mass * d.powi(2) */
}

/*
 * RigidBody::update_mass_and_inertia (rigid_body.rs:96-128) on the list that
 * ColliderSet::insert_with_parent (collider.rs:179-181) and
 * Physics::insert_collider_with_parent (physics.rs:144) leave behind: every
 * collider handle pushed twice in a row, so the sums run c0,c0,c1,c1,...
 */
void oracle_mass_setup_f32(int32_t n_bodies, const int32_t *body_col_start, const int32_t *body_col,
                           const float *radius, const float *mass_override, const uint8_t *has_override,
                           const uint8_t *sensor, const float *off_t, float *mass, float *inertia, float *com) {
    for (int32_t b = 0; b < n_bodies; ++b) {
        float m = 0.0f, I = 0.0f, wx = 0.0f, wy = 0.0f;
        for (int32_t k = body_col_start[b]; k < body_col_start[b + 1]; ++k) {
            int32_t c = body_col[k];
            for (int rep = 0; rep < 2; ++rep) {
                if (sensor[c]) continue;
                float cm = col_mass(radius[c], mass_override[c], has_override[c]);
                m = m + cm;
                I = I + col_inertia(radius[c], mass_override[c], has_override[c], off_t[2 * c], off_t[2 * c + 1]);
                wx = wx + off_t[2 * c] * cm;
                wy = wy + off_t[2 * c + 1] * cm;
            }
        }
        if (m == 0.0f) m = 1.0f;
        if (I == 0.0f) I = 1.0f;
        mass[b] = m;
        inertia[b] = I;
        com[2 * b] = wx / m;
        com[2 * b + 1] = wy / m;
    }
}

void oracle_joint_setup_f32(const float *pa, const float *pb, const float *anchor_a, const float *anchor_b,
                            float rot_a, float rot_b, float *distance, float *target) {
    /* (rbd_a.position + anchor_a - rbd_b.position - anchor_b).length() */
    float x = ((pa[0] + anchor_a[0]) - pb[0]) - anchor_b[0];
    float y = ((pa[1] + anchor_a[1]) - pb[1]) - anchor_b[1];
    *distance = sqrtf(x * x + y * y);
    *target = rot_b - rot_a;
}

void oracle_body_transform_mul_f32(float rot, const float *pos, const float *off_m, const float *off_t,
                                   float *out_m, float *out_t) {
    float s = sinf(rot), c = cosf(rot);
    /* self.matrix2 * rhs.matrix2: each column of rhs through Mat2*Vec2 */
    for (int col = 0; col < 2; ++col) {
        float x = off_m[2 * col], y = off_m[2 * col + 1];
        out_m[2 * col] = c * x + (-s) * y;
        out_m[2 * col + 1] = s * x + c * y;
    }
    out_t[0] = (c * off_t[0] + (-s) * off_t[1]) + pos[0];
    out_t[1] = (s * off_t[0] + c * off_t[1]) + pos[1];
}
