/* A C program driving the MI355X step through include/blobs_gpu.h only: the
 * calls a Rust binding (or any FFI) makes.  Two overlapping spheres under
 * gravity, ten steps of Physics::step (physics.rs:78), then the checks a
 * caller would rely on.  Exit status: 0 ok, 3 no HIP device (the step
 * refuses to run: there is no CPU fallback), 1 a check failed. */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "blobs_gpu.h"

#define CHECK(cond, msg)                                \
    do {                                                \
        if (!(cond)) {                                  \
            fprintf(stderr, "check failed: %s\n", msg); \
            return 1;                                   \
        }                                               \
    } while (0)

/* RigidBodyBuilder::new().position(p).build() (rigid_body.rs:287-400). */
static bg_rigid_body body_at(float x, float y) {
    bg_rigid_body b;
    memset(&b, 0, sizeof b);
    b.position[0] = b.position_old[0] = x;
    b.position[1] = b.position_old[1] = y;
    b.calculated_mass = 1.0f;
    b.gravity_mod = 1.0f;
    b.inertia = 1.0f;
    b.scale[0] = b.scale[1] = 1.0f;
    b.body_type = BG_DYNAMIC;
    return b;
}

/* ColliderBuilder::new().radius(r).absolute_transform_translation(p).build() (collider.rs:195-283). */
static bg_collider ball_at(float x, float y, float r) {
    bg_collider c;
    memset(&c, 0, sizeof c);
    c.offset_matrix[0] = c.offset_matrix[3] = 1.0f;
    c.absolute_matrix[0] = c.absolute_matrix[3] = 1.0f;
    c.absolute_translation[0] = x;
    c.absolute_translation[1] = y;
    c.radius = r;
    c.memberships = 0xFFFFFFFFu;
    c.filter = 0xFFFFFFFFu;
    return c;
}

int main(void) {
    bg_world *w = NULL;
    CHECK(bg_world_create(0.0f, -9.81f, 0, &w) == BG_OK && w, "bg_world_create");
    bg_handle ha, hb, ca, cb;
    bg_rigid_body a = body_at(0.0f, 0.0f), b = body_at(0.8f, 0.0f);
    CHECK(bg_insert_rbd(w, &a, &ha) == BG_OK, "insert a");
    CHECK(bg_insert_rbd(w, &b, &hb) == BG_OK, "insert b");
    bg_collider c1 = ball_at(0.0f, 0.0f, 0.5f), c2 = ball_at(0.8f, 0.0f, 0.5f);
    CHECK(bg_insert_collider_with_parent(w, &c1, ha, &ca) == BG_OK, "collider a");
    CHECK(bg_insert_collider_with_parent(w, &c2, hb, &cb) == BG_OK, "collider b");

    /* reference panics become statuses with the reference's message */
    bg_handle hj;
    const float anchor[2] = {0.0f, 0.0f};
    CHECK(bg_create_fixed_joint(w, ha, ha, anchor, anchor, &hj) == BG_ERR_PANIC, "joint on one body panics");
    char msg[256];
    bg_last_error(w, msg, sizeof msg);
    CHECK(msg[0] != '\0', "panic message");
    bg_rigid_body none;
    CHECK(bg_get_rbd(w, ((bg_handle)7 << 32) | 9u, &none) == BG_ERR_NOT_FOUND, "stale handle");

    int st = bg_step(w, 1.0 / 60.0, 10);
    if (st == BG_ERR_NO_DEVICE) {
        printf("c-abi example: no HIP device (step refused)\n");
        bg_world_destroy(w);
        return 3;
    }
    CHECK(st == BG_OK, "bg_step");
    bg_rigid_body ra, rb;
    CHECK(bg_get_rbd(w, ha, &ra) == BG_OK && bg_get_rbd(w, hb, &rb) == BG_OK, "bg_get_rbd");
    const float dx = rb.position[0] - ra.position[0], dy = rb.position[1] - ra.position[1];
    const float d = sqrtf(dx * dx + dy * dy);
    CHECK(isfinite(d) && d > 0.99f, "contacts separated the spheres");
    CHECK(ra.position[1] < 0.0f && rb.position[1] < 0.0f, "gravity pulled both down");
    uint64_t contacts = 0;
    CHECK(bg_contact_count(w, &contacts) == BG_OK && contacts >= 1, "contact counter");
    printf("c-abi example ok: separation %.4f after 10 steps, %llu contacts\n", d, (unsigned long long)contacts);
    CHECK(bg_world_destroy(w) == BG_OK, "destroy");
    return 0;
}
