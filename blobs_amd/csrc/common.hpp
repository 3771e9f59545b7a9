// common.hpp -- shared device helpers for the gfx950 kernels.
//
// Every kernel source in this directory is compiled with -ffp-contract=off:
// the reference (Rust) never fuses a*b+c, and bit-exact parity of the
// integration and of the contact predicate depends on separate roundings.
// f32 '/' and sqrtf are correctly rounded (HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt), matching Rust's f32 ops.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BG_WAVE 64

namespace bg {

// glam Vec2::length: sqrt((x*x) + (y*y)).
__device__ __forceinline__ float vlen(float x, float y) { return sqrtf(x * x + y * y); }

__device__ __forceinline__ bool finitef(float x) { return __builtin_isfinite(x); }

// sqrtf(d2) < r, decided without the square root.  The correctly rounded
// sqrtf(d2) is below r exactly when sqrt(d2) is below m, the midpoint of r
// and its predecessor p, i.e. when d2 < m * m: m = (p + r) / 2 and m * m are
// exact in double (m has at most 26 significant bits), and d2 never equals
// m * m (m is an odd multiple of half an ulp, so m * m needs more than 24
// bits).  A NaN or infinite d2 gives false, as sqrtf does; r <= 0 gives false.
__device__ __forceinline__ bool sqrt_lt(float d2, float r) {
    if (!(r > 0.0f)) return false;
    if (!finitef(r)) return d2 < r;
    const float p = __uint_as_float(__float_as_uint(r) - 1u);
    const double m = ((double)p + (double)r) * 0.5;
    return (double)d2 < m * m;
}

// Order-preserving map float -> uint32 for atomicMin/atomicMax.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__host__ inline uint32_t f2ord_host(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// sin/cos of a body rotation for Affine2::from_angle_translation
// (rigid_body.rs:88-90 -> glam Mat2::from_angle -> f32::sin_cos).  Zero is
// exact (sin(+-0) = +-0, cos = 1); otherwise evaluated in f64 and rounded,
// which is the correctly rounded f32 result except in rare double-rounding
// cases -- glibc's sinf/cosf agree with it to the last bit almost always.
__device__ __forceinline__ void rot_sincos(float rot, float* s, float* c) {
    if (rot == 0.0f) {
        *s = rot;
        *c = 1.0f;
    } else {
        double ds, dc;
        sincos((double)rot, &ds, &dc);
        *s = (float)ds;
        *c = (float)dc;
    }
}

__device__ __forceinline__ float atan2_f32(float y, float x) { return (float)atan2((double)y, (double)x); }

// Affine2(rot, p) * offset, translation part: Mat2*Vec2 = c0*x + c1*y with
// columns (cos, sin), (-sin, cos), then + p (physics.rs:363).
__device__ __forceinline__ float2 body_xform(float rot, float2 p, float2 off) {
    float s, c;
    rot_sincos(rot, &s, &c);
    float rx = c * off.x + (-s) * off.y;
    float ry = s * off.x + c * off.y;
    return make_float2(rx + p.x, ry + p.y);
}

// Out-of-line copy for rarely taken paths, so its f64 sincos does not
// inflate the register allocation of the calling kernel.
static __device__ __noinline__ float2 body_xform_slow(float rot, float2 p, float2 off) {
    return body_xform(rot, p, off);
}

// Same for a collider whose offset translation is (0,0): with a finite
// rotation R*0 is a signed zero, and x + (+-0) == x for every x != 0, so only
// zero components need the full evaluation; a rotation that is not finite
// makes R*0 NaN and the translation NaN, as in the reference.
__device__ __forceinline__ float2 body_xform_zero_off(float rot, float2 p) {
    if (p.x != 0.0f && p.y != 0.0f && finitef(rot)) return p;
    return body_xform_slow(rot, p, make_float2(0.0f, 0.0f));
}

// ---- wave64 / block reductions ------------------------------------------

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, BG_WAVE);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, BG_WAVE);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BG_WAVE);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BG_WAVE);
    return v;
}

// Inclusive prefix sum across a wave (Hillis-Steele over __shfl_up).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    int lane = threadIdx.x & (BG_WAVE - 1);
    for (int o = 1; o < BG_WAVE; o <<= 1) {
        uint32_t w = __shfl_up(v, o, BG_WAVE);
        if (lane >= o) v += w;
    }
    return v;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    int lane = threadIdx.x & (BG_WAVE - 1);
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Per-step reduction record: max displacement since the last rebuild, the
// bounding box of collider snapshots (order-mapped floats) and the contact
// count.  Block-reduced, then one atomic per block.
struct StepAccum {
    uint32_t dmax;      // f2ord(max |x - x0 - drift|), monotone within a rebuild period
    uint32_t pad0;
    uint32_t bbox[2][4];  // [parity]{minx, miny, maxx, maxy} as f2ord
    uint32_t n_outliers;  // sphere path: bodies beyond the skin since the last rebuild
    uint32_t error_flags;
    unsigned long long event_count;  // collision events recorded since the last drain (may exceed the buffer)
    uint32_t push_outs;
    float old_dt;  // Physics::old_dt (physics.rs:33), world scalar
};
static_assert(sizeof(StepAccum) == 64, "StepAccum layout");

// One collision event (physics.rs:302-311): the substep it happened in, the
// dense indices of collider A (the higher index) and B, and the bodies'
// calculated_velocity before the contact pass.
struct DevEvent {
    uint32_t sub, a, b, pad;
    float2 va, vb;
};

enum : uint32_t {
    ERR_NAN_NORMAL = 1u << 0,   // physics.rs:293
    ERR_JOINT_MASS = 1u << 1,   // physics.rs:447-448
    ERR_JOINT_ROT = 1u << 2,    // physics.rs:471-474
    ERR_CONTACT_CAP = 1u << 3,
    ERR_ORDERED_CAP = 1u << 4,  // a push-out cluster outgrew the ordered contact pass's buffers
    INFO_PUSH_OUT = 1u << 8,    // physics.rs:272-286: a pair's snapshots coincided
    INFO_PUSH_OUT_PAIRWISE = 1u << 9,  // ... on the fused sphere path, resolved pair by pair
};

// Body flag bits (device SoA `bflags`).
enum : uint32_t {
    BF_STATIC = 1u << 0,
    BF_HAS_VREQ = 1u << 1,
    BF_INERTIA_POS = 1u << 2,  // inertia is positive and finite
    BF_SENSOR = 1u << 3,       // sphere path: the body's collider is a sensor
    BF_FIRST_DYN = 1u << 4,    // sphere path: first non-static body of its world (Q1, physics.rs:338-340)
    BF_ROT_STILL = 1u << 5,    // sphere path: a substep without user torque leaves rotation and angular
                               // velocity bit-unchanged (static, or omega +0, finite rotation that is
                               // not -0, positive inertia: physics.rs:345-349 then adds +0)
};

struct GridParams {
    float ox, oy, cell, inv_cell;
    int32_t nx, ny;
    uint32_t ncells;
    uint32_t pad;
};

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Monotone cell coordinate: clamp(floor((x - o) * inv), 0, n-1).  Monotone
// non-decreasing in x, so range queries over clamped cells stay supersets.
__device__ __forceinline__ int32_t cell_coord(float x, float o, float inv, int32_t n) {
    float f = floorf((x - o) * inv);
    if (!(f >= 0.0f)) return 0;  // also NaN
    if (f >= (float)(n - 1)) return n - 1;
    return (int32_t)f;
}

}  // namespace bg
