"""ctypes binding of the C ABI in include/blobs_gpu.h (libblobs_gpu.so).

The library is the product: every GPU entry point runs HIP kernels, and
there is no CPU fallback.  Loading fails loudly when the built library is
missing; stepping fails loudly (BG_ERR_NO_DEVICE) when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libblobs_gpu.so"
SCENES_PATH = PKG / "libblobs_scenes.so"

BG_OK = 0
BG_ERR_INVALID_ARGUMENT = 1
BG_ERR_NO_DEVICE = 2
BG_ERR_HIP = 3
BG_ERR_NOT_FOUND = 4
BG_ERR_PANIC = 5
BG_ERR_CAPACITY = 6
BG_ERR_UNSUPPORTED = 7

STATUS_NAMES = {
    0: "BG_OK", 1: "BG_ERR_INVALID_ARGUMENT", 2: "BG_ERR_NO_DEVICE", 3: "BG_ERR_HIP",
    4: "BG_ERR_NOT_FOUND", 5: "BG_ERR_PANIC", 6: "BG_ERR_CAPACITY", 7: "BG_ERR_UNSUPPORTED",
}

handle_t = C.c_uint64
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class RigidBodyC(C.Structure):
    _fields_ = [
        ("position", C.c_float * 2),
        ("position_old", C.c_float * 2),
        ("center_of_mass", C.c_float * 2),
        ("calculated_mass", C.c_float),
        ("gravity_mod", C.c_float),
        ("rotation", C.c_float),
        ("angular_velocity", C.c_float),
        ("torque", C.c_float),
        ("inertia", C.c_float),
        ("scale", C.c_float * 2),
        ("acceleration", C.c_float * 2),
        ("velocity_request", C.c_float * 2),
        ("has_velocity_request", C.c_uint32),
        ("calculated_velocity", C.c_float * 2),
        ("body_type", C.c_uint32),
        ("user_data", C.c_uint64 * 2),
    ]


class ColliderC(C.Structure):
    _fields_ = [
        ("offset_matrix", C.c_float * 4),
        ("offset_translation", C.c_float * 2),
        ("absolute_matrix", C.c_float * 4),
        ("absolute_translation", C.c_float * 2),
        ("user_data", C.c_uint64 * 2),
        ("parent", handle_t),
        ("has_parent", C.c_uint32),
        ("radius", C.c_float),
        ("mass_override", C.c_float),
        ("has_mass_override", C.c_uint32),
        ("is_sensor", C.c_uint32),
        ("memberships", C.c_uint32),
        ("filter", C.c_uint32),
    ]


class SpringC(C.Structure):
    _fields_ = [
        ("rigid_body_a", handle_t),
        ("rigid_body_b", handle_t),
        ("rest_length", C.c_float),
        ("stiffness", C.c_float),
        ("damping", C.c_float),
    ]


class FixedJointC(C.Structure):
    _fields_ = [
        ("rigid_body_a", handle_t),
        ("rigid_body_b", handle_t),
        ("anchor_a", C.c_float * 2),
        ("anchor_b", C.c_float * 2),
        ("distance", C.c_float),
        ("target_angle", C.c_float),
    ]


class ParamsC(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 2),
        ("substeps", C.c_uint32),
        ("joint_iterations", C.c_uint32),
        ("collisions_enabled", C.c_uint32),
        ("use_spatial_hash", C.c_uint32),
        ("accumulator", C.c_double),
        ("time", C.c_double),
        ("old_dt", C.c_float),
    ]


class BodyBatchC(C.Structure):
    _fields_ = [
        ("position", f32p),
        ("velocity_request", f32p),
        ("has_velocity_request", u8p),
        ("gravity_mod", f32p),
        ("rotation", f32p),
        ("body_type", u8p),
    ]


class ColliderBatchC(C.Structure):
    _fields_ = [
        ("offset_translation", f32p),
        ("absolute_translation", f32p),
        ("radius", f32p),
        ("mass_override", f32p),
        ("has_mass_override", u8p),
        ("is_sensor", u8p),
        ("memberships", u32p),
        ("filter", u32p),
    ]


class CollisionEventC(C.Structure):
    _fields_ = [
        ("col_handle_a", handle_t),
        ("col_handle_b", handle_t),
        ("impact_vel_a", C.c_float * 2),
        ("impact_vel_b", C.c_float * 2),
    ]


class Affine2C(C.Structure):
    _fields_ = [("matrix", C.c_float * 4), ("translation", C.c_float * 2)]


# (name, argtypes) of every entry point in include/blobs_gpu.h
WORLD = C.c_void_p
SIGNATURES = {
    "bg_world_create": [C.c_float, C.c_float, C.c_int, C.POINTER(WORLD)],
    "bg_world_destroy": [WORLD],
    "bg_world_reset": [WORLD],
    "bg_get_params": [WORLD, C.POINTER(ParamsC)],
    "bg_set_params": [WORLD, C.POINTER(ParamsC)],
    "bg_insert_rbd": [WORLD, C.POINTER(RigidBodyC), C.POINTER(handle_t)],
    "bg_insert_collider_with_parent": [WORLD, C.POINTER(ColliderC), handle_t, C.POINTER(handle_t)],
    "bg_insert_rbds": [WORLD, C.c_uint32, C.POINTER(BodyBatchC), C.POINTER(handle_t)],
    "bg_insert_colliders_with_parents": [WORLD, C.c_uint32, C.POINTER(ColliderBatchC), C.POINTER(handle_t),
                                         C.POINTER(handle_t)],
    "bg_create_fixed_joint": [WORLD, handle_t, handle_t, f32p, f32p, C.POINTER(handle_t)],
    "bg_create_fixed_joint_with_distance": [WORLD, handle_t, handle_t, f32p, f32p, C.c_float, C.POINTER(handle_t)],
    "bg_insert_spring": [WORLD, C.POINTER(SpringC), C.POINTER(handle_t)],
    "bg_push_constraint": [WORLD, C.c_float, C.c_float, C.c_float],
    "bg_remove_rbd": [WORLD, handle_t],
    "bg_remove_col": [WORLD, handle_t],
    "bg_step": [WORLD, C.c_double, C.c_uint32],
    "bg_fixed_step": [WORLD, C.c_double],
    "bg_synchronize": [WORLD],
    "bg_rbd_count": [WORLD, u32p],
    "bg_get_rbd": [WORLD, handle_t, C.POINTER(RigidBodyC)],
    "bg_set_rbd": [WORLD, handle_t, C.POINTER(RigidBodyC)],
    "bg_get_col": [WORLD, handle_t, C.POINTER(ColliderC)],
    "bg_get_joint": [WORLD, handle_t, C.POINTER(FixedJointC)],
    "bg_get_spring": [WORLD, handle_t, C.POINTER(SpringC)],
    "bg_get_bodies": [WORLD, f32p, f32p, f32p, f32p, f32p, C.POINTER(handle_t)],
    "bg_get_collider_positions": [WORLD, f32p, u32p],
    "bg_contact_count": [WORLD, u64p],
    "bg_last_error": [WORLD, C.c_char_p, C.c_size_t],
    "bg_set_broadphase": [WORLD, C.c_float, C.c_float, C.c_uint32],
    "bg_step_path": [WORLD, u32p],
    "bg_set_profiling": [WORLD, C.c_int],
    "bg_set_deferred_checks": [WORLD, C.c_int],
    "bg_stage_stats": [WORLD, C.c_uint32, u64p, C.POINTER(C.c_double)],
    "bg_batch_run": [WORLD, C.c_double, C.c_uint32, C.c_uint32],
    "bg_batch_create": [C.POINTER(WORLD), C.c_uint32, C.POINTER(WORLD)],
    "bg_batch_writeback": [WORLD, C.POINTER(WORLD), C.c_uint32],
    "bg_batch_size": [WORLD, u32p],
    "bg_batch_set_persistent": [WORLD, C.c_int],
    "bg_batch_persistent": [WORLD, C.POINTER(C.c_int)],
    "bg_autotune": [WORLD, C.c_double, C.c_uint32, f32p, f32p, u32p],
    "bg_set_event_capacity": [WORLD, C.c_uint32],
    "bg_drain_events": [WORLD, C.POINTER(CollisionEventC), C.c_uint32, u32p, u64p],
    "bg_debug_data": [WORLD, C.POINTER(Affine2C), C.POINTER(Affine2C), f32p, f32p, f32p, u32p],
}

ROLLOUT = C.c_void_p
SIGNATURES.update({
    "bg_rollout_create": [C.POINTER(WORLD), C.c_uint32, f32p, C.c_uint32, C.POINTER(ROLLOUT)],
    "bg_rollout_destroy": [ROLLOUT],
    "bg_rollout_sizes": [ROLLOUT, u32p, u32p, u32p],
    "bg_rollout_set_params": [ROLLOUT, f32p, f32p],
    "bg_rollout_forward": [ROLLOUT, C.c_double, C.c_uint32, f32p],
    "bg_rollout_backward": [ROLLOUT, f32p, f32p, f32p],
    "bg_rollout_final_positions": [ROLLOUT, f32p],
    "bg_rollout_nonfinite": [ROLLOUT, C.POINTER(C.c_uint8)],
})

# SpatialHash query (device pointers, passed as integers)
SIGNATURES.update({
    "bg_spatial_hash_query": [C.c_float, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                              C.c_uint32, u64p],
    "bg_spatial_hash_query_noalloc": [C.c_float, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_void_p, C.c_uint32, u64p],
    "bg_spatial_hash_query_host": [C.c_float, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_void_p, C.c_uint32, u64p],
})
# entry points bound only when present in the loaded library
EXTRA_SIGNATURES: dict[str, list] = {}
# device-buffer test hooks (internal)
EXTRA_SIGNATURES.update({
    "bgx_device_alloc": [C.c_uint64, C.POINTER(C.c_void_p)],
    "bgx_device_free": [C.c_void_p],
    "bgx_copy_to_device": [C.c_void_p, C.c_void_p, C.c_uint64],
    "bgx_copy_to_host": [C.c_void_p, C.c_void_p, C.c_uint64],
})


_lib: C.CDLL | None = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libblobs_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m blobs_amd.build` (hipcc, gfx950). "
            "There is no CPU fallback for the GPU step.")
    lib = C.CDLL(str(p))
    for name, args in {**SIGNATURES, **EXTRA_SIGNATURES}.items():
        fn = getattr(lib, name, None)
        if fn is None:
            if name in SIGNATURES:
                raise RuntimeError(f"{p} does not export {name}")
            continue
        fn.argtypes = args
        fn.restype = C.c_int
    if path is None:
        _lib = lib
    return lib


class BlobsError(RuntimeError):
    """A non-OK bg_status.  `code` is the status; PANIC mirrors a reference panic."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def check(code: int, world=None) -> None:
    if code == BG_OK:
        return
    buf = C.create_string_buffer(512)
    load().bg_last_error(world, buf, 512)
    raise BlobsError(code, buf.value.decode(errors="replace"))


def fptr(a):
    """float32 numpy array -> POINTER(c_float) (None passes through)."""
    if a is None:
        return None
    return a.ctypes.data_as(f32p)


def u8ptr(a):
    return None if a is None else a.ctypes.data_as(u8p)


def u32ptr(a):
    return None if a is None else a.ctypes.data_as(u32p)


def hptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(handle_t))
