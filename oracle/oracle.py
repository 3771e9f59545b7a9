"""ctypes wrapper of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY -- the checker for the HIP implementation.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; nothing in blobs_amd/ does.

`OracleWorld(scene, np.float32)` is the parity oracle (the reference step in
f32, blobs/src/physics.rs:78-477); `OracleWorld(scene, np.float64)` is its
FP64 twin, the error oracle.  Masses and joint rest data are computed by
the oracle's own restatement of the construction path
(rigid_body.rs:96-128 with the double push of collider.rs:179-181 +
physics.rs:144; physics.rs:184-239), not taken from the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
# libm="cr": the f32 restatement with sinf/cosf/atan2f evaluated in double and
# rounded (the platform libm's last bit is not part of the reference's semantics)
LIBS = {"glibc": LIB, "cr": HERE / "liboracle_cr.so"}

_libs = {}


def build(force: bool = False) -> Path:
    """Compile liboracle.so (make; `force` recompiles even when it is up to date)."""
    subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []), check=True)
    return LIB


def lib(libm: str = "glibc"):
    _lib = _libs.get(libm)
    if _lib is None:
        if not LIBS[libm].exists():
            build()
        _lib = C.CDLL(str(LIBS[libm]))
        _libs[libm] = _lib
        for sfx, W in (("f32", World32), ("f64", World64)):
            getattr(_lib, f"oracle_step_{sfx}").argtypes = [C.POINTER(W), C.c_double, C.c_int32]
            getattr(_lib, f"oracle_fixed_step_{sfx}").argtypes = [C.POINTER(W), C.c_double,
                                                                   C.POINTER(C.c_double)]
        _lib.oracle_contact_pairs_f32.argtypes = [C.POINTER(World32), C.POINTER(C.c_int32), C.c_int64]
        _lib.oracle_contact_pairs_f32.restype = C.c_int64
        _lib.oracle_contact_pairs_f64.argtypes = [C.POINTER(World64), C.POINTER(C.c_int32), C.c_int64]
        _lib.oracle_contact_pairs_f64.restype = C.c_int64
        _lib.oracle_bruteforce_rows_f32.argtypes = [C.POINTER(World32), C.c_int32, C.c_int32]
        _lib.oracle_bruteforce_rows_f32.restype = C.c_int64
        f = C.POINTER(C.c_float)
        i = C.POINTER(C.c_int32)
        u8 = C.POINTER(C.c_uint8)
        _lib.oracle_mass_setup_f32.argtypes = [C.c_int32, i, i, f, f, u8, u8, f, f, f, f]
        _lib.oracle_joint_setup_f32.argtypes = [f, f, f, f, C.c_float, C.c_float, f, f]
        _lib.oracle_body_transform_mul_f32.argtypes = [C.c_float, f, f, f, f, f]
    return _lib


def _fields(R):
    P = C.POINTER
    return [
        ("n_bodies", C.c_int32), ("n_colliders", C.c_int32), ("n_springs", C.c_int32), ("n_joints", C.c_int32),
        ("n_constraints", C.c_int32),
        ("pos", P(R)), ("pos_old", P(R)), ("vel", P(R)), ("acc", P(R)), ("vreq", P(R)),
        ("rot", P(R)), ("omega", P(R)), ("torque", P(R)), ("inertia", P(R)), ("mass", P(R)),
        ("gravity_mod", P(R)),
        ("has_vreq", P(C.c_uint8)), ("body_type", P(C.c_uint8)),
        ("body_col_start", P(C.c_int32)), ("body_col", P(C.c_int32)),
        ("parent", P(C.c_int32)), ("off_t", P(R)), ("abs_t", P(R)), ("off_m", P(R)), ("radius", P(R)),
        ("memberships", P(C.c_uint32)), ("filter", P(C.c_uint32)), ("sensor", P(C.c_uint8)),
        ("spring_a", P(C.c_int32)), ("spring_b", P(C.c_int32)), ("spring_rest", P(R)), ("spring_k", P(R)),
        ("spring_c", P(R)),
        ("joint_a", P(C.c_int32)), ("joint_b", P(C.c_int32)), ("joint_anchor_a", P(R)),
        ("joint_anchor_b", P(R)), ("joint_dist", P(R)), ("joint_target", P(R)),
        ("con_pos", P(R)), ("con_radius", P(R)),
        ("gravity", R * 2), ("substeps", C.c_int32), ("joint_iterations", C.c_int32),
        ("collisions_enabled", C.c_int32), ("old_dt", R), ("time", C.c_double),
        ("contact_count", C.c_uint64), ("error_flags", C.c_uint32), ("push_outs", C.c_uint32),
        ("event_cap", C.c_int64), ("n_events", C.c_int64), ("event_cols", P(C.c_int32)),
        ("event_vels", P(R)), ("use_grid", C.c_int32),
    ]


class World32(C.Structure):
    _fields_ = _fields(C.c_float)


class World64(C.Structure):
    _fields_ = _fields(C.c_double)


ERR_NAN_NORMAL = 1 << 0
ERR_JOINT_MASS = 1 << 1
ERR_JOINT_ROT = 1 << 2
INFO_PUSH_OUT = 1 << 8


def mass_setup(scene):
    """Oracle restatement of the mass/inertia setup, f32."""
    L = lib()
    start, order = scene.body_col_csr()
    n = scene.n_bodies
    mass = np.zeros(n, np.float32)
    inertia = np.zeros(n, np.float32)
    com = np.zeros(2 * n, np.float32)
    r = np.ascontiguousarray(scene.col_r, np.float32)
    mo = np.ascontiguousarray(scene.col_mass_override, np.float32)
    hmo = np.ascontiguousarray(scene.col_has_mo, np.uint8)
    sen = np.ascontiguousarray(scene.col_sensor, np.uint8)
    off = np.ascontiguousarray(scene.col_off, np.float32).reshape(-1)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    i = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    u8 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint8))
    L.oracle_mass_setup_f32(n, i(start), i(order), f(r), f(mo), u8(hmo), u8(sen), f(off), f(mass), f(inertia),
                            f(com))
    return mass, inertia, com.reshape(-1, 2)


def joint_setup(scene, libm="glibc"):
    """distance / target_angle of create_fixed_joint for every joint (f32)."""
    L = lib(libm)
    nj = scene.joint_a.shape[0]
    dist = np.zeros(nj, np.float32)
    targ = np.zeros(nj, np.float32)
    pa = np.zeros(2, np.float32)
    pb = np.zeros(2, np.float32)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    for k in range(nj):
        a, b = int(scene.joint_a[k]), int(scene.joint_b[k])
        pa[:] = scene.pos[a]
        pb[:] = scene.pos[b]
        aa = np.ascontiguousarray(scene.joint_anchor_a[k], np.float32)
        ab = np.ascontiguousarray(scene.joint_anchor_b[k], np.float32)
        d = C.c_float()
        t = C.c_float()
        L.oracle_joint_setup_f32(f(pa), f(pb), f(aa), f(ab), float(scene.rotation[a]), float(scene.rotation[b]),
                                 C.byref(d), C.byref(t))
        dist[k] = d.value
        targ[k] = t.value
    return dist, targ


class OracleWorld:
    """A scene stepped by the CPU restatement (float32 or float64)."""

    def __init__(self, scene, dtype=np.float32, use_grid=True, event_cap=0, substeps=8, joint_iterations=4,
                 collisions_enabled=True, libm="glibc"):
        self.dtype = np.dtype(dtype)
        self._L = lib(libm)
        R = self.dtype
        self._W = World32 if R == np.float32 else World64
        self._sfx = "f32" if R == np.float32 else "f64"
        n, c = scene.n_bodies, scene.n_colliders
        mass, inertia, _ = mass_setup(scene)
        jd, jt = joint_setup(scene, libm)
        start, order = scene.body_col_csr()
        a = {}
        a["pos"] = np.ascontiguousarray(scene.pos, R).reshape(-1).copy()
        a["pos_old"] = a["pos"].copy()
        a["vel"] = np.zeros(2 * n, R)
        a["acc"] = np.zeros(2 * n, R)
        a["vreq"] = np.ascontiguousarray(scene.vreq, R).reshape(-1).copy()
        a["rot"] = np.ascontiguousarray(scene.rotation, R).copy()
        a["omega"] = np.zeros(n, R)
        a["torque"] = np.zeros(n, R)
        a["inertia"] = inertia.astype(R)
        a["mass"] = mass.astype(R)
        a["gravity_mod"] = np.ascontiguousarray(scene.gravity_mod, R).copy()
        a["has_vreq"] = np.ascontiguousarray(scene.has_vreq, np.uint8).copy()
        a["body_type"] = np.ascontiguousarray(scene.body_type, np.uint8).copy()
        a["body_col_start"] = start
        a["body_col"] = order
        a["parent"] = np.ascontiguousarray(scene.col_parent, np.int32)
        a["off_t"] = np.ascontiguousarray(scene.col_off, R).reshape(-1).copy()
        a["abs_t"] = np.ascontiguousarray(scene.col_abs, R).reshape(-1).copy()
        om = np.zeros((c, 4), R)
        om[:, 0] = 1
        om[:, 3] = 1
        a["off_m"] = om.reshape(-1)
        a["radius"] = np.ascontiguousarray(scene.col_r, R).copy()
        a["memberships"] = np.ascontiguousarray(scene.col_mem, np.uint32)
        a["filter"] = np.ascontiguousarray(scene.col_filt, np.uint32)
        a["sensor"] = np.ascontiguousarray(scene.col_sensor, np.uint8)
        a["spring_a"] = np.ascontiguousarray(scene.spring_a, np.int32)
        a["spring_b"] = np.ascontiguousarray(scene.spring_b, np.int32)
        a["spring_rest"] = np.ascontiguousarray(scene.spring_rest, R)
        a["spring_k"] = np.ascontiguousarray(scene.spring_k, R)
        a["spring_c"] = np.ascontiguousarray(scene.spring_c, R)
        a["joint_a"] = np.ascontiguousarray(scene.joint_a, np.int32)
        a["joint_b"] = np.ascontiguousarray(scene.joint_b, np.int32)
        a["joint_anchor_a"] = np.ascontiguousarray(scene.joint_anchor_a, R).reshape(-1)
        a["joint_anchor_b"] = np.ascontiguousarray(scene.joint_anchor_b, R).reshape(-1)
        a["joint_dist"] = jd.astype(R)
        a["joint_target"] = jt.astype(R)
        a["con_pos"] = np.ascontiguousarray(scene.con_pos, R).reshape(-1)
        a["con_radius"] = np.ascontiguousarray(scene.con_r, R)
        if event_cap:
            a["event_cols"] = np.zeros(2 * event_cap, np.int32)
            a["event_vels"] = np.zeros(4 * event_cap, R)
        self.a = a
        w = self._W()
        w.n_bodies, w.n_colliders = n, c
        w.n_springs, w.n_joints = scene.spring_a.shape[0], scene.joint_a.shape[0]
        w.n_constraints = scene.con_r.shape[0]
        for name, arr in a.items():
            ctype = dict(self._W._fields_)[name]._type_
            setattr(w, name, arr.ctypes.data_as(C.POINTER(ctype)))
        w.gravity[0], w.gravity[1] = scene.gravity
        w.substeps = substeps
        w.joint_iterations = joint_iterations
        w.collisions_enabled = int(collisions_enabled)
        w.old_dt = 1.0
        w.time = 0.0
        w.event_cap = event_cap
        w.use_grid = int(use_grid)
        self.w = w

    def step(self, delta, n=1):
        getattr(self._L, f"oracle_step_{self._sfx}")(C.byref(self.w), float(delta), int(n))

    def fixed_step(self, frame_time, accumulator: C.c_double):
        getattr(self._L, f"oracle_fixed_step_{self._sfx}")(C.byref(self.w), float(frame_time), C.byref(accumulator))

    def contact_pairs(self, cap=1 << 22):
        """Run one contact pass (mutating positions) and return its (i, j) pairs."""
        out = np.zeros(2 * cap, np.int32)
        fn = getattr(self._L, f"oracle_contact_pairs_{self._sfx}")
        n = fn(C.byref(self.w), out.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        return out[: 2 * min(n, cap)].reshape(-1, 2)

    def bruteforce_rows(self, i0, i1):
        """Run rows [i0, i1) of the reference O(C^2) contact loop; returns pair tests."""
        return self._L.oracle_bruteforce_rows_f32(C.byref(self.w), int(i0), int(i1))

    @property
    def pos(self):
        return self.a["pos"].reshape(-1, 2)

    @property
    def pos_old(self):
        return self.a["pos_old"].reshape(-1, 2)

    @property
    def vel(self):
        return self.a["vel"].reshape(-1, 2)

    @property
    def rot(self):
        return self.a["rot"]

    @property
    def abs_t(self):
        return self.a["abs_t"].reshape(-1, 2)

    @property
    def events(self):
        n = min(self.w.n_events, self.w.event_cap)
        return self.a["event_cols"][: 2 * n].reshape(-1, 2), self.a["event_vels"][: 4 * n].reshape(-1, 4)

    @property
    def contact_count(self):
        return self.w.contact_count

    @property
    def error_flags(self):
        return self.w.error_flags
