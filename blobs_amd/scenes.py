"""Synthetic scenes for the benchmark and the parity tests.

A `Scene` holds builder-level inputs in arena order (what a user passes to
RigidBodyBuilder / ColliderBuilder / create_fixed_joint / springs.insert),
so the GPU world and the oracle are both built from the same arrays through
the reference's own construction recipe (demo/src/simulation.rs:87-146):
`RigidBodyBuilder.position(p).velocity_request(v)`, then a collider whose
absolute_transform is Affine2(rot, p) * offset (the demo sets
Affine2::from_translation(position), simulation.rs:107), inserted with
`insert_collider_with_parent`.

Generators (libblobs_scenes.so) use xorshift128 seeded like the reference's
`XorShiftRng::from_seed([k; 16])` (blobs/benches/spatial_hash.rs:9).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

_sg = None


def _scenes_lib():
    global _sg
    if _sg is None:
        if not L.SCENES_PATH.exists():
            raise RuntimeError(f"{L.SCENES_PATH} missing: run `python -m blobs_amd.build`")
        _sg = C.CDLL(str(L.SCENES_PATH))
        u32a = C.POINTER(C.c_uint32)
        f32 = C.POINTER(C.c_float)
        i32 = C.POINTER(C.c_int32)
        _sg.sg_seed.argtypes = [u32a, C.c_uint32]
        _sg.sg_uniform.argtypes = [u32a, C.c_int64, C.c_double, C.c_double, C.c_int, f32]
        _sg.sg_normal.argtypes = [u32a, C.c_int64, C.c_double, f32]
        _sg.sg_randint.argtypes = [u32a, C.c_int64, C.c_int64, C.c_int64, i32]
        _sg.sg_spheres.argtypes = [u32a, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_int64, f32, f32, f32]
        _sg.sg_spheres.restype = C.c_int64
        _sg.sg_multi.argtypes = [u32a, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, f32, f32,
                                 i32, f32, f32, C.POINTER(C.c_int64)]
        _sg.sg_multi.restype = C.c_int64
        _sg.sg_pairs.argtypes = [u32a, C.c_int64, f32, C.c_int64, C.c_double, i32, i32, f32]
        _sg.sg_pairs.restype = C.c_int64
    return _sg


class Rng:
    """xorshift128 stream (state shared with the C generator)."""

    def __init__(self, seed_byte: int):
        self.st = (C.c_uint32 * 4)()
        _scenes_lib().sg_seed(self.st, seed_byte)

    def uniform(self, n, a, b, log=False):
        out = np.zeros(n, np.float32)
        _scenes_lib().sg_uniform(self.st, n, a, b, int(log), out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def normal(self, n, sigma):
        out = np.zeros(n, np.float32)
        _scenes_lib().sg_normal(self.st, n, sigma, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def randint(self, n, lo, hi):
        out = np.zeros(n, np.int32)
        _scenes_lib().sg_randint(self.st, n, lo, hi, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@dataclass
class Scene:
    # bodies
    pos: np.ndarray
    vreq: np.ndarray
    has_vreq: np.ndarray
    gravity_mod: np.ndarray
    body_type: np.ndarray
    rotation: np.ndarray
    # colliders (grouped by parent in insertion order)
    col_parent: np.ndarray
    col_off: np.ndarray
    col_abs: np.ndarray
    col_r: np.ndarray
    col_mass_override: np.ndarray
    col_has_mo: np.ndarray
    col_sensor: np.ndarray
    col_mem: np.ndarray
    col_filt: np.ndarray
    # springs / joints / constraints
    spring_a: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    spring_b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    spring_rest: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    spring_k: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    spring_c: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    joint_a: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    joint_b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    joint_anchor_a: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    joint_anchor_b: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    con_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    con_r: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    gravity: tuple = (0.0, -9.81)
    dt: float = 1.0 / 60.0
    steps: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def n_bodies(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_colliders(self) -> int:
        return int(self.col_r.shape[0])

    def build_physics(self, physics=None):
        """Insert everything into a GPU `Physics` world through the bulk C ABI
        (n sequential insert_rbd / insert_collider_with_parent calls)."""
        from .physics import Physics
        p = physics or Physics(self.gravity, False)
        bh = p.insert_bodies(self.pos, self.vreq, self.gravity_mod, self.body_type, self.rotation)
        has = self.has_vreq
        if has is not None and not np.all(has):
            raise ValueError("bulk insert assumes every body carries a velocity request")
        p.insert_colliders(bh[self.col_parent], self.col_r, self.col_off, self.col_abs, self.col_mass_override,
                           self.col_has_mo, self.col_sensor, self.col_mem, self.col_filt)
        for k in range(self.spring_a.shape[0]):
            from .physics import Spring
            p.springs.insert(Spring(int(bh[self.spring_a[k]]), int(bh[self.spring_b[k]]), float(self.spring_rest[k]),
                                    float(self.spring_k[k]), float(self.spring_c[k])))
        for k in range(self.joint_a.shape[0]):
            p.create_fixed_joint(int(bh[self.joint_a[k]]), int(bh[self.joint_b[k]]), self.joint_anchor_a[k],
                                 self.joint_anchor_b[k])
        for k in range(self.con_r.shape[0]):
            from .physics import Constraint
            p.constraints.push(Constraint(tuple(self.con_pos[k]), float(self.con_r[k])))
        return p, bh

    def body_col_csr(self):
        """Colliders of each body in insertion order (CSR over body index)."""
        n = self.n_bodies
        order = np.argsort(self.col_parent, kind="stable").astype(np.int32)
        counts = np.bincount(self.col_parent, minlength=n)
        start = np.zeros(n + 1, np.int32)
        start[1:] = np.cumsum(counts)
        return start, order


def _single_sphere_scene(pos, vel, radius, gravity, dt, steps, meta, static_mask=None, mass_override=None):
    n = pos.shape[0]
    bt = np.zeros(n, np.uint8)
    if static_mask is not None:
        bt[static_mask] = 1
    has_mo = np.zeros(n, np.uint8)
    mo = np.zeros(n, np.float32)
    if mass_override is not None:
        has_mo[:] = ~np.isnan(mass_override)
        mo[:] = np.nan_to_num(mass_override)
    return Scene(
        pos=pos.astype(np.float32), vreq=vel.astype(np.float32), has_vreq=np.ones(n, np.uint8),
        gravity_mod=np.ones(n, np.float32), body_type=bt, rotation=np.zeros(n, np.float32),
        col_parent=np.arange(n, dtype=np.int32), col_off=np.zeros((n, 2), np.float32),
        col_abs=pos.astype(np.float32).copy(), col_r=radius.astype(np.float32), col_mass_override=mo,
        col_has_mo=has_mo, col_sensor=np.zeros(n, np.uint8), col_mem=np.full(n, 0xFFFFFFFF, np.uint32),
        col_filt=np.full(n, 0xFFFFFFFF, np.uint32), gravity=gravity, dt=dt, steps=steps, meta=meta)


def spheres(n, rmin, rmax, box, vel_sigma, seed, reject=True, max_tries=200):
    """n single spheres, radius U[rmin,rmax], centre U over box=(x0,y0,x1,y1)."""
    st = Rng(seed)
    pos = np.zeros((n, 2), np.float32)
    vel = np.zeros((n, 2), np.float32)
    rad = np.zeros(n, np.float32)
    forced = _scenes_lib().sg_spheres(st.st, n, rmin, rmax, *box, vel_sigma, int(reject), max_tries, _f(pos),
                                      _f(vel), _f(rad))
    return pos, vel, rad, int(forced), st


def box_side_for_packing(n, rmin, rmax, phi):
    """Side L with n * pi * E[r^2] / L^2 = phi for r ~ U[rmin, rmax]."""
    er2 = (rmax ** 3 - rmin ** 3) / (3.0 * (rmax - rmin)) if rmax > rmin else rmin * rmin
    return math.sqrt(n * math.pi * er2 / phi)


def plain_small_scene(n=1024, steps=600, walls=True, seed=1):
    """1,024 spheres r~U[0.1,0.3] in a 64x64 box, v~N(0,1), g=(0,-9.81),
    dt=1/60.  Walls: static spheres r=0.5 one unit apart just outside the box,
    mass_override 1e30 so contacts leave them in place (blobs has no plane
    collider; static bodies are still moved by contacts, physics.rs:296-299)."""
    pos, vel, rad, forced, _ = spheres(n, 0.1, 0.3, (0.0, 0.0, 64.0, 64.0), 1.0, seed)
    meta = {"config": "plain_small", "forced_overlaps": forced, "walls": walls}
    if walls:
        ws = []
        for k in range(67):
            t = -1.0 + k
            ws += [(t, -1.0), (t, 65.0)]
        for k in range(1, 66):
            t = -1.0 + k
            ws += [(-1.0, t), (65.0, t)]
        wp = np.array(ws, np.float32)
        m = wp.shape[0]
        pos = np.concatenate([pos, wp])
        vel = np.concatenate([vel, np.zeros((m, 2), np.float32)])
        rad = np.concatenate([rad, np.full(m, 0.5, np.float32)])
        static = np.zeros(n + m, bool)
        static[n:] = True
        mo = np.full(n + m, np.nan, np.float32)
        mo[n:] = 1e30
        meta["n_walls"] = m
        return _single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1.0 / 60.0, steps, meta, static, mo)
    return _single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1.0 / 60.0, steps, meta)


def plain_large_scene(n=1 << 20, steps=100, seed=2, phi=0.4):
    """n spheres r~U[0.5,1.0], square box at packing fraction phi, centres U
    with overlap rejection, v~N(0,2), gravity on, dt=1/120, no walls."""
    side = box_side_for_packing(n, 0.5, 1.0, phi)
    pos, vel, rad, forced, _ = spheres(n, 0.5, 1.0, (0.0, 0.0, side, side), 2.0, seed)
    meta = {"config": "plain_large", "box_side": side, "forced_overlaps": forced, "walls": False}
    return _single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1.0 / 120.0, steps, meta)


def multi_collider_scene(n=65536, n_springs=32768, n_joints=16384, steps=300, seed=3, phi=0.2):
    """n bodies with U{1..8} sphere colliders (offsets uniform in a disc of
    radius 2, collider r~U[0.2,0.5]), n_springs springs and n_joints fixed
    joints between random bodies within distance 6 (rest = initial distance,
    stiffness logU[10,1e4], damping 0; joints create_fixed_joint(a,b,0,0)),
    dt=1/60, gravity on.  Square box centred on the origin, side chosen for
    collider packing fraction phi (centring keeps f32 position resolution
    fine: joint rotations reach positions through collider offsets)."""
    rng = Rng(seed)
    cap = n * 8
    bpos = np.zeros((n, 2), np.float32)
    bvel = np.zeros((n, 2), np.float32)
    ncol = np.zeros(n, np.int32)
    off = np.zeros((cap, 2), np.float32)
    cr = np.zeros(cap, np.float32)
    er2 = (0.5 ** 3 - 0.2 ** 3) / (3 * 0.3)
    side = math.sqrt(n * 4.5 * math.pi * er2 / phi)
    forced = C.c_int64()
    half = side / 2.0
    nc = _scenes_lib().sg_multi(rng.st, n, 1, 8, 2.0, 0.2, 0.5, -half, -half, half, half, 1.0, 100, _f(bpos),
                                _f(bvel), _i(ncol), _f(off), _f(cr), C.byref(forced))
    off = off[:nc].copy()
    cr = cr[:nc].copy()
    parent = np.repeat(np.arange(n, dtype=np.int32), ncol)
    cabs = (bpos[parent] + off).astype(np.float32)  # Affine2(0, p) * offset = p + offset (rot 0)
    pairs_a = np.zeros(n_springs + n_joints, np.int32)
    pairs_b = np.zeros(n_springs + n_joints, np.int32)
    rest = np.zeros(n_springs + n_joints, np.float32)
    made = _scenes_lib().sg_pairs(rng.st, n, _f(bpos), n_springs + n_joints, 6.0, _i(pairs_a), _i(pairs_b),
                                  _f(rest))
    assert made == n_springs + n_joints
    k = rng.uniform(n_springs, 10.0, 1e4, log=True)
    sc = Scene(
        pos=bpos, vreq=bvel, has_vreq=np.ones(n, np.uint8), gravity_mod=np.ones(n, np.float32),
        body_type=np.zeros(n, np.uint8), rotation=np.zeros(n, np.float32), col_parent=parent, col_off=off,
        col_abs=cabs, col_r=cr, col_mass_override=np.zeros(nc, np.float32), col_has_mo=np.zeros(nc, np.uint8),
        col_sensor=np.zeros(nc, np.uint8), col_mem=np.full(nc, 0xFFFFFFFF, np.uint32),
        col_filt=np.full(nc, 0xFFFFFFFF, np.uint32),
        spring_a=pairs_a[:n_springs].copy(), spring_b=pairs_b[:n_springs].copy(),
        spring_rest=rest[:n_springs].copy(), spring_k=k, spring_c=np.zeros(n_springs, np.float32),
        joint_a=pairs_a[n_springs:].copy(), joint_b=pairs_b[n_springs:].copy(),
        joint_anchor_a=np.zeros((n_joints, 2), np.float32), joint_anchor_b=np.zeros((n_joints, 2), np.float32),
        gravity=(0.0, -9.81), dt=1.0 / 60.0, steps=steps,
        meta={"config": "multi_collider", "box_side": side, "forced_overlaps": int(forced.value),
              "spring_damping": 0.0})
    return sc


def batch_world_sizes(n_worlds=4096, lo=64, hi=16384, seed=4):
    """Body counts log-uniform in [lo, hi] (rounded)."""
    rng = Rng(seed)
    return np.rint(rng.uniform(n_worlds, lo, hi, log=True)).astype(np.int64), rng


def uneven_batch_scenes(n_worlds=4096, lo=64, hi=16384, seed=4, size_cap=None):
    """The batched scene: n_worlds single-sphere worlds, N_w log-uniform in
    [lo, hi], packing U[0.2, 0.5] per world, all drawn from one seeded
    stream (sizes, then packings, then each world's bodies)."""
    sizes, rng = batch_world_sizes(n_worlds, lo, hi, seed)
    packs = rng.uniform(n_worlds, 0.2, 0.5)
    if size_cap is not None:
        sizes = np.minimum(sizes, size_cap)
    return [uneven_batch_world(int(sizes[k]), float(packs[k]), rng) for k in range(n_worlds)]


def uneven_batch_world(size, packing, seed_rng: Rng, vel_sigma=1.0):
    """One batched world: single spheres r~U[0.1,0.5] at the given packing."""
    side = box_side_for_packing(int(size), 0.1, 0.5, float(packing))
    pos = np.zeros((size, 2), np.float32)
    vel = np.zeros((size, 2), np.float32)
    rad = np.zeros(size, np.float32)
    forced = _scenes_lib().sg_spheres(seed_rng.st, int(size), 0.1, 0.5, 0.0, 0.0, side, side, vel_sigma, 1, 200,
                                      _f(pos), _f(vel), _f(rad))
    return _single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1.0 / 60.0, 1000,
                                {"config": "batched_world", "box_side": side, "forced_overlaps": int(forced)})


def rollout_world(n, n_springs, rng: Rng, packing=0.4, vel_sigma=2.0, kmin=10.0, kmax=1e3, damping=0.0,
                  n_joints=0, anchor=0.0):
    """One differentiable-rollout world: n single spheres r~U[0.5,1.0] at the
    given packing (square box centred on the origin: position Verlet forms
    velocities as p - p_old in f32, so small coordinates keep them
    accurate), velocity requests N(0, vel_sigma),
    n_springs springs between random bodies within distance 6 (rest =
    initial distance, stiffness logU[kmin, kmax], damping 0 or, with
    damping > 0, U[0, damping]), gravity on, dt = 1/120.  n_joints > 0 adds
    fixed joints between random bodies within distance 6, anchors
    U[-anchor, anchor]^2 (distance and target angle from create_fixed_joint).
    The extra draws come after the default ones, so the default world is
    unchanged."""
    side = box_side_for_packing(int(n), 0.5, 1.0, float(packing))
    pos = np.zeros((n, 2), np.float32)
    vel = np.zeros((n, 2), np.float32)
    rad = np.zeros(n, np.float32)
    half = side / 2.0
    forced = _scenes_lib().sg_spheres(rng.st, int(n), 0.5, 1.0, -half, -half, half, half, vel_sigma, 1, 200,
                                      _f(pos), _f(vel), _f(rad))
    sa = np.zeros(n_springs, np.int32)
    sb = np.zeros(n_springs, np.int32)
    rest = np.zeros(n_springs, np.float32)
    made = _scenes_lib().sg_pairs(rng.st, int(n), _f(pos), int(n_springs), 6.0, _i(sa), _i(sb), _f(rest))
    assert made == n_springs
    k = rng.uniform(n_springs, kmin, kmax, log=True)
    sc = _single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1.0 / 120.0, 200,
                              {"config": "differentiable_world", "box_side": side, "forced_overlaps": int(forced),
                               "spring_damping": 0.0})
    sc.spring_a, sc.spring_b, sc.spring_rest = sa, sb, rest
    sc.spring_k, sc.spring_c = k, np.zeros(n_springs, np.float32)
    if damping > 0.0:
        sc.spring_c = rng.uniform(n_springs, 0.0, damping).astype(np.float32)
        sc.meta["spring_damping"] = float(damping)
    if n_joints > 0:
        ja = np.zeros(n_joints, np.int32)
        jb = np.zeros(n_joints, np.int32)
        jr = np.zeros(n_joints, np.float32)
        made = _scenes_lib().sg_pairs(rng.st, int(n), _f(pos), int(n_joints), 6.0, _i(ja), _i(jb), _f(jr))
        assert made == n_joints
        sc.joint_a, sc.joint_b = ja, jb
        sc.joint_anchor_a = rng.uniform(2 * n_joints, -anchor, anchor).reshape(n_joints, 2).astype(np.float32)
        sc.joint_anchor_b = rng.uniform(2 * n_joints, -anchor, anchor).reshape(n_joints, 2).astype(np.float32)
    return sc


def rollout_scenes(n_worlds=256, n=2048, n_springs=1024, seed=5, **kw):
    """The differentiable rollout: n_worlds worlds of n spheres + n_springs
    springs, drawn from one seeded stream, plus one seeded target point per
    world (uniform over the world's box)."""
    rng = Rng(seed)
    worlds = [rollout_world(n, n_springs, rng, **kw) for _ in range(n_worlds)]
    side = np.array([w.meta["box_side"] for w in worlds], np.float32)
    u = rng.uniform(2 * n_worlds, -0.5, 0.5).reshape(n_worlds, 2)
    targets = (u * side[:, None]).astype(np.float32)
    return worlds, targets
