"""Python mirror of the reference `Physics` API over the C ABI.

Names, argument meaning and error behaviour follow blobs/src/physics.rs,
rigid_body.rs, collider.rs, springs.rs and joints.rs, so code written against
the reference reads the same:

    physics = Physics((0.0, -9.81), False)            # Physics::new
    rbd = RigidBodyBuilder().position((1, 2)).velocity_request((0, 1)).build()
    h = physics.insert_rbd(rbd)                       # physics.rs:121
    physics.insert_collider_with_parent(
        ColliderBuilder().radius(0.5).absolute_transform_translation((1, 2)).build(), h)
    physics.step(1 / 60)                              # physics.rs:78

Stepping runs on the GPU (libblobs_gpu.so); where the reference panics,
these calls raise BlobsError with code BG_ERR_PANIC.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

DYNAMIC, STATIC, KINEMATIC_POSITION_BASED, KINEMATIC_VELOCITY_BASED = 0, 1, 2, 3


class RigidBodyType:
    """rigid_body.rs:221-242 (same discriminants)."""
    Dynamic = DYNAMIC
    Static = STATIC
    KinematicPositionBased = KINEMATIC_POSITION_BASED
    KinematicVelocityBased = KINEMATIC_VELOCITY_BASED


@dataclass
class RigidBody:
    """RigidBody fields, rigid_body.rs:42-74."""
    position: tuple = (0.0, 0.0)
    position_old: tuple = (0.0, 0.0)
    center_of_mass: tuple = (0.0, 0.0)
    calculated_mass: float = 1.0
    gravity_mod: float = 1.0
    rotation: float = 0.0
    angular_velocity: float = 0.0
    torque: float = 0.0
    inertia: float = 1.0
    scale: tuple = (1.0, 1.0)
    acceleration: tuple = (0.0, 0.0)
    velocity_request: tuple | None = None
    calculated_velocity: tuple = (0.0, 0.0)
    body_type: int = DYNAMIC
    user_data: int = 0

    def to_c(self) -> L.RigidBodyC:
        c = L.RigidBodyC()
        c.position[:] = self.position
        c.position_old[:] = self.position_old
        c.center_of_mass[:] = self.center_of_mass
        c.calculated_mass = self.calculated_mass
        c.gravity_mod = self.gravity_mod
        c.rotation = self.rotation
        c.angular_velocity = self.angular_velocity
        c.torque = self.torque
        c.inertia = self.inertia
        c.scale[:] = self.scale
        c.acceleration[:] = self.acceleration
        if self.velocity_request is not None:
            c.velocity_request[:] = self.velocity_request
            c.has_velocity_request = 1
        c.calculated_velocity[:] = self.calculated_velocity
        c.body_type = self.body_type
        c.user_data[0] = self.user_data & 0xFFFFFFFFFFFFFFFF
        c.user_data[1] = (self.user_data >> 64) & 0xFFFFFFFFFFFFFFFF
        return c

    @staticmethod
    def from_c(c: L.RigidBodyC) -> "RigidBody":
        return RigidBody(
            position=tuple(c.position), position_old=tuple(c.position_old),
            center_of_mass=tuple(c.center_of_mass), calculated_mass=c.calculated_mass,
            gravity_mod=c.gravity_mod, rotation=c.rotation, angular_velocity=c.angular_velocity,
            torque=c.torque, inertia=c.inertia, scale=tuple(c.scale), acceleration=tuple(c.acceleration),
            velocity_request=tuple(c.velocity_request) if c.has_velocity_request else None,
            calculated_velocity=tuple(c.calculated_velocity), body_type=c.body_type,
            user_data=c.user_data[0] | (c.user_data[1] << 64))

    # RigidBody methods used around the step (rigid_body.rs:130-213)
    def is_static(self) -> bool:
        return self.body_type == STATIC

    def apply_force(self, force) -> None:
        """rigid_body.rs:155-160: acceleration += force / mass (not static)."""
        if not self.is_static():
            m = np.float32(self.calculated_mass)
            self.acceleration = (float(np.float32(self.acceleration[0]) + np.float32(force[0]) / m),
                                 float(np.float32(self.acceleration[1]) + np.float32(force[1]) / m))

    def _lever(self, world_point):
        # world_point - (position + center_of_mass), f32 per component
        f = np.float32
        return (f(world_point[0]) - (f(self.position[0]) + f(self.center_of_mass[0])),
                f(world_point[1]) - (f(self.position[1]) + f(self.center_of_mass[1])))

    @staticmethod
    def _perp_dot(a, b):
        # glam Vec2::perp_dot: a.x * b.y - a.y * b.x
        f = np.float32
        return f(a[0]) * f(b[1]) - f(a[1]) * f(b[0])

    def apply_impulse(self, impulse) -> None:
        """rigid_body.rs:130-135: velocity_request = calculated_velocity + impulse / mass."""
        if not self.is_static():
            m = np.float32(self.calculated_mass)
            self.add_velocity((np.float32(impulse[0]) / m, np.float32(impulse[1]) / m))

    def apply_impulse_at_point(self, impulse, world_point) -> None:
        """rigid_body.rs:137-148."""
        if not self.is_static():
            self.apply_impulse(impulse)
            ang = self._perp_dot(self._lever(world_point), impulse)
            self.angular_velocity = float(np.float32(self.angular_velocity) + ang / np.float32(self.inertia))

    def add_velocity(self, velocity) -> None:
        """rigid_body.rs:150-152."""
        v = self.get_velocity()
        self.set_velocity((np.float32(v[0]) + np.float32(velocity[0]), np.float32(v[1]) + np.float32(velocity[1])))

    def apply_force_at_point(self, force, world_point) -> None:
        """rigid_body.rs:162-171."""
        if not self.is_static():
            self.apply_force(force)
            self.torque = float(np.float32(self.torque) + self._perp_dot(self._lever(world_point), force))

    def apply_torque_at_point(self, force, world_point) -> None:
        """rigid_body.rs:173-179: torque only."""
        if not self.is_static():
            self.torque = float(np.float32(self.torque) + self._perp_dot(self._lever(world_point), force))

    def set_velocity(self, v) -> None:
        """rigid_body.rs:182-184."""
        self.velocity_request = (float(v[0]), float(v[1]))

    def get_velocity(self):
        return self.calculated_velocity


class RigidBodyBuilder:
    """rigid_body.rs:287-401."""

    def __init__(self):
        self._b = RigidBody()

    def position(self, p):
        self._b.position = (float(p[0]), float(p[1]))
        self._b.position_old = self._b.position
        return self

    def gravity_mod(self, g):
        self._b.gravity_mod = float(g)
        return self

    def rotation(self, r):
        self._b.rotation = float(r)
        return self

    def scale(self, s):
        self._b.scale = (float(s[0]), float(s[1]))
        return self

    def acceleration(self, a):
        self._b.acceleration = (float(a[0]), float(a[1]))
        return self

    def velocity_request(self, v):
        self._b.velocity_request = (float(v[0]), float(v[1]))
        return self

    def calculated_velocity(self, v):
        self._b.calculated_velocity = (float(v[0]), float(v[1]))
        return self

    def user_data(self, u):
        self._b.user_data = int(u)
        return self

    def body_type(self, t):
        self._b.body_type = int(t)
        return self

    def build(self) -> RigidBody:
        b = self._b
        return RigidBody(**{**b.__dict__, "calculated_mass": 1.0, "inertia": 1.0, "angular_velocity": 0.0,
                            "torque": 0.0, "center_of_mass": (0.0, 0.0)})


@dataclass
class InteractionGroups:
    """groups.rs:7-64."""
    memberships: int = 0xFFFFFFFF
    filter: int = 0xFFFFFFFF

    def test(self, rhs: "InteractionGroups") -> bool:
        return (self.memberships & rhs.filter) != 0 and (rhs.memberships & self.filter) != 0


@dataclass
class Collider:
    """Collider fields, collider.rs:4-20 (Affine2 as (mat2 columns, translation))."""
    offset_matrix: tuple = (1.0, 0.0, 0.0, 1.0)
    offset_translation: tuple = (0.0, 0.0)
    absolute_matrix: tuple = (1.0, 0.0, 0.0, 1.0)
    absolute_translation: tuple = (0.0, 0.0)
    user_data: int = 0
    parent: int | None = None
    radius: float = 0.5
    mass_override: float | None = None
    is_sensor: bool = False
    collision_groups: InteractionGroups = field(default_factory=InteractionGroups)

    def to_c(self) -> L.ColliderC:
        c = L.ColliderC()
        c.offset_matrix[:] = self.offset_matrix
        c.offset_translation[:] = self.offset_translation
        c.absolute_matrix[:] = self.absolute_matrix
        c.absolute_translation[:] = self.absolute_translation
        c.user_data[0] = self.user_data & 0xFFFFFFFFFFFFFFFF
        c.user_data[1] = (self.user_data >> 64) & 0xFFFFFFFFFFFFFFFF
        if self.parent is not None:
            c.parent = self.parent
            c.has_parent = 1
        c.radius = self.radius
        if self.mass_override is not None:
            c.mass_override = self.mass_override
            c.has_mass_override = 1
        c.is_sensor = 1 if self.is_sensor else 0
        c.memberships = self.collision_groups.memberships
        c.filter = self.collision_groups.filter
        return c

    @staticmethod
    def from_c(c: L.ColliderC) -> "Collider":
        return Collider(
            offset_matrix=tuple(c.offset_matrix), offset_translation=tuple(c.offset_translation),
            absolute_matrix=tuple(c.absolute_matrix), absolute_translation=tuple(c.absolute_translation),
            user_data=c.user_data[0] | (c.user_data[1] << 64), parent=c.parent if c.has_parent else None,
            radius=c.radius, mass_override=c.mass_override if c.has_mass_override else None,
            is_sensor=bool(c.is_sensor), collision_groups=InteractionGroups(c.memberships, c.filter))

    def mass(self) -> float:
        """collider.rs:40-42."""
        return self.mass_override if self.mass_override is not None else float(np.float32(self.radius) * 2)

    def absolute_translation_vec(self):
        return self.absolute_translation


class ColliderBuilder:
    """collider.rs:199-284 (shape omitted: physics reads `radius`)."""

    def __init__(self):
        self._c = Collider()

    def offset_translation(self, t):
        self._c.offset_translation = (float(t[0]), float(t[1]))
        return self

    def offset(self, matrix, translation):
        self._c.offset_matrix = tuple(float(v) for v in matrix)
        self._c.offset_translation = (float(translation[0]), float(translation[1]))
        return self

    def absolute_transform(self, matrix, translation):
        self._c.absolute_matrix = tuple(float(v) for v in matrix)
        self._c.absolute_translation = (float(translation[0]), float(translation[1]))
        return self

    def absolute_transform_translation(self, t):
        """Affine2::from_translation(t) as in demo/src/simulation.rs:107."""
        self._c.absolute_matrix = (1.0, 0.0, 0.0, 1.0)
        self._c.absolute_translation = (float(t[0]), float(t[1]))
        return self

    def mass_override(self, m):
        self._c.mass_override = float(m)
        return self

    def user_data(self, u):
        self._c.user_data = int(u)
        return self

    def parent(self, p):
        self._c.parent = int(p)
        return self

    def radius(self, r):
        self._c.radius = float(r)
        return self

    def sensor(self, s=True):
        self._c.is_sensor = bool(s)
        return self

    def collision_groups(self, g: InteractionGroups):
        self._c.collision_groups = g
        return self

    def build(self) -> Collider:
        return Collider(**self._c.__dict__)


@dataclass
class Spring:
    """springs.rs:16-22."""
    rigid_body_a: int
    rigid_body_b: int
    rest_length: float
    stiffness: float
    damping: float = 0.0


@dataclass
class Constraint:
    """Circle constraint, lib.rs:189-193."""
    position: tuple
    radius: float


class _SpringArena:
    """Stand-in for the pub `springs: Arena<Spring>` (physics.rs:14)."""

    def __init__(self, physics: "Physics"):
        self._p = physics

    def insert(self, s: Spring) -> int:
        c = L.SpringC(s.rigid_body_a, s.rigid_body_b, s.rest_length, s.stiffness, s.damping)
        h = L.handle_t()
        L.check(self._p._lib.bg_insert_spring(self._p._w, C.byref(c), C.byref(h)), self._p._w)
        return h.value


class _ConstraintVec:
    """Stand-in for the pub `constraints: Vec<Constraint>` (physics.rs:15)."""

    def __init__(self, physics: "Physics"):
        self._p = physics

    def push(self, c: Constraint) -> None:
        L.check(self._p._lib.bg_push_constraint(self._p._w, c.position[0], c.position[1], c.radius), self._p._w)

    append = push


class Physics:
    """The reference `Physics` world (physics.rs:3-34) with the step on the GPU."""

    def __init__(self, gravity=(0.0, 0.0), use_spatial_hash: bool = False):
        self._lib = L.load()
        w = C.c_void_p()
        L.check(self._lib.bg_world_create(float(gravity[0]), float(gravity[1]), int(bool(use_spatial_hash)),
                                          C.byref(w)))
        self._w = w
        self.springs = _SpringArena(self)
        self.constraints = _ConstraintVec(self)

    def __del__(self):
        w = getattr(self, "_w", None)
        if w:
            self._lib.bg_world_destroy(w)
            self._w = None

    def _chk(self, code: int) -> None:
        L.check(code, self._w)

    # ---- pub fields (physics.rs:6-9,20,27,30-33)
    def _params(self) -> L.ParamsC:
        p = L.ParamsC()
        self._chk(self._lib.bg_get_params(self._w, C.byref(p)))
        return p

    def _set(self, **kw) -> None:
        p = self._params()
        for k, v in kw.items():
            if k == "gravity":
                p.gravity[:] = (float(v[0]), float(v[1]))
            else:
                setattr(p, k, v)
        self._chk(self._lib.bg_set_params(self._w, C.byref(p)))

    gravity = property(lambda s: tuple(s._params().gravity), lambda s, v: s._set(gravity=v))
    substeps = property(lambda s: s._params().substeps, lambda s, v: s._set(substeps=int(v)))
    joint_iterations = property(lambda s: s._params().joint_iterations,
                                lambda s, v: s._set(joint_iterations=int(v)))
    collisions_enabled = property(lambda s: bool(s._params().collisions_enabled),
                                  lambda s, v: s._set(collisions_enabled=int(bool(v))))
    use_spatial_hash = property(lambda s: bool(s._params().use_spatial_hash),
                                lambda s, v: s._set(use_spatial_hash=int(bool(v))))
    time = property(lambda s: s._params().time, lambda s, v: s._set(time=float(v)))
    accumulator = property(lambda s: s._params().accumulator, lambda s, v: s._set(accumulator=float(v)))
    old_dt = property(lambda s: s._params().old_dt, lambda s, v: s._set(old_dt=float(v)))

    # ---- construction (physics.rs:121-239)
    def insert_rbd(self, rbd: RigidBody) -> int:
        c = rbd.to_c()
        h = L.handle_t()
        self._chk(self._lib.bg_insert_rbd(self._w, C.byref(c), C.byref(h)))
        return h.value

    def insert_collider_with_parent(self, collider: Collider, rbd_handle: int) -> int:
        c = collider.to_c()
        h = L.handle_t()
        self._chk(self._lib.bg_insert_collider_with_parent(self._w, C.byref(c), rbd_handle, C.byref(h)))
        return h.value

    def create_fixed_joint(self, a: int, b: int, anchor_a, anchor_b) -> int:
        aa = (C.c_float * 2)(*anchor_a)
        ab = (C.c_float * 2)(*anchor_b)
        h = L.handle_t()
        self._chk(self._lib.bg_create_fixed_joint(self._w, a, b, aa, ab, C.byref(h)))
        return h.value

    def create_fixed_joint_with_distance(self, a: int, b: int, anchor_a, anchor_b, distance: float) -> int:
        aa = (C.c_float * 2)(*anchor_a)
        ab = (C.c_float * 2)(*anchor_b)
        h = L.handle_t()
        self._chk(self._lib.bg_create_fixed_joint_with_distance(self._w, a, b, aa, ab, float(distance),
                                                                 C.byref(h)))
        return h.value

    def remove_rbd(self, h: int) -> None:
        self._chk(self._lib.bg_remove_rbd(self._w, h))

    def remove_col(self, h: int) -> None:
        self._chk(self._lib.bg_remove_col(self._w, h))

    def reset(self) -> None:
        self._chk(self._lib.bg_world_reset(self._w))

    # ---- stepping (physics.rs:78-99)
    def step(self, delta: float, n_steps: int = 1) -> None:
        self._chk(self._lib.bg_step(self._w, float(delta), int(n_steps)))

    def set_deferred_checks(self, enable: bool) -> None:
        """step() returns once its kernels are queued; a panic surfaces at a
        later step() or synchronize()."""
        self._chk(self._lib.bg_set_deferred_checks(self._w, int(bool(enable))))

    def synchronize(self) -> None:
        self._chk(self._lib.bg_synchronize(self._w))

    def fixed_step(self, frame_time: float) -> None:
        self._chk(self._lib.bg_fixed_step(self._w, float(frame_time)))

    def run_chunks(self, delta: float, chunk_steps: int, n_chunks: int) -> None:
        """n_chunks x chunk_steps steps, each chunk a replayed hipGraph (same
        results as step(delta, chunk_steps * n_chunks))."""
        self._chk(self._lib.bg_batch_run(self._w, float(delta), int(chunk_steps), int(n_chunks)))

    def autotune(self, delta: float, steps_per_trial: int = 2):
        """Pick the fastest broadphase settings over real steps; returns
        (cell_scale, skin, rebuild_interval)."""
        cs, sk, iv = C.c_float(), C.c_float(), C.c_uint32()
        self._chk(self._lib.bg_autotune(self._w, float(delta), int(steps_per_trial), C.byref(cs), C.byref(sk),
                                        C.byref(iv)))
        return cs.value, sk.value, iv.value

    # ---- batches of single-sphere worlds
    @classmethod
    def batch(cls, worlds) -> "Physics":
        """One device world stepping every world of `worlds` (each a
        single-sphere Physics) in one launch per substep; bodies in world
        order.  Results per world are identical to stepping it alone."""
        worlds = list(worlds)
        lib = L.load()
        arr = (C.c_void_p * len(worlds))(*[w._w.value for w in worlds])
        out = C.c_void_p()
        code = lib.bg_batch_create(arr, len(worlds), C.byref(out))
        L.check(code, None)
        self = cls.__new__(cls)
        self._lib = lib
        self._w = out
        self.springs = _SpringArena(self)
        self.constraints = _ConstraintVec(self)
        self._batch_worlds = worlds
        return self

    def batch_size(self) -> int:
        n = C.c_uint32()
        self._chk(self._lib.bg_batch_size(self._w, C.byref(n)))
        return n.value

    def set_persistent(self, enable: bool) -> None:
        """Step this batch with the persistent kernel (one workgroup per
        world, one launch per step call or chunk) when it allows it.  Off
        (the default): the grid path."""
        self._chk(self._lib.bg_batch_set_persistent(self._w, int(bool(enable))))

    def persistent_active(self) -> bool:
        v = C.c_int()
        self._chk(self._lib.bg_batch_persistent(self._w, C.byref(v)))
        return bool(v.value)

    def writeback(self, worlds=None) -> None:
        """Copy a batch's state back into its source worlds."""
        worlds = list(worlds) if worlds is not None else self._batch_worlds
        arr = (C.c_void_p * len(worlds))(*[w._w.value for w in worlds])
        self._chk(self._lib.bg_batch_writeback(self._w, arr, len(worlds)))

    # ---- access (physics.rs:101-157)
    def rbd_count(self) -> int:
        n = C.c_uint32()
        self._chk(self._lib.bg_rbd_count(self._w, C.byref(n)))
        return n.value

    def get_rbd(self, h: int) -> RigidBody | None:
        c = L.RigidBodyC()
        code = self._lib.bg_get_rbd(self._w, h, C.byref(c))
        if code == L.BG_ERR_NOT_FOUND:
            return None
        self._chk(code)
        return RigidBody.from_c(c)

    def set_rbd(self, h: int, rbd: RigidBody) -> None:
        """Write-back half of get_mut_rbd (physics.rs:109-111)."""
        c = rbd.to_c()
        self._chk(self._lib.bg_set_rbd(self._w, h, C.byref(c)))

    def get_col(self, h: int) -> Collider | None:
        c = L.ColliderC()
        code = self._lib.bg_get_col(self._w, h, C.byref(c))
        if code == L.BG_ERR_NOT_FOUND:
            return None
        self._chk(code)
        return Collider.from_c(c)

    def rbd_position(self, h: int):
        b = self.get_rbd(h)
        return None if b is None else b.position

    def col_position(self, h: int):
        c = self.get_col(h)
        return None if c is None else c.absolute_translation

    def contact_count(self) -> int:
        """perf counter "collisions" summed since creation (physics.rs:316)."""
        n = C.c_uint64()
        self._chk(self._lib.bg_contact_count(self._w, C.byref(n)))
        return n.value

    # ---- bulk helpers (SoA, arena order)
    def insert_bodies(self, position, velocity_request=None, gravity_mod=None, body_type=None,
                      rotation=None) -> np.ndarray:
        pos = np.ascontiguousarray(position, dtype=np.float32).reshape(-1, 2)
        n = pos.shape[0]
        vr = None if velocity_request is None else np.ascontiguousarray(velocity_request, np.float32).reshape(-1, 2)
        gm = None if gravity_mod is None else np.ascontiguousarray(gravity_mod, np.float32)
        bt = None if body_type is None else np.ascontiguousarray(body_type, np.uint8)
        ro = None if rotation is None else np.ascontiguousarray(rotation, np.float32)
        batch = L.BodyBatchC(L.fptr(pos), L.fptr(vr), None, L.fptr(gm), L.fptr(ro), L.u8ptr(bt))
        out = np.zeros(n, np.uint64)
        self._chk(self._lib.bg_insert_rbds(self._w, n, C.byref(batch), L.hptr(out)))
        return out

    def insert_colliders(self, parents, radius, offset_translation=None, absolute_translation=None,
                         mass_override=None, has_mass_override=None, is_sensor=None, memberships=None,
                         filter=None) -> np.ndarray:
        par = np.ascontiguousarray(parents, np.uint64)
        n = par.shape[0]
        r = np.ascontiguousarray(radius, np.float32)
        off = None if offset_translation is None else np.ascontiguousarray(offset_translation, np.float32)
        ab = None if absolute_translation is None else np.ascontiguousarray(absolute_translation, np.float32)
        mo = None if mass_override is None else np.ascontiguousarray(mass_override, np.float32)
        hmo = None if has_mass_override is None else np.ascontiguousarray(has_mass_override, np.uint8)
        sen = None if is_sensor is None else np.ascontiguousarray(is_sensor, np.uint8)
        mem = None if memberships is None else np.ascontiguousarray(memberships, np.uint32)
        fil = None if filter is None else np.ascontiguousarray(filter, np.uint32)
        batch = L.ColliderBatchC(L.fptr(off), L.fptr(ab), L.fptr(r), L.fptr(mo), L.u8ptr(hmo), L.u8ptr(sen),
                                 L.u32ptr(mem), L.u32ptr(fil))
        out = np.zeros(n, np.uint64)
        self._chk(self._lib.bg_insert_colliders_with_parents(self._w, n, C.byref(batch), L.hptr(par),
                                                             L.hptr(out)))
        return out

    def bodies(self) -> dict:
        """Body state in arena order as numpy arrays."""
        n = self.rbd_count()
        pos = np.zeros((n, 2), np.float32)
        pos_old = np.zeros((n, 2), np.float32)
        vel = np.zeros((n, 2), np.float32)
        rot = np.zeros(n, np.float32)
        om = np.zeros(n, np.float32)
        hs = np.zeros(n, np.uint64)
        self._chk(self._lib.bg_get_bodies(self._w, L.fptr(pos), L.fptr(pos_old), L.fptr(vel), L.fptr(rot),
                                          L.fptr(om), L.hptr(hs)))
        return {"position": pos, "position_old": pos_old, "velocity": vel, "rotation": rot,
                "angular_velocity": om, "handles": hs}

    def collider_positions(self) -> np.ndarray:
        n = C.c_uint32()
        self._chk(self._lib.bg_get_collider_positions(self._w, None, C.byref(n)))
        out = np.zeros((n.value, 2), np.float32)
        self._chk(self._lib.bg_get_collider_positions(self._w, L.fptr(out), C.byref(n)))
        return out

    def set_broadphase(self, cell_scale: float = 1.25, rebuild_interval: int = 1, skin: float = 0.25) -> None:
        """Broadphase tuning (results never depend on it): cell size as a
        multiple of 2*max radius, steps between grid rebuilds, and the skin
        (fraction of a cell) beyond which a body is tracked as an outlier."""
        self._chk(self._lib.bg_set_broadphase(self._w, float(cell_scale), float(skin), int(rebuild_interval)))

    def set_event_capacity(self, capacity: int) -> None:
        """Record collision events (physics.rs:304-311) into a device buffer of
        `capacity` entries between drains; 0 turns recording off."""
        self._chk(self._lib.bg_set_event_capacity(self._w, int(capacity)))

    def drain_events(self, max_events: int | None = None):
        """Drain `collision_recv`: (handles [n, 2] uint64 as (col_handle_a,
        col_handle_b), impact velocities [n, 4] float32 as (vel_a, vel_b),
        events dropped to a full buffer since the last drain), in the order the
        reference sends them."""
        chunk = 1 << 16
        hs, vs = [], []
        dropped = 0
        left = max_events
        while left is None or left > 0:
            cap = chunk if left is None else min(chunk, left)
            buf = (L.CollisionEventC * cap)()
            n = C.c_uint32()
            d = C.c_uint64()
            self._chk(self._lib.bg_drain_events(self._w, buf, cap, C.byref(n), C.byref(d)))
            dropped += d.value
            arr = np.frombuffer(buf, dtype=np.dtype([("a", np.uint64), ("b", np.uint64), ("va", np.float32, 2),
                                                     ("vb", np.float32, 2)]), count=n.value)
            hs.append(np.stack([arr["a"], arr["b"]], 1))
            vs.append(np.concatenate([arr["va"], arr["vb"]], 1))
            if left is not None:
                left -= n.value
            if n.value < cap:
                break
        return np.concatenate(hs).reshape(-1, 2), np.concatenate(vs).reshape(-1, 4), dropped

    def debug_data(self) -> dict:
        """Physics::debug_data (debug.rs:34-91) as arrays in arena order:
        body and collider transforms as (column-major Mat2, translation),
        collider radii, and (body_a, body_b) positions of joints and springs."""
        cnt = np.zeros(4, np.uint32)
        self._chk(self._lib.bg_debug_data(self._w, None, None, None, None, None, L.u32ptr(cnt)))
        nb, nc, nj, ns = (int(x) for x in cnt)
        bodies = (L.Affine2C * max(nb, 1))()
        cols = (L.Affine2C * max(nc, 1))()
        rad = np.zeros(nc, np.float32)
        joints = np.zeros((nj, 4), np.float32)
        springs = np.zeros((ns, 4), np.float32)
        self._chk(self._lib.bg_debug_data(self._w, bodies, cols, L.fptr(rad), L.fptr(joints), L.fptr(springs),
                                          L.u32ptr(cnt)))
        aff = np.dtype([("m", np.float32, 4), ("t", np.float32, 2)])
        b = np.frombuffer(bodies, dtype=aff, count=nb)
        c = np.frombuffer(cols, dtype=aff, count=nc)
        return {"body_matrix": b["m"].copy(), "body_translation": b["t"].copy(), "collider_matrix": c["m"].copy(),
                "collider_translation": c["t"].copy(), "collider_radius": rad, "joints": joints,
                "springs": springs}

    def set_profiling(self, enable: bool = True) -> None:
        self._chk(self._lib.bg_set_profiling(self._w, int(bool(enable))))

    def stage_stats(self, stage: int = 0):
        """(launches, total_ms) of stage 0 (substep) or 1 (broadphase rebuild)."""
        n = C.c_uint64()
        ms = C.c_double()
        self._chk(self._lib.bg_stage_stats(self._w, int(stage), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def step_path(self) -> str:
        n = C.c_uint32()
        self._chk(self._lib.bg_step_path(self._w, C.byref(n)))
        return {0: "none", 1: "general", 2: "sphere"}[n.value]
