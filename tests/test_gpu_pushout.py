"""Push-out (physics.rs:272-286) through the C ABI, against the oracle.

A pair whose snapshots coincide (dist < 1e-6) moves both bodies by -+0.01 in
x and rewrites both colliders' snapshots from the bodies' live positions;
every later pair of the (i, j<i) loop reads the rewritten values.  The
device replays those pairs in reference order (the ordered contact pass of
the general path; a single-sphere world is stepped there while its snapshot
holds coincident pairs).  Bar: bit-exact positions and contact counts.

The demo test follows the reference's only caller of this branch:
BallsDemo::update (demo/src/demos/balls.rs:92-181) calls fixed_step, then
every 0.1 s spawns 20 balls at one point random_around((1,1), 0.1, 0.2) with
random_circle(3.0) velocities and radii U[0.05, 0.1] through
spawn_rbd_entity (demo/src/simulation.rs:85-146), in a world of gravity
(0, -30) with a circle constraint of radius 4 at the origin
(demo/src/main.rs:48-56).  Its RNG is macroquad's; a seeded numpy RNG draws
the same distributions here.
"""
import ctypes as C

import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.physics import ColliderBuilder, Constraint, Physics, RigidBodyBuilder, Spring, STATIC
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(got, want, what):
    bad = np.argwhere(bits(got) != bits(want))
    assert bad.shape[0] == 0, (f"{what}: {bad.shape[0]} coordinates differ, first {bad[:3].tolist()}: "
                               f"gpu {np.asarray(got)[tuple(bad[0])]} oracle {np.asarray(want)[tuple(bad[0])]}")


class GrowingOracle:
    """Oracle world that bodies are appended to between frames (no removals:
    arena order is insertion order).  Each append rebuilds the oracle arrays
    and carries every existing body's state over."""

    def __init__(self, gravity, constraint=None, static_spring_pair=False, event_cap=0):
        self.gravity = gravity
        self.event_cap = event_cap
        self.con = constraint
        self.pos = np.zeros((0, 2), np.float32)
        self.vreq = np.zeros((0, 2), np.float32)
        self.rad = np.zeros(0, np.float32)
        self.static = np.zeros(0, np.uint8)
        self.has_col = np.zeros(0, bool)
        self.spring = static_spring_pair
        self.o = None
        self.acc = C.c_double(0.0)
        self.time = 0.0  # of the empty world, before the first body (physics.rs:81, 84-99)
        self.contacts = 0
        if static_spring_pair:  # two static bodies without colliders, far away
            self._append([[100.0, 100.0], [101.0, 100.0]], [[0, 0], [0, 0]], [0.0, 0.0], static=True, cols=False)

    def _append(self, pos, vel, rad, static=False, cols=True):
        n0 = self.pos.shape[0]
        self.pos = np.concatenate([self.pos, np.asarray(pos, np.float32)])
        self.vreq = np.concatenate([self.vreq, np.asarray(vel, np.float32)])
        self.rad = np.concatenate([self.rad, np.asarray(rad, np.float32)])
        k = len(pos)
        self.static = np.concatenate([self.static, np.full(k, 1 if static else 0, np.uint8)])
        self.has_col = np.concatenate([self.has_col, np.full(k, cols)])
        old = self.o
        n = self.pos.shape[0]
        cb = np.flatnonzero(self.has_col).astype(np.int32)
        pos_now = self.pos.copy()
        abs_now = self.pos[cb].copy()
        if old is not None:
            pos_now[:n0] = old.pos
            abs_now[: old.abs_t.shape[0]] = old.abs_t
        sc = scenes.Scene(
            pos=pos_now, vreq=self.vreq.copy(), has_vreq=np.ones(n, np.uint8), gravity_mod=np.ones(n, np.float32),
            body_type=self.static.copy(), rotation=np.zeros(n, np.float32), col_parent=cb,
            col_off=np.zeros((cb.size, 2), np.float32), col_abs=abs_now, col_r=self.rad[cb].copy(),
            col_mass_override=np.zeros(cb.size, np.float32), col_has_mo=np.zeros(cb.size, np.uint8),
            col_sensor=np.zeros(cb.size, np.uint8), col_mem=np.full(cb.size, 0xFFFFFFFF, np.uint32),
            col_filt=np.full(cb.size, 0xFFFFFFFF, np.uint32), gravity=self.gravity, dt=1 / 60)
        if self.spring:
            sc.spring_a = np.array([0], np.int32)
            sc.spring_b = np.array([1], np.int32)
            sc.spring_rest = np.array([1.0], np.float32)
            sc.spring_k = np.array([0.0], np.float32)
            sc.spring_c = np.array([0.0], np.float32)
        if self.con is not None:
            sc.con_pos = np.array([self.con[0]], np.float32)
            sc.con_r = np.array([self.con[1]], np.float32)
        o = OracleWorld(sc, np.float32, event_cap=self.event_cap)
        if old is not None:
            for name in ("pos_old", "vel", "acc", "vreq"):
                o.a[name][: 2 * n0] = old.a[name][: 2 * n0]
            for name in ("rot", "omega", "torque"):
                o.a[name][:n0] = old.a[name][:n0]
            o.a["has_vreq"][:n0] = old.a["has_vreq"][:n0]
            o.w.old_dt = old.w.old_dt
            o.w.time = old.w.time
            self.contacts += old.contact_count
        else:
            o.w.time = self.time
        self.o = o

    def spawn(self, pos, vel, rad):
        self._append(pos, vel, rad)

    def fixed_step(self, frame):
        if self.o is not None:
            self.o.fixed_step(frame, self.acc)
            return
        # an empty world only advances its clock (physics.rs:84-99)
        self.acc.value += frame
        n = 3
        while self.acc.value >= 1.0 / 60.0 and n > 0:
            self.time += 1.0 / 60.0
            self.acc.value -= 1.0 / 60.0
            n -= 1

    @property
    def contact_count(self):
        return self.contacts + self.o.contact_count


def gpu_world(gravity, constraint=None, static_spring_pair=False):
    p = Physics(gravity, False)
    if constraint is not None:
        p.constraints.push(Constraint(constraint[0], constraint[1]))
    handles = []
    if static_spring_pair:
        for x in (100.0, 101.0):
            handles.append(p.insert_rbd(RigidBodyBuilder().position((x, 100.0)).body_type(STATIC).build()))
        p.springs.insert(Spring(handles[0], handles[1], 1.0, 0.0, 0.0))
    return p, handles


def gpu_spawn(p, handles, pos, vel, rad):
    for (x, y), (vx, vy), r in zip(pos, vel, rad):
        h = p.insert_rbd(RigidBodyBuilder().position((float(x), float(y))).velocity_request((float(vx), float(vy)))
                         .build())
        p.insert_collider_with_parent(
            ColliderBuilder().radius(float(r)).absolute_transform_translation((float(x), float(y))).build(), h)
        handles.append(h)


@pytest.mark.parametrize("path", ["sphere", "general"])
def test_three_coincident_spheres_bitexact(path):
    """Three equal spheres at one point: the second and third pairs read the
    snapshots the first pair's push-out rewrote."""
    p, hs = gpu_world((0.0, 0.0), static_spring_pair=path == "general")
    o = GrowingOracle((0.0, 0.0), static_spring_pair=path == "general")
    pos = np.zeros((3, 2), np.float32)
    vel = np.zeros((3, 2), np.float32)
    rad = np.full(3, 0.1, np.float32)
    gpu_spawn(p, hs, pos, vel, rad)
    o.spawn(pos, vel, rad)
    for _ in range(3):
        p.step(1 / 60)
        o.o.step(1 / 60)
    assert p.step_path() == "general"  # the single-sphere world too: its snapshot had coincident pairs
    got = np.array([p.get_rbd(h).position for h in hs], np.float32)
    assert_bitexact(got, o.o.pos, "three coincident spheres")
    assert p.contact_count() == o.contact_count


def _demo_burst(rng):
    ang = rng.uniform(0, 2 * np.pi)
    d = rng.uniform(0.1, 0.2)
    at = np.array([1.0 + d * np.cos(ang), 1.0 + d * np.sin(ang)], np.float32)
    pos = np.repeat(at[None], 20, axis=0)
    rr = 3.0 * np.sqrt(rng.uniform(0, 1, 20))
    th = rng.uniform(0, 2 * np.pi, 20)
    vel = np.stack([rr * np.cos(th), rr * np.sin(th)], 1).astype(np.float32)
    rad = rng.uniform(0.05, 0.1, 20).astype(np.float32)
    return pos, vel, rad


@pytest.mark.parametrize("path", ["sphere", "general"])
def test_demo_spawn_60_frames_bitexact(path):
    """BallsDemo: fixed_step(1/60) per frame, then 20 coincident balls every
    6th frame while fewer than 200 bodies exist; 60 frames.  Positions,
    velocities and the contact count after every frame equal the oracle's."""
    rng = np.random.default_rng(2024)
    general = path == "general"
    p, hs = gpu_world((0.0, -30.0), ((0.0, 0.0), 4.0), static_spring_pair=general)
    o = GrowingOracle((0.0, -30.0), ((0.0, 0.0), 4.0), static_spring_pair=general)
    n_balls = 0
    paths = set()
    for frame in range(60):
        p.fixed_step(1 / 60)
        o.fixed_step(1 / 60)
        if n_balls:
            paths.add(p.step_path())
            b = p.bodies()
            assert_bitexact(b["position"], o.o.pos, f"frame {frame} position")
            assert_bitexact(b["velocity"], o.o.vel, f"frame {frame} velocity")
            assert p.contact_count() == o.contact_count, f"frame {frame} contacts"
        if frame % 6 == 0 and n_balls < 200:
            pos, vel, rad = _demo_burst(rng)
            gpu_spawn(p, hs, pos, vel, rad)
            o.spawn(pos, vel, rad)
            n_balls += 20
    assert n_balls == 200
    assert "general" in paths
    if not general:
        # between bursts the world returns to the fused path once push-outs stop
        assert "sphere" in paths
    assert p.time == o.o.w.time


def test_push_out_cluster_in_a_crowd_bitexact():
    """20 coincident spheres inside 1,200 others on the general path (a
    spring elsewhere): the ordered pass replays only the pairs the push-outs
    make dirty and leaves every other body's snapshot sum in place."""
    rng = np.random.default_rng(5)
    crowd = rng.uniform(-8, 8, (1200, 2)).astype(np.float32)
    pos = np.concatenate([crowd[:600], np.full((20, 2), 0.5, np.float32), crowd[600:]])
    vel = rng.normal(0, 1, pos.shape).astype(np.float32)
    rad = rng.uniform(0.1, 0.3, pos.shape[0]).astype(np.float32)
    p, hs = gpu_world((0.0, -9.81), static_spring_pair=True)
    o = GrowingOracle((0.0, -9.81), static_spring_pair=True)
    gpu_spawn(p, hs, pos, vel, rad)
    o.spawn(pos, vel, rad)
    for k in range(4):
        p.step(1 / 60)
        o.o.step(1 / 60)
        assert_bitexact(p.bodies()["position"], o.o.pos, f"step {k}")
        assert p.contact_count() == o.contact_count


def test_push_out_events_match_oracle():
    """Collision events of push-out substeps: a pair that stops (or starts)
    being a contact once snapshots are rewritten has its event withdrawn (or
    added), so the drained stream is the oracle's: same colliders, order and
    impact velocities."""
    rng = np.random.default_rng(11)
    p, hs = gpu_world((0.0, -30.0), ((0.0, 0.0), 4.0), static_spring_pair=True)
    p.set_event_capacity(1 << 16)
    o = GrowingOracle((0.0, -30.0), ((0.0, 0.0), 4.0), static_spring_pair=True, event_cap=1 << 16)
    for burst in range(3):
        pos, vel, rad = _demo_burst(rng)
        if burst:
            pos = pos + np.float32(0.5 * burst)
        gpu_spawn(p, hs, pos, vel, rad)
    # one oracle world holding all 60 balls (spawned together here)
    allp, allv, allr = [], [], []
    rng = np.random.default_rng(11)
    for burst in range(3):
        pos, vel, rad = _demo_burst(rng)
        if burst:
            pos = pos + np.float32(0.5 * burst)
        allp.append(pos), allv.append(vel), allr.append(rad)
    o.spawn(np.concatenate(allp), np.concatenate(allv), np.concatenate(allr))
    for _ in range(3):
        p.step(1 / 60)
        o.o.step(1 / 60)
    got_h, got_v, dropped = p.drain_events()
    cols, vels = o.o.events
    assert dropped == 0 and o.o.w.n_events == cols.shape[0]
    assert got_h.shape[0] == cols.shape[0] == p.contact_count() == o.contact_count
    got = (got_h & np.uint64(0xFFFFFFFF)).astype(np.int64)  # collider slots 0.. in insertion order
    assert np.array_equal(got, cols.astype(np.int64))
    assert np.array_equal(bits(got_v), bits(vels))
