"""Variable step sizes, fixed_step, user forces and impulses, and mid-run
inserts/removes on the GPU, against the oracle.

* variable delta: h/old_dt != 1 on every step after the first exercises the
  Q1 ratio (physics.rs:336-338) on both step paths;
* fixed_step (physics.rs:84-99): accumulator, at most three integrations per
  call, time advanced by 1/60 each;
* RigidBody force/impulse/velocity helpers (rigid_body.rs:130-184) applied
  through get_rbd/set_rbd between steps: acceleration, torque -> angular
  velocity -> rotation, velocity_request re-entry;
* remove_rbd / remove_col / insert mid-run (physics.rs:121-172,
  collider.rs:134-164): thunderdome slot reuse decides the new arena order,
  mass is recomputed, and the state of every surviving body carries over.
The oracle is advanced through the same changes on its own arrays.  Bar:
bit-exact positions (sphere path; general path without joints).
"""
import ctypes as C

import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.physics import ColliderBuilder, RigidBodyBuilder
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(got, want, what):
    bad = np.argwhere(bits(got) != bits(want))
    assert bad.shape[0] == 0, f"{what}: {bad.shape[0]} coordinates differ, first {bad[:3].tolist()}"


def sphere_scene(seed=21, n=2048):
    return scenes.plain_large_scene(n=n, steps=8, seed=seed)


def general_scene(seed=22):
    return scenes.multi_collider_scene(n=1024, n_springs=512, n_joints=0, steps=8, seed=seed, phi=0.3)


@pytest.mark.parametrize("kind", ["sphere", "general"])
def test_variable_delta_sequence(kind):
    sc = sphere_scene() if kind == "sphere" else general_scene()
    p, _ = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    for delta in (1 / 60, 1 / 30, 1 / 144, 0.01, 1 / 60, 1 / 75):
        p.step(delta)
        o.step(delta)
    assert p.step_path() == kind
    assert_bitexact(p.bodies()["position"], o.pos, f"{kind} variable delta")
    assert np.float32(p.old_dt) == np.float32(o.w.old_dt)


def test_fixed_step_matches_oracle():
    sc = sphere_scene(seed=23)
    p, _ = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    acc = C.c_double(0.0)
    for frame in (0.005, 0.02, 0.1, 0.016, 1 / 60, 0.0):
        p.fixed_step(frame)
        o.fixed_step(frame, acc)
    assert p.accumulator == acc.value
    assert p.time == o.w.time
    assert_bitexact(p.bodies()["position"], o.pos, "fixed_step")


def test_forces_impulses_and_velocity_requests_between_steps():
    """RigidBody helpers on the GPU world (through get_rbd/set_rbd) and the
    same f32 arithmetic on the oracle's arrays."""
    sc = sphere_scene(seed=24)
    p, bh = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    f = np.float32
    p.step(sc.dt, 2)
    o.step(sc.dt, 2)
    rng = np.random.default_rng(3)
    picks = rng.choice(sc.n_bodies, 12, replace=False)
    for k, i in enumerate(picks):
        b = p.get_rbd(int(bh[i]))
        F = (f(rng.normal() * 50), f(rng.normal() * 50))
        pt = (f(b.position[0] + 0.3), f(b.position[1] - 0.2))
        m, inertia = o.a["mass"][i], o.a["inertia"][i]
        pos = o.pos[i]
        lever = (pt[0] - (pos[0] + f(0)), pt[1] - (pos[1] + f(0)))  # single centred collider: com = 0
        cross = lever[0] * F[1] - lever[1] * F[0]
        mode = k % 4
        if mode == 0:  # apply_force: acceleration += F / m
            b.apply_force(F)
            o.a["acc"][2 * i] += F[0] / m
            o.a["acc"][2 * i + 1] += F[1] / m
        elif mode == 1:  # apply_force_at_point: + torque
            b.apply_force_at_point(F, pt)
            o.a["acc"][2 * i] += F[0] / m
            o.a["acc"][2 * i + 1] += F[1] / m
            o.a["torque"][i] += cross
        elif mode == 2:  # apply_impulse_at_point: velocity request + angular velocity
            b.apply_impulse_at_point(F, pt)
            v = o.vel[i]
            o.a["vreq"][2 * i] = v[0] + F[0] / m
            o.a["vreq"][2 * i + 1] = v[1] + F[1] / m
            o.a["has_vreq"][i] = 1
            o.a["omega"][i] += cross / inertia
        else:  # set_velocity
            b.set_velocity((F[0] / 10, F[1] / 10))
            o.a["vreq"][2 * i] = F[0] / 10
            o.a["vreq"][2 * i + 1] = F[1] / 10
            o.a["has_vreq"][i] = 1
        p.set_rbd(int(bh[i]), b)
    p.step(sc.dt, 3)
    o.step(sc.dt, 3)
    got = p.bodies()
    assert_bitexact(got["position"], o.pos, "forces and impulses")
    assert_bitexact(got["rotation"], o.rot, "rotation from torque and angular impulse")
    assert np.any(got["rotation"][picks] != 0.0)


def _slot(h):
    return int(h) & 0xFFFFFFFF


def _carry_over(o_old, sc_old, body_src, col_src, new_bodies, springs_keep):
    """An oracle world for the mutated arena: `body_src` / `col_src` give, for
    each dense index of the new arena, the old dense index or ('new', k)."""
    nb, nc = len(body_src), len(col_src)
    old_b = [s for s in body_src if not isinstance(s, tuple)]
    pos = np.zeros((nb, 2), np.float32)
    for d, s in enumerate(body_src):
        pos[d] = o_old.pos[s] if not isinstance(s, tuple) else new_bodies[s[1]]["pos"]
    dense_of_old = {s: d for d, s in enumerate(body_src) if not isinstance(s, tuple)}

    def pick(arr, src, new_val):
        out = np.zeros((len(src),) + arr.shape[1:], arr.dtype)
        for d, s in enumerate(src):
            out[d] = arr[s] if not isinstance(s, tuple) else new_val(s[1])
        return out

    col_parent = np.zeros(nc, np.int32)
    for d, s in enumerate(col_src):
        if isinstance(s, tuple):
            col_parent[d] = body_src.index(("new", s[1]))
        else:
            col_parent[d] = dense_of_old[int(sc_old.col_parent[s])]
    nbd = new_bodies
    sc = scenes.Scene(
        pos=pos, vreq=pick(sc_old.vreq, body_src, lambda k: nbd[k]["vel"]),
        has_vreq=pick(sc_old.has_vreq, body_src, lambda k: 1),
        gravity_mod=pick(sc_old.gravity_mod, body_src, lambda k: 1.0),
        body_type=pick(sc_old.body_type, body_src, lambda k: 0),
        rotation=pick(sc_old.rotation, body_src, lambda k: 0.0),
        col_parent=col_parent, col_off=pick(sc_old.col_off, col_src, lambda k: (0.0, 0.0)),
        col_abs=pick(o_old.abs_t, col_src, lambda k: nbd[k]["pos"]),
        col_r=pick(sc_old.col_r, col_src, lambda k: nbd[k]["r"]),
        col_mass_override=pick(sc_old.col_mass_override, col_src, lambda k: 0.0),
        col_has_mo=pick(sc_old.col_has_mo, col_src, lambda k: 0),
        col_sensor=pick(sc_old.col_sensor, col_src, lambda k: 0),
        col_mem=pick(sc_old.col_mem, col_src, lambda k: 0xFFFFFFFF),
        col_filt=pick(sc_old.col_filt, col_src, lambda k: 0xFFFFFFFF),
        spring_a=np.array([dense_of_old[int(sc_old.spring_a[q])] for q in springs_keep], np.int32),
        spring_b=np.array([dense_of_old[int(sc_old.spring_b[q])] for q in springs_keep], np.int32),
        spring_rest=sc_old.spring_rest[springs_keep].copy(), spring_k=sc_old.spring_k[springs_keep].copy(),
        spring_c=sc_old.spring_c[springs_keep].copy(), gravity=sc_old.gravity, dt=sc_old.dt)
    o = OracleWorld(sc, np.float32)
    for d, s in enumerate(body_src):
        if isinstance(s, tuple):
            continue
        for name in ("pos_old", "vel", "acc"):
            o.a[name][2 * d:2 * d + 2] = o_old.a[name][2 * s:2 * s + 2]
        for name in ("rot", "omega", "torque"):
            o.a[name][d] = o_old.a[name][s]
        o.a["has_vreq"][d] = o_old.a["has_vreq"][s]
        o.a["vreq"][2 * d:2 * d + 2] = o_old.a["vreq"][2 * s:2 * s + 2]
    o.w.old_dt = o_old.w.old_dt
    assert len(old_b) + len(new_bodies) == nb
    return o, sc


def test_remove_and_insert_bodies_mid_run_sphere_path():
    sc = sphere_scene(seed=25)
    p, bh = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    p.step(sc.dt, 3)
    o.step(sc.dt, 3)
    removed = [5, 700, 701, 1500]
    for i in removed:
        p.remove_rbd(int(bh[i]))
    new_bodies = [{"pos": (10.0, 12.0), "vel": (1.0, -2.0), "r": 0.75},
                  {"pos": (30.5, 8.25), "vel": (-0.5, 0.0), "r": 0.6}]
    new_h, new_c = [], []
    for nbd in new_bodies:
        h = p.insert_rbd(RigidBodyBuilder().position(nbd["pos"]).velocity_request(nbd["vel"]).build())
        c = p.insert_collider_with_parent(
            ColliderBuilder().radius(nbd["r"]).absolute_transform_translation(nbd["pos"]).build(), h)
        new_h.append(h)
        new_c.append(c)
    # thunderdome: the last freed slot is reused first
    assert [_slot(h) for h in new_h] == [1500, 701]
    slot_src = {i: i for i in range(sc.n_bodies) if i not in removed}
    for k, h in enumerate(new_h):
        slot_src[_slot(h)] = ("new", k)
    body_src = [slot_src[s] for s in sorted(slot_src)]
    col_slot_src = {i: i for i in range(sc.n_colliders) if i not in removed}
    for k, c in enumerate(new_c):
        col_slot_src[_slot(c)] = ("new", k)
    col_src = [col_slot_src[s] for s in sorted(col_slot_src)]
    o2, _ = _carry_over(o, sc, body_src, col_src, new_bodies, [])
    p.step(sc.dt, 4)
    o2.step(sc.dt, 4)
    assert p.step_path() == "sphere"
    assert_bitexact(p.bodies()["position"], o2.pos, "after removes and inserts")
    assert p.contact_count() == o.contact_count + o2.contact_count


def test_remove_colliders_mid_run_recomputes_mass():
    """remove_col on multi-collider bodies (general path): the parent's mass
    and centre of mass are recomputed from its remaining colliders
    (collider.rs:134-141); a body's last collider takes the body with it."""
    sc = general_scene(seed=26)
    p, bh = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    p.step(sc.dt, 2)
    o.step(sc.dt, 2)
    counts = np.bincount(sc.col_parent, minlength=sc.n_bodies)
    start = np.concatenate([[0], np.cumsum(counts)])
    multi = np.flatnonzero(counts >= 3)[:3]
    in_springs = set(sc.spring_a.tolist()) | set(sc.spring_b.tolist())
    single = [b for b in np.flatnonzero(counts == 1) if b not in in_springs][:1]
    assert len(multi) == 3 and len(single) == 1
    drop_cols = [int(start[b]) + 1 for b in multi] + [int(start[single[0]])]
    mass_before = [p.get_rbd(int(bh[b])).calculated_mass for b in multi]
    for c in drop_cols:
        p.remove_col((1 << 32) | c)  # a fresh arena's handles: generation 1, slot = index
    assert p.get_rbd(int(bh[single[0]])) is None
    for b, m0 in zip(multi, mass_before):
        assert p.get_rbd(int(bh[b])).calculated_mass < m0
    body_src = [i for i in range(sc.n_bodies) if i != single[0]]
    col_src = [c for c in range(sc.n_colliders) if c not in drop_cols]
    o2, sc2 = _carry_over(o, sc, body_src, col_src, [], list(range(sc.spring_a.shape[0])))
    for d, b in enumerate(body_src):
        assert np.float32(p.get_rbd(int(bh[b])).calculated_mass) == o2.a["mass"][d]
    p.step(sc.dt, 3)
    o2.step(sc.dt, 3)
    assert p.step_path() == "general"
    assert_bitexact(p.bodies()["position"], o2.pos, "after collider removal")


@pytest.mark.parametrize("kind", ["sphere", "general"])
def test_non_finite_rotation_makes_the_collider_translation_nan(kind):
    """A body whose rotation is inf or NaN: the reference's collider
    translation position + R(rotation) * offset is NaN even for a zero offset
    (physics.rs:360-366), and that collider then touches nothing; every
    other coordinate stays bit-exact against the oracle."""
    sc = sphere_scene(seed=26, n=1024) if kind == "sphere" else general_scene(seed=27)
    p, bh = sc.build_physics()
    o = OracleWorld(sc, np.float32)
    for i, v in ((5, np.inf), (9, np.nan), (12, -np.inf)):
        b = p.get_rbd(int(bh[i]))
        b.rotation = float(v)
        p.set_rbd(int(bh[i]), b)
        o.a["rot"][i] = v
    p.step(sc.dt, 3)
    o.step(sc.dt, 3)
    assert p.step_path() == kind
    cols = p.collider_positions()
    nan = np.isnan(o.abs_t)
    assert nan.any() and np.array_equal(np.isnan(cols), nan)
    assert_bitexact(np.where(nan, 0.0, cols), np.where(nan, 0.0, o.abs_t), "collider translation")
    assert_bitexact(p.bodies()["position"], o.pos, "position")
    assert p.contact_count() == o.contact_count
