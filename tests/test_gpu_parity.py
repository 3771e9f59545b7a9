"""Parity of the HIP step (through the C ABI) with the oracle.

Bars (from the path's definition):
  * integration-only steps and the contact pair set: bit-exact;
  * contact / spring / joint resolution: max |dx| <= 1e-4 * radius after one
    step.  Contacts are summed per body in the reference's (i, j<i) order,
    so sphere scenes are expected bit-exact too and are asserted so.
"""
import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.physics import (ColliderBuilder, Constraint, InteractionGroups, Physics, RigidBodyBuilder, Spring,
                               STATIC)
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(gpu, ora, what):
    g, o = bits(gpu), bits(ora)
    bad = np.argwhere(g != o)
    assert bad.shape[0] == 0, (f"{what}: {bad.shape[0]} of {g.size} differ; first at {bad[:3].tolist()}: "
                               f"gpu {np.asarray(gpu)[tuple(bad[0])]} oracle {np.asarray(ora)[tuple(bad[0])]}")


def gpu_world(scene, **params):
    p, handles = scene.build_physics()
    for k, v in params.items():
        setattr(p, k, v)
    return p


def compare_state(p, o, tag):
    b = p.bodies()
    assert_bitexact(b["position"], o.pos, f"{tag} position")
    assert_bitexact(b["position_old"], o.pos_old, f"{tag} position_old")
    assert_bitexact(b["velocity"], o.vel, f"{tag} velocity")
    assert_bitexact(b["rotation"], o.rot, f"{tag} rotation")
    assert_bitexact(p.collider_positions(), o.abs_t, f"{tag} collider translation")


def test_integration_only_bitexact_600_steps():
    """Gravity + Verlet + velocity requests (Q1 included), 600 x step(1/60)."""
    sc = scenes.plain_small_scene(steps=600)
    p = gpu_world(sc, collisions_enabled=False)
    o = OracleWorld(sc, np.float32, collisions_enabled=False)
    p.step(sc.dt, 600)
    o.step(sc.dt, 600)
    compare_state(p, o, "integration-only")
    assert p.contact_count() == 0


@pytest.mark.parametrize("steps", [1, 30])
def test_plain_small_contacts_bitexact(steps):
    """1,024 spheres + static walls, contacts on: positions and contact count bit-exact."""
    sc = scenes.plain_small_scene(steps=steps)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for k in range(steps):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.contact_count() == o.contact_count, f"step {k}: contact count"
    compare_state(p, o, f"{steps} steps")
    assert p.step_path() == "sphere"


def test_general_path_equals_sphere_path():
    """A zero spring forces the general (multi-kernel) path; it adds +0 to
    accelerations, so both device paths must agree bit for bit."""
    sc = scenes.plain_small_scene(steps=20)
    p1 = gpu_world(sc)
    p2, bh = sc.build_physics()
    p2.springs.insert(Spring(int(bh[0]), int(bh[1]), 0.0, 0.0, 0.0))
    p1.step(sc.dt, 20)
    p2.step(sc.dt, 20)
    assert p1.step_path() == "sphere" and p2.step_path() == "general"
    assert_bitexact(p1.bodies()["position"], p2.bodies()["position"], "sphere vs general")
    assert p1.contact_count() == p2.contact_count()


@pytest.mark.parametrize("interval,scale,skin", [(1, 1.0, 0.125), (3, 1.0, 0.01), (8, 1.5, 0.5), (50, 0.5, 0.02)])
def test_broadphase_settings_do_not_change_results(interval, scale, skin):
    """Rebuild interval, cell size and skin (tiny skins force many outliers)
    change only speed: positions and contact counts stay bit-exact."""
    sc = scenes.plain_small_scene(steps=12)
    o = OracleWorld(sc, np.float32)
    o.step(sc.dt, 12)
    p = gpu_world(sc)
    p.set_broadphase(scale, interval, skin)
    p.step(sc.dt, 12)
    compare_state(p, o, f"interval {interval} scale {scale} skin {skin}")
    assert p.contact_count() == o.contact_count


def test_plain_large_one_step_bitexact_1m():
    """1,048,576 spheres at packing 0.4: one step(1/120) bit-exact vs the
    oracle (grid-enumerated, same (i, j) visit order as the O(C^2) loop)."""
    sc = scenes.plain_large_scene(steps=1)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32, use_grid=True)
    p.step(sc.dt)
    o.step(sc.dt)
    assert p.contact_count() == o.contact_count
    compare_state(p, o, "1M one step")


def test_plain_large_invariants_100_steps():
    """Size-independent properties at the full 1M shape over 100 steps.

    The reference's f32 trajectory is not a free fall of the centroid: for
    |y| above ~700 the gravity increment g*h^2 ~ 1e-5 is below half an ulp of
    the position and `p + (disp + a h^2)` drops it (the f32 oracle shows the
    same), so the checks are exact ones instead: the result does not depend
    on the broadphase rebuild interval or cell size (the pair set is exact
    for every setting), positions stay finite, and along x (no gravity) the
    mass-weighted centroid follows the initial velocities, because contacts
    conserve sum(m x)."""
    sc = scenes.plain_large_scene(steps=100)
    p1 = gpu_world(sc)
    p2 = gpu_world(sc)
    p2.set_broadphase(1.5, 1, 0.05)
    p1.step(sc.dt, 100)
    p2.step(sc.dt, 100)
    x = p1.bodies()["position"]
    assert np.isfinite(x).all()
    assert_bitexact(x, p2.bodies()["position"], "rebuild every 4 steps vs every step (cell x1.5)")
    assert p1.contact_count() == p2.contact_count()
    m = 4.0 * sc.col_r.astype(np.float64)  # Q2: mass = 2 * (2 r)
    h = sc.dt / 8.0
    k = 800
    vx = sc.vreq[:, 0].astype(np.float64)
    vx[0] *= h  # Q1: the first body's request is scaled by h / old_dt = h
    pred = sc.pos[:, 0].astype(np.float64) + k * h * vx
    cm = (m * x[:, 0]).sum() / m.sum()
    assert abs(cm - (m * pred).sum() / m.sum()) < 1e-3


def test_plain_large_sphere_vs_general_path_20_steps():
    """At 1M the fused sphere kernel (cell-ordered bodies) and the general
    multi-kernel path (arena-ordered bodies, sorted collider index) are two
    independent device implementations of the same step: bit-identical."""
    sc = scenes.plain_large_scene(steps=20)
    p1 = gpu_world(sc)
    p2, bh = sc.build_physics()
    p2.springs.insert(Spring(int(bh[0]), int(bh[1]), 0.0, 0.0, 0.0))
    p1.step(sc.dt, 20)
    p2.step(sc.dt, 20)
    assert p1.step_path() == "sphere" and p2.step_path() == "general"
    assert_bitexact(p1.bodies()["position"], p2.bodies()["position"], "sphere vs general at 1M")
    assert p1.contact_count() == p2.contact_count()


def test_multi_collider_springs_joints_tolerance():
    """Multi-collider bodies + springs + fixed joints (general path).

    Springs and contacts are summed in reference order and joints run as
    sequential Gauss-Seidel per component, so the only source of difference
    is the last bit of glibc's atan2f/sinf/cosf (not correctly rounded; the
    device evaluates in f64 and rounds).  Bar: max |dx| <= 1e-4 * (smallest
    collider radius) after one step; bodies that never rotate are exact."""
    sc = scenes.multi_collider_scene(n=4096, n_springs=2048, n_joints=1024, steps=10)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    p.step(sc.dt)
    o.step(sc.dt)
    b = p.bodies()
    err = np.abs(b["position"].astype(np.float64) - o.pos.astype(np.float64)).max()
    assert err <= 1e-4 * float(sc.col_r.min()), err
    assert p.step_path() == "general"
    jointed = np.zeros(sc.n_bodies, bool)
    jointed[sc.joint_a] = True
    jointed[sc.joint_b] = True
    rot_err = np.abs(b["rotation"].astype(np.float64) - o.rot.astype(np.float64))
    assert rot_err[~jointed].max() == 0.0
    assert rot_err.max() < 1e-5


def test_groups_sensors_constraints_static():
    """Collision groups, sensors, circle constraints and static bodies."""
    rng = np.random.default_rng(5)
    n = 600
    pos = rng.uniform(0, 20, (n, 2)).astype(np.float32)
    vel = rng.normal(0, 1, (n, 2)).astype(np.float32)
    rad = rng.uniform(0.2, 0.6, n).astype(np.float32)
    sc = scenes._single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1 / 60, 10, {})
    sc.col_mem = rng.choice(np.array([1, 2, 3, 0xFFFFFFFF], np.uint32), n)
    sc.col_filt = rng.choice(np.array([1, 2, 0xFFFFFFFF], np.uint32), n)
    sc.col_sensor = (rng.uniform(size=n) < 0.1).astype(np.uint8)
    sc.body_type = (rng.uniform(size=n) < 0.1).astype(np.uint8)
    sc.con_pos = np.array([[10.0, 10.0]], np.float32)
    sc.con_r = np.array([9.0], np.float32)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for _ in range(10):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.contact_count() == o.contact_count
    compare_state(p, o, "groups/sensors/constraints/static")


def test_reference_api_construction_and_step():
    """The reference's own construction recipe, one body at a time."""
    p = Physics((0.0, -30.0), False)
    p.constraints.push(Constraint((0.0, 0.0), 4.0))
    hs = []
    for k in range(50):
        x, y = (k % 10) * 0.3 - 1.5, (k // 10) * 0.3
        h = p.insert_rbd(RigidBodyBuilder().position((x, y)).velocity_request((0.1 * k, 0.0)).build())
        p.insert_collider_with_parent(
            ColliderBuilder().radius(0.1 + 0.01 * k).absolute_transform_translation((x, y)).build(), h)
        hs.append(h)
    for _ in range(5):
        p.fixed_step(1 / 60)
    b = p.get_rbd(hs[3])
    assert np.all(np.isfinite(b.position))
    assert p.get_rbd(hs[3]).velocity_request is None


def test_push_out_pair_matches_oracle():
    """Two colliders of different bodies at the same point (physics.rs:272-286):
    +-0.01 push-out, then the contact correction from the new axis; an
    isolated pair has no order dependence, so the result is exact."""
    p = Physics((0.0, 0.0), False)
    hs = []
    for x in (1.0, 1.0):
        h = p.insert_rbd(RigidBodyBuilder().position((x, 2.0)).build())
        p.insert_collider_with_parent(ColliderBuilder().radius(0.5).absolute_transform_translation((x, 2.0)).build(),
                                      h)
        hs.append(h)
    p.substeps = 1
    p.step(1 / 60)
    pos = np.array([[1, 2], [1, 2]], np.float32)
    sc = scenes._single_sphere_scene(pos, np.zeros((2, 2), np.float32), np.array([0.5, 0.5], np.float32),
                                     (0.0, 0.0), 1 / 60, 1, {})
    sc.has_vreq[:] = 0
    o = OracleWorld(sc, np.float32, substeps=1)
    o.step(1 / 60)
    got = np.array([p.get_rbd(h).position for h in hs], np.float32)
    assert_bitexact(got, o.pos, "push-out pair")
    assert p.contact_count() == o.contact_count == 1


def test_empty_and_single_body_worlds():
    p = Physics((0.0, -9.81), False)
    p.step(1 / 60)  # nothing to do
    h = p.insert_rbd(RigidBodyBuilder().position((0.0, 0.0)).build())
    p.step(1 / 60)
    assert p.get_rbd(h).position[1] < 0.0


def test_joint_components_small_and_large():
    """Joint components on both sides of the size at which the device keeps a
    component in LDS (16 bodies and 16 joints): chains of 1-40 joints, stars
    around a static hub with 15 and 16 spokes, a ring and a repeated pair.
    Colliders have no offset, so rotations (whose atan2 may differ from
    glibc's in the last bit) never reach positions: positions are bit-exact
    over several steps; rotations agree to 1e-5."""
    rng = np.random.default_rng(11)
    pos, ja, jb, static = [], [], [], []

    def body(x, y, is_static=False):
        pos.append((x, y))
        static.append(is_static)
        return len(pos) - 1

    row = 0.0
    for length in (1, 2, 5, 15, 16, 40):  # 15 joints: 16 bodies (LDS); 16: 17 bodies (global)
        ids = [body(1.0 * k, row) for k in range(length + 1)]
        ja += ids[:-1]
        jb += ids[1:]
        row += 3.0
    for spokes in (15, 16):
        hub = body(0.0, row, True)
        for k in range(spokes):
            ang = 2 * np.pi * k / spokes
            ja.append(hub)
            jb.append(body(2.5 * np.cos(ang), row + 2.5 * np.sin(ang)))
        row += 8.0
    ring = [body(2.0 * np.cos(2 * np.pi * k / 8), row + 2.0 * np.sin(2 * np.pi * k / 8)) for k in range(8)]
    ja += ring
    jb += ring[1:] + ring[:1]
    row += 6.0
    a, b = body(0.0, row), body(1.2, row)
    ja += [a, b]
    jb += [b, a]
    pos = np.array(pos, np.float32)
    n = pos.shape[0]
    vel = rng.normal(0, 2, (n, 2)).astype(np.float32)
    rad = np.full(n, 0.4, np.float32)
    sc = scenes._single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1 / 60, 3, {}, static_mask=np.array(static))
    sc.joint_a = np.array(ja, np.int32)
    sc.joint_b = np.array(jb, np.int32)
    nj = sc.joint_a.shape[0]
    sc.joint_anchor_a = rng.uniform(-0.2, 0.2, (nj, 2)).astype(np.float32)
    sc.joint_anchor_b = rng.uniform(-0.2, 0.2, (nj, 2)).astype(np.float32)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for step in range(3):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.step_path() == "general"
        b = p.bodies()
        assert_bitexact(b["position"], o.pos, f"step {step} position")
        rot_err = np.abs(b["rotation"].astype(np.float64) - o.rot.astype(np.float64))
        assert rot_err.max() < 1e-5, rot_err.max()
    moved = np.abs(p.bodies()["position"] - pos).max(axis=1)
    assert (moved[~np.array(static)] > 0).all()


def test_crowded_bodies_beyond_the_contact_lists_bitexact():
    """160 spheres of radius 0.8-1.6 in a 6 x 6 box: every body overlaps most
    of the others (about 75 contacts each), far more than either path's
    register list holds.  The sphere path and the general path (forced by a
    zero spring, which adds +0 to accelerations) apply the rest in further
    passes over the candidates, in the same ascending order, so both agree
    bit for bit with the oracle."""
    rng = np.random.default_rng(7)
    n = 160
    pos = rng.uniform(0.0, 6.0, (n, 2)).astype(np.float32)
    vel = rng.normal(0.0, 0.5, (n, 2)).astype(np.float32)
    rad = rng.uniform(0.8, 1.6, n).astype(np.float32)
    sc = scenes._single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1 / 60, 3, {})
    o = OracleWorld(sc, np.float32)
    p1 = gpu_world(sc)
    p2, bh = sc.build_physics()
    p2.springs.insert(Spring(int(bh[0]), int(bh[1]), 0.0, 0.0, 0.0))
    for step in range(3):
        o.step(sc.dt)
        p1.step(sc.dt)
        p2.step(sc.dt)
        assert p1.step_path() == "sphere" and p2.step_path() == "general"
        assert p1.contact_count() == p2.contact_count() == o.contact_count
        assert_bitexact(p1.bodies()["position"], o.pos, f"step {step} sphere path")
        assert_bitexact(p2.bodies()["position"], o.pos, f"step {step} general path")
    assert o.contact_count > 25 * n  # pairs: over 50 contacts per body on average


@pytest.mark.parametrize("masses", ["default", "override_half_r", "override_r", "one_off", "gravity_mod"])
def test_partner_masses_from_radius_or_loaded(masses):
    """The sphere path takes a partner's mass from the radius in its
    snapshot when every body's mass is exactly 2 r or 4 r (checked at
    upload), and loads the mass records otherwise.  Crowded spheres with
    forced overlaps, both ways, bit-exact against the oracle over 20 steps:
    default masses (4 r), overrides of r / 2 on every collider (mass r after
    the double push: neither 2 r nor 4 r), overrides of r (mass 2 r),
    default masses but for one body (loaded), and one body with a
    gravity_mod of 0.5 (a steady substep skips the mass record only while
    every mass follows the radius and every gravity_mod is 1)."""
    rng = np.random.default_rng(11)
    n = 3000
    pos = rng.uniform(0.0, 60.0, (n, 2)).astype(np.float32)
    vel = rng.normal(0.0, 1.0, (n, 2)).astype(np.float32)
    rad = rng.uniform(0.3, 0.6, n).astype(np.float32)
    sc = scenes._single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1 / 60, 20, {})
    if masses == "override_half_r":
        sc.col_has_mo[:] = 1
        sc.col_mass_override[:] = rad * np.float32(0.5)
    elif masses == "override_r":
        sc.col_has_mo[:] = 1
        sc.col_mass_override[:] = rad
    elif masses == "one_off":
        sc.col_has_mo[17] = 1
        sc.col_mass_override[17] = np.float32(3.0)
    elif masses == "gravity_mod":
        sc.gravity_mod[5] = np.float32(0.5)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for step in range(20):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.contact_count() == o.contact_count, f"step {step}: contact count"
    assert p.step_path() == "sphere"
    assert o.contact_count > n // 4
    compare_state(p, o, masses)


def test_bodies_with_many_colliders_bitexact():
    """Bodies of 12 colliders each (the general contact pass stages a body's
    first 8 colliders and reads the rest from memory).  Without joints or
    torques rotations stay exactly 0, so collider transforms are exact and
    the general path is bit-exact against the oracle."""
    rng = np.random.default_rng(12)
    n, per = 400, 12
    bpos = rng.uniform(0.0, 30.0, (n, 2)).astype(np.float32)
    bvel = rng.normal(0.0, 1.0, (n, 2)).astype(np.float32)
    ang = 2 * np.pi * np.arange(per) / per
    ring = np.stack([0.8 * np.cos(ang), 0.8 * np.sin(ang)], 1).astype(np.float32)
    parent = np.repeat(np.arange(n, dtype=np.int32), per)
    off = np.tile(ring, (n, 1))
    nc = n * per
    cr = rng.uniform(0.2, 0.4, nc).astype(np.float32)
    sc = scenes.Scene(
        pos=bpos, vreq=bvel, has_vreq=np.ones(n, np.uint8), gravity_mod=np.ones(n, np.float32),
        body_type=np.zeros(n, np.uint8), rotation=np.zeros(n, np.float32), col_parent=parent, col_off=off,
        col_abs=(bpos[parent] + off).astype(np.float32), col_r=cr, col_mass_override=np.zeros(nc, np.float32),
        col_has_mo=np.zeros(nc, np.uint8), col_sensor=np.zeros(nc, np.uint8),
        col_mem=np.full(nc, 0xFFFFFFFF, np.uint32), col_filt=np.full(nc, 0xFFFFFFFF, np.uint32),
        gravity=(0.0, -9.81), dt=1.0 / 60.0, steps=3, meta={})
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for step in range(3):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.step_path() == "general"
        assert p.contact_count() == o.contact_count
        assert_bitexact(p.bodies()["position"], o.pos, f"step {step} position")
    assert o.contact_count > n
    assert (p.bodies()["rotation"] == 0.0).all()


def test_joint_block_beyond_the_lds_staging():
    """A block of 64 joint components stages its joints and small components'
    bodies in LDS when they fit (512 joints, 640 bodies).  Here a chain of 600
    joints comes first, so the block that holds it and 40 small chains (1-15
    joints) exceeds the staging and every component of it iterates in global
    memory: positions bit-exact against the oracle."""
    rng = np.random.default_rng(12)
    pos, ja, jb = [], [], []
    row = 0.0
    for length in [600] + [1 + k % 15 for k in range(40)]:
        base = len(pos)
        pos += [(0.8 * k, row) for k in range(length + 1)]
        ja += list(range(base, base + length))
        jb += list(range(base + 1, base + length + 1))
        row += 3.0
    pos = np.array(pos, np.float32)
    n = pos.shape[0]
    vel = rng.normal(0, 2, (n, 2)).astype(np.float32)
    rad = np.full(n, 0.3, np.float32)
    sc = scenes._single_sphere_scene(pos, vel, rad, (0.0, -9.81), 1 / 60, 3, {})
    sc.joint_a = np.array(ja, np.int32)
    sc.joint_b = np.array(jb, np.int32)
    nj = sc.joint_a.shape[0]
    sc.joint_anchor_a = rng.uniform(-0.2, 0.2, (nj, 2)).astype(np.float32)
    sc.joint_anchor_b = rng.uniform(-0.2, 0.2, (nj, 2)).astype(np.float32)
    p = gpu_world(sc)
    o = OracleWorld(sc, np.float32)
    for step in range(3):
        p.step(sc.dt)
        o.step(sc.dt)
        assert p.step_path() == "general"
        assert_bitexact(p.bodies()["position"], o.pos, f"step {step} position")
    assert np.abs(p.bodies()["position"] - pos).max(axis=1).min() > 0
