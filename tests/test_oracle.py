"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

The reference is a Rust crate that cannot be built here and ships no
trajectory fixtures; its only test on this path is the body-transform check
in blobs/src/collider.rs:340-389 (a golden vector below).  Every other stage
is pinned by what the mathematics fixes -- closed forms, invariants,
brute force, and special cases -- chosen so that a dropped term, a wrong
sign, a wrong index or a transposed operand fails one of them.
"""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

from blobs_amd import scenes
from oracle import oracle
from oracle.oracle import OracleWorld

GOLDEN = Path(__file__).resolve().parent / "golden"


def one_body_scene(pos, vel=None, radius=0.5, gravity=(0.0, -9.81), static=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 2)
    n = pos.shape[0]
    vel = np.zeros((n, 2), np.float32) if vel is None else np.asarray(vel, np.float32).reshape(-1, 2)
    rad = np.broadcast_to(np.asarray(radius, np.float32), (n,)).copy()
    sc = scenes._single_sphere_scene(pos, vel, rad, gravity, 1 / 60, 1, {})
    if static is not None:
        sc.body_type = np.asarray(static, np.uint8)
    return sc


# ---------------------------------------------------------------- golden

def test_reference_transform_golden_vectors():
    """collider.rs:340-389: body (2,3) rotated pi/2 composed with offset
    (identity, (1,1)); compared with assert_relative_eq (f32 epsilon)."""
    kat = json.loads((GOLDEN / "reference_transform_kat.json").read_text())
    L = oracle.lib()
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    for case in kat["cases"]:
        pos = np.array(case["body_position"], np.float32)
        off_m = np.array(case["offset_matrix"], np.float32)
        off_t = np.array(case["offset_translation"], np.float32)
        out_m = np.zeros(4, np.float32)
        out_t = np.zeros(2, np.float32)
        L.oracle_body_transform_mul_f32(case["body_rotation"], f(pos), f(off_m), f(off_t), f(out_m), f(out_t))
        eps = np.finfo(np.float32).eps
        for got, want in ((out_m, case["expected_matrix"]), (out_t, case["expected_translation"])):
            want = np.array(want, np.float32)
            # approx::relative_eq with epsilon = max_relative = f32::EPSILON
            tol = np.maximum(eps, eps * np.maximum(np.abs(got), np.abs(want)))
            assert np.all(np.abs(got - want) <= tol), (case["name"], got, want)


# ---------------------------------------------------------- integration

def test_free_fall_closed_form_and_q1():
    """Position Verlet under constant g: x_k = x0 + k v h + g h^2 k(k+1)/2.
    Q1 (physics.rs:338-339): the first non-static body in arena order sees
    h / old_dt with old_dt = 1.0 on its first substep, so its request is
    scaled by h; a static body ahead of it does not take the role."""
    g = (0.0, -9.81)
    sc = one_body_scene([[0, 0], [5, 0], [10, 0]], [[1, 1], [2, 3], [-1, 0.5]], gravity=g, static=[1, 0, 0])
    o = OracleWorld(sc, np.float64, collisions_enabled=False)
    steps = 30
    o.step(1 / 60, steps)
    h = (1 / 60) / 8
    k = 8 * steps
    for b, scale in ((1, h), (2, 1.0)):
        v = sc.vreq[b].astype(np.float64) * scale
        pred = sc.pos[b] + k * h * v + np.array(g) * h * h * k * (k + 1) / 2
        assert np.allclose(o.pos[b], pred, rtol=0, atol=1e-9), (b, o.pos[b], pred)
    # static body: never moves, velocity 0, pos_old == pos
    assert np.array_equal(o.pos[0], sc.pos[0]) and np.array_equal(o.vel[0], [0, 0])
    assert np.array_equal(o.pos_old[0], o.pos[0])
    # f32 restatement agrees with the f64 twin to rounding
    o32 = OracleWorld(sc, np.float32, collisions_enabled=False)
    o32.step(1 / 60, steps)
    assert np.allclose(o32.pos, o.pos, atol=2e-4)


def test_calculated_velocity_is_displacement_over_h():
    sc = one_body_scene([[0, 0], [3, 0]], [[1, 2], [-2, 4]], gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64, collisions_enabled=False)
    o.step(1 / 60, 3)
    assert np.allclose(o.vel[1], sc.vreq[1], rtol=1e-12)
    assert np.allclose(o.pos[1] - o.pos_old[1], sc.vreq[1] * (1 / 480), rtol=1e-9)


# -------------------------------------------------------------- contacts

def test_equal_spheres_separate_to_contact_distance():
    """rho = 1/2 for equal masses: one contact pass moves each sphere by
    delta/2 along the axis, leaving them exactly touching (to rounding),
    and m_a dp_a + m_b dp_b = 0."""
    sc = one_body_scene([[0, 0], [0.6, 0.2]], radius=0.5, gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64)
    before = o.pos.copy()
    pairs = o.contact_pairs()
    assert pairs.tolist() == [[1, 0]]
    d = np.linalg.norm(o.pos[1] - o.pos[0])
    assert abs(d - 1.0) < 1e-12
    dp = o.pos - before
    assert np.allclose(dp[0] + dp[1], 0, atol=1e-15)
    # the higher arena index is 'A' and moves along +axis = A - B
    assert np.dot(dp[1], before[1] - before[0]) > 0


def test_unequal_masses_split_by_ratio():
    """ratio = 1 - m_A/(m_A+m_B) (physics.rs:319-321): |dp_A|/|dp_B| = m_B/m_A."""
    sc = one_body_scene([[0, 0], [1.0, 0]], radius=[0.375, 0.75], gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64)
    before = o.pos.copy()
    o.contact_pairs()
    dp = np.abs(o.pos - before)[:, 0]
    m = 4 * np.array([0.375, 0.75])  # Q2: 2 * (2 r)
    assert math.isclose(dp[1] / dp[0], m[0] / m[1], rel_tol=1e-12)
    assert math.isclose(dp.sum(), 0.375 + 0.75 - 1.0, rel_tol=1e-12)


def _random_multi_scene(seed=7, n=300):
    rng = np.random.default_rng(seed)
    ncol = rng.integers(1, 4, n)
    parent = np.repeat(np.arange(n, dtype=np.int32), ncol)
    c = parent.shape[0]
    pos = rng.uniform(0, 18, (n, 2)).astype(np.float32)
    off = rng.uniform(-0.6, 0.6, (c, 2)).astype(np.float32)
    sc = scenes.Scene(
        pos=pos, vreq=rng.normal(0, 1, (n, 2)).astype(np.float32), has_vreq=np.ones(n, np.uint8),
        gravity_mod=np.ones(n, np.float32), body_type=(rng.uniform(size=n) < 0.1).astype(np.uint8),
        rotation=np.zeros(n, np.float32), col_parent=parent, col_off=off, col_abs=(pos[parent] + off),
        col_r=rng.uniform(0.2, 0.6, c).astype(np.float32), col_mass_override=np.zeros(c, np.float32),
        col_has_mo=np.zeros(c, np.uint8), col_sensor=(rng.uniform(size=c) < 0.1).astype(np.uint8),
        col_mem=rng.choice(np.array([1, 2, 3, 0xFFFFFFFF], np.uint32), c),
        col_filt=rng.choice(np.array([1, 2, 0xFFFFFFFF], np.uint32), c))
    return sc


def _numpy_pairs(sc):
    """Independent brute force of the reference predicate (physics.rs:250-269):
    parents differ, groups test, |abs_i - abs_j| < r_i + r_j in f32."""
    a = sc.col_abs.astype(np.float32)
    i, j = np.tril_indices(sc.n_colliders, -1)
    dx = a[i, 0] - a[j, 0]
    dy = a[i, 1] - a[j, 1]
    dist = np.sqrt(dx * dx + dy * dy)
    ok = dist < (sc.col_r[i] + sc.col_r[j])
    ok &= sc.col_parent[i] != sc.col_parent[j]
    ok &= ((sc.col_mem[i] & sc.col_filt[j]) != 0) & ((sc.col_mem[j] & sc.col_filt[i]) != 0)
    return set(zip(i[ok].tolist(), j[ok].tolist()))


@pytest.mark.parametrize("use_grid", [0, 1])
def test_contact_pair_set_matches_bruteforce(use_grid):
    sc = _random_multi_scene()
    o = OracleWorld(sc, np.float32, use_grid=bool(use_grid))
    pairs = o.contact_pairs()
    got = [tuple(p) for p in pairs.tolist()]
    assert len(got) == len(set(got))
    assert set(got) == _numpy_pairs(sc)
    # visit order is lexicographic in (i, j<i)
    assert got == sorted(got)


def test_grid_enumeration_equals_reference_loop():
    """The grid variant (used for large scenes) visits the same (i, j)
    sequence as the O(C^2) loop: trajectories are bit-identical."""
    for sc in (scenes.plain_small_scene(n=400, steps=1), _random_multi_scene(11)):
        a = OracleWorld(sc, np.float32, use_grid=False)
        b = OracleWorld(sc, np.float32, use_grid=True)
        a.step(1 / 60, 15)
        b.step(1 / 60, 15)
        assert np.array_equal(a.pos.view(np.uint32), b.pos.view(np.uint32))
        assert a.contact_count == b.contact_count


def test_sensor_contacts_counted_not_resolved():
    sc = one_body_scene([[0, 0], [0.6, 0]], radius=0.5, gravity=(0.0, 0.0))
    sc.col_sensor[:] = [1, 0]
    o = OracleWorld(sc, np.float32)
    before = o.pos.copy()
    assert o.contact_pairs().tolist() == [[1, 0]]
    assert np.array_equal(o.pos, before)


def test_push_out_coincident_centres():
    """dist < 1e-6: +-0.01 along x, then the correction from the new axis."""
    sc = one_body_scene([[1, 2], [1, 2]], radius=0.5, gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64)
    o.contact_pairs()
    assert o.error_flags & oracle.INFO_PUSH_OUT
    assert np.allclose(o.pos, [[0.5, 2.0], [1.5, 2.0]], atol=1e-12)


def test_push_out_later_pairs_read_the_rewritten_snapshot():
    """Three coincident equal spheres (r = 0.1) at the origin, worked by hand
    through physics.rs:248-300 in (i, j<i) order:
      (1,0): push-out -> x1 = 0.01, x0 = -0.01; both snapshots are rewritten to
             those live positions; axis 0.02, delta 0.18, split 0.09 each
             -> x1 = 0.1, x0 = -0.1;
      (2,0): snapshot 0 now reads -0.01: axis 0.01, delta 0.19 -> x2 = 0.095,
             x0 = -0.195;
      (2,1): snapshot 1 reads 0.01: axis -0.01 -> x2 = 0, x1 = 0.195.
    A per-pair push-out (every pair +-0.01 from the unchanged snapshot) gives
    (-0.2, 0, 0.2) instead."""
    sc = one_body_scene([[0, 0], [0, 0], [0, 0]], radius=0.1, gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64)
    o.contact_pairs()
    assert np.allclose(o.pos, [[-0.195, 0.0], [0.195, 0.0], [0.0, 0.0]], atol=1e-12)
    assert o.w.push_outs == 1


def test_push_out_rewrites_with_the_unrotated_offset():
    """physics.rs:280-283 rebuilds a pushed collider's snapshot as
    position + offset.translation, without the body's rotation.  Body 0 sits
    at the origin rotated pi/2 with its collider offset (1, 0), so its
    snapshot is (0, 1) -- on top of body 1's collider.  Pair (1, 0), by hand:
    push -> body 1 at (0.01, 1), body 0 at (-0.01, 0); snapshots (0.01, 1) and
    (-0.01, 0) + (1, 0) = (0.99, 0); axis (-0.98, 1), rs = 2, equal masses
    split delta in half.  The rotated offset would give axis (0.02, 0)."""
    sc = scenes.Scene(
        pos=np.array([[0, 0], [0, 1]], np.float32), vreq=np.zeros((2, 2), np.float32),
        has_vreq=np.zeros(2, np.uint8), gravity_mod=np.ones(2, np.float32), body_type=np.zeros(2, np.uint8),
        rotation=np.array([math.pi / 2, 0.0], np.float32), col_parent=np.array([0, 1], np.int32),
        col_off=np.array([[1, 0], [0, 0]], np.float32), col_abs=np.array([[0, 1], [0, 1]], np.float32),
        col_r=np.array([1.0, 1.0], np.float32), col_mass_override=np.zeros(2, np.float32),
        col_has_mo=np.zeros(2, np.uint8), col_sensor=np.zeros(2, np.uint8), col_mem=np.full(2, 0xFFFFFFFF, np.uint32),
        col_filt=np.full(2, 0xFFFFFFFF, np.uint32), gravity=(0.0, 0.0), dt=1 / 60)
    o = OracleWorld(sc, np.float64)
    o.contact_pairs()
    axis = np.array([-0.98, 1.0])
    dist = math.hypot(*axis)
    half = 0.5 * (2.0 - dist) * axis / dist
    assert np.allclose(o.pos[1], np.array([0.01, 1.0]) + half, atol=1e-12)
    assert np.allclose(o.pos[0], np.array([-0.01, 0.0]) - half, atol=1e-12)


def test_push_out_cluster_grid_equals_bruteforce():
    """The grid variant of the oracle's contact pass re-bins pushed colliders
    and re-queries the row; on 20 coincident spheres inside a crowd it visits
    the reference loop's exact sequence (bit-identical f32 positions)."""
    rng = np.random.default_rng(7)
    crowd = rng.uniform(-1.5, 1.5, (80, 2))
    pos = np.concatenate([crowd[:30], np.full((20, 2), 0.25), crowd[30:]]).astype(np.float32)
    rad = rng.uniform(0.05, 0.2, 100).astype(np.float32)
    sc = one_body_scene(pos, radius=rad, gravity=(0.0, 0.0))
    a = OracleWorld(sc, np.float32, use_grid=True)
    b = OracleWorld(sc, np.float32, use_grid=False)
    pa, pb = a.contact_pairs(), b.contact_pairs()
    assert np.array_equal(pa, pb)
    assert np.array_equal(a.pos.view(np.uint32), b.pos.view(np.uint32))
    assert a.w.push_outs == b.w.push_outs > 1


# --------------------------------------------------------------- springs

def test_spring_two_body_oscillator_frequency():
    """F = k (L - rest) along the axis on bodies of mass 4r (Q2): the relative
    coordinate oscillates with omega^2 = k (1/m_a + 1/m_b)."""
    sc = one_body_scene([[0, 0], [2.1, 0]], radius=[0.25, 0.5], gravity=(0.0, 0.0))
    sc.has_vreq[:] = 0
    k = 40.0
    sc.spring_a = np.array([0], np.int32)
    sc.spring_b = np.array([1], np.int32)
    sc.spring_rest = np.array([2.0], np.float32)
    sc.spring_k = np.array([k], np.float32)
    sc.spring_c = np.array([0.0], np.float32)
    o = OracleWorld(sc, np.float64, collisions_enabled=False)
    m = 4 * np.array([0.25, 0.5])
    omega = math.sqrt(k * (1 / m[0] + 1 / m[1]))
    dt = 1 / 240
    L = []
    for _ in range(600):
        o.step(dt)
        L.append(o.pos[1, 0] - o.pos[0, 0])
    L = np.array(L) - 2.0
    crossings = np.nonzero((L[:-1] > 0) & (L[1:] <= 0))[0]
    period = (crossings[-1] - crossings[0]) * dt / (len(crossings) - 1)
    assert math.isclose(period, 2 * math.pi / omega, rel_tol=1e-2)
    # equal and opposite: the mass-weighted centre stays put
    cm = (m[0] * o.pos[0] + m[1] * o.pos[1]) / m.sum()
    cm0 = (m[0] * sc.pos[0] + m[1] * sc.pos[1]) / m.sum()
    assert np.allclose(cm, cm0, atol=1e-9)


def test_spring_damping_is_componentwise():
    """springs.rs:38-43: force = u * (k (L - rest) - (c (v_rel . u)) u)
    (scalar minus Vec2, then Vec2 * Vec2): with u = (1,0) the y force is 0
    and the x force is k(L - rest) - c v_rel.x."""
    sc = one_body_scene([[0, 0], [3, 1]], radius=0.5, gravity=(0.0, 0.0))
    sc.has_vreq[:] = 0
    sc.spring_a = np.array([0], np.int32)
    sc.spring_b = np.array([1], np.int32)
    sc.spring_rest = np.array([2.0], np.float32)
    sc.spring_k = np.array([10.0], np.float32)
    sc.spring_c = np.array([3.0], np.float32)
    o = OracleWorld(sc, np.float64, collisions_enabled=False, substeps=1)
    o.a["vel"][:] = [0, 0, 1, 0.5]  # calculated_velocity read by the spring
    h = 1 / 60
    o.step(h)
    # body 0 starts at rest: p' = p + acc h^2 (Q1 scales its zero displacement)
    acc = o.pos[0] / (h * h)
    d = np.array([3.0, 1.0])
    L = np.linalg.norm(d)
    u = d / L
    rv = np.array([0.0, 0.0]) - np.array([1.0, 0.5])
    damping = 3.0 * np.dot(rv, u) * u                 # Vec2
    force = u * (10.0 * (L - 2.0) - damping)          # f32 - Vec2, then Vec2 * Vec2
    assert np.allclose(acc, force / 2.0, rtol=1e-9)   # m_a = 2 * (2 * 0.5)
    textbook = u * (10.0 * (L - 2.0) - 3.0 * np.dot(rv, u))
    assert not np.allclose(acc, textbook / 2.0, rtol=1e-3)


# ---------------------------------------------------------------- joints

@pytest.mark.parametrize("dy", [0.375, -0.375])
def test_joint_projects_distance_and_rotates_by_half_pi(dy):
    """After the joint pass the anchor distance equals the rest distance
    (pos_old keeps the pre-Verlet position); the rotation correction is
    0.5 * (angle_b - angle_a - target) * h with angle_b - angle_a = -+pi by
    the sign of d.y (anchors unrotated, Q7)."""
    sc = one_body_scene([[0, 0], [1.0, dy]], radius=0.1, gravity=(0.0, 0.0))
    sc.has_vreq[:] = 0
    sc.joint_a = np.array([0], np.int32)
    sc.joint_b = np.array([1], np.int32)
    sc.joint_anchor_a = np.zeros((1, 2), np.float32)
    sc.joint_anchor_b = np.zeros((1, 2), np.float32)
    o = OracleWorld(sc, np.float64, collisions_enabled=False, substeps=1, joint_iterations=1)
    rest = float(o.a["joint_dist"][0])  # create_fixed_joint's f32 |pa - pb|
    assert abs(rest - math.hypot(1.0, dy)) < 1e-6
    # move body b away, then one joint pass must restore the rest distance
    o.a["pos"][2:] = [1.5, dy * 1.5]
    o.a["pos_old"][2:] = [1.5, dy * 1.5]
    o.step(1 / 60)
    d = np.linalg.norm(o.pos_old[1] - o.pos_old[0])
    assert abs(d - rest) < 1e-12
    # equal masses: each end moved half the error, centre fixed
    assert np.allclose(o.pos_old.mean(0), [0.75, dy * 0.75], atol=1e-12)
    h = 1 / 60
    expect = 0.5 * (-math.pi if dy > 0 else math.pi) * h
    assert math.isclose(o.rot[0], expect, rel_tol=1e-9)
    assert math.isclose(o.rot[1], -expect, rel_tol=1e-9)


def test_joint_static_branch_uses_inverse_mass_sum():
    """physics.rs:452-455: with a static, the other body moves by
    (1/m_a + 1/m_b) * correction (not the full correction)."""
    sc = one_body_scene([[0, 0], [2.0, 0]], radius=0.5, gravity=(0.0, 0.0), static=[1, 0])
    sc.has_vreq[:] = 0
    sc.joint_a = np.array([0], np.int32)
    sc.joint_b = np.array([1], np.int32)
    sc.joint_anchor_a = np.zeros((1, 2), np.float32)
    sc.joint_anchor_b = np.zeros((1, 2), np.float32)
    o = OracleWorld(sc, np.float64, collisions_enabled=False, substeps=1, joint_iterations=1)
    o.a["pos"][2] = 2.5  # stretch by 0.5
    o.a["pos_old"][2] = 2.5
    o.step(1 / 60)
    ims = 1 / 2.0 + 1 / 2.0
    assert math.isclose(o.pos_old[1, 0], 2.5 - ims * 0.5, rel_tol=1e-12)
    assert o.pos[0, 0] == 0.0


# ----------------------------------------------------------- constraints

def test_circle_constraint_projects_centre_onto_radius():
    sc = one_body_scene([[0, 0], [9, 1]], radius=0.5, gravity=(0.0, 0.0), static=[1, 0])
    sc.has_vreq[:] = 0
    sc.con_pos = np.array([[0.0, 0.0]], np.float32)
    sc.con_r = np.array([4.0], np.float32)
    o = OracleWorld(sc, np.float64, collisions_enabled=False, substeps=1)
    o.step(1 / 60)
    assert abs(np.linalg.norm(o.pos[1]) - 4.0) < 1e-12
    assert np.allclose(o.pos[1] / np.linalg.norm(o.pos[1]), np.array([9, 1]) / np.hypot(9, 1))


# ------------------------------------------------------------ mass setup

def test_mass_setup_double_push_closed_form():
    """Q2: every collider handle sits twice in rbd.colliders, so
    m = 2 * sum(2 r) and I = 2 * sum(0.5 m_c r^2 + m_c |offset|^2);
    sensors are skipped, mass_override replaces 2 r, zero becomes 1."""
    sc = scenes.Scene(
        pos=np.zeros((4, 2), np.float32), vreq=np.zeros((4, 2), np.float32), has_vreq=np.ones(4, np.uint8),
        gravity_mod=np.ones(4, np.float32), body_type=np.zeros(4, np.uint8), rotation=np.zeros(4, np.float32),
        col_parent=np.array([0, 1, 1, 2, 3], np.int32),
        col_off=np.array([[0, 0], [1, 0], [0, 2], [0, 0], [3, 4]], np.float32), col_abs=np.zeros((5, 2), np.float32),
        col_r=np.array([0.5, 0.25, 0.5, 0.75, 1.0], np.float32),
        col_mass_override=np.array([0, 0, 0, 0, 7.0], np.float32), col_has_mo=np.array([0, 0, 0, 0, 1], np.uint8),
        col_sensor=np.array([0, 0, 0, 1, 0], np.uint8), col_mem=np.full(5, 0xFFFFFFFF, np.uint32),
        col_filt=np.full(5, 0xFFFFFFFF, np.uint32))
    m, I, com = oracle.mass_setup(sc)
    assert m.tolist() == [2.0, 2 * (0.5 + 1.0), 1.0, 14.0]
    i1 = 2 * ((0.5 * 0.5 * 0.25 ** 2 + 0.5 * 1.0) + (0.5 * 1.0 * 0.5 ** 2 + 1.0 * 4.0))
    assert math.isclose(I[1], i1, rel_tol=1e-6)
    assert I[2] == 1.0  # sensor only: inertia 0 -> 1
    assert math.isclose(I[3], 2 * (0.5 * 7.0 * 1.0 + 7.0 * 25.0), rel_tol=1e-6)
    assert np.allclose(com[1], [(0.5 * 1.0) / 1.5, (1.0 * 2.0) / 1.5], rtol=1e-6)


# --------------------------------------------------------------- twins

def test_f32_restatement_tracks_fp64_twin():
    sc = scenes.plain_small_scene(n=512, steps=2)
    a = OracleWorld(sc, np.float32)
    b = OracleWorld(sc, np.float64)
    a.step(sc.dt, 2)
    b.step(sc.dt, 2)
    assert np.abs(a.pos - b.pos).max() < 5e-4


def test_fixed_step_runs_at_most_three_integrations():
    sc = one_body_scene([[0, 0]], [[1, 0]], gravity=(0.0, 0.0))
    o = OracleWorld(sc, np.float64, collisions_enabled=False)
    acc = C.c_double(0.0)
    o.fixed_step(0.1, acc)  # 6 frames due, max 3 run
    assert math.isclose(acc.value, 0.1 - 3 / 60, rel_tol=1e-12)
    assert math.isclose(o.w.time, 3 / 60, rel_tol=1e-12)
