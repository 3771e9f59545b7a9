"""The persistent batched step on the GPU (batch_kernels.hip): worlds of
1,025 to 4,096 bodies, one workgroup per world from a largest-first queue,
every substep of a launch inside it with the world on chip (LDS +
registers); the batch's other worlds on the grid path in the same steps.

Bar: bit-identical to the same batch on the grid path and to each world
stepped alone (both pinned bit-exactly to the oracle elsewhere), and a
sampled world bit-exact against the oracle directly; contact counts equal.
"""
import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.physics import Physics
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def state(p):
    b = p.bodies()
    return ({k: b[k] for k in ("position", "position_old", "velocity", "rotation", "angular_velocity")},
            p.collider_positions())


def assert_state_equal(p, q, tag):
    (sa, ca), (sb, cb) = state(p), state(q)
    for k in sa:
        ga, gb = bits(sa[k]), bits(sb[k])
        bad = np.argwhere(ga != gb)
        assert bad.shape[0] == 0, f"{tag} {k}: {bad.shape[0]} of {ga.size} differ, first {bad[:3].tolist()}"
    assert np.array_equal(bits(ca), bits(cb)), f"{tag} collider translation"


def mixed_batch(seed=80):
    """Worlds of every size: tiny (below a wave), grid-path (up to 1,024 and
    beyond 4,096) and persistent (1,025 to 4,096), at both class edges."""
    rng = scenes.Rng(seed)
    sizes = [64, 100, 1024, 1025, 2048, 2049, 4096, 4097, 9000, 16384]
    packs = rng.uniform(len(sizes), 0.2, 0.5)
    return [scenes.uneven_batch_world(n, float(pk), rng) for n, pk in zip(sizes, packs)]


def test_persistent_equals_grid_path_and_worlds_alone():
    scs = mixed_batch()
    a = Physics.batch([sc.build_physics()[0] for sc in scs])
    a.set_persistent(True)
    g = Physics.batch([sc.build_physics()[0] for sc in scs])
    assert not g.persistent_active()  # the default: the grid path
    a.step(scs[0].dt, 1)
    g.step(scs[0].dt, 1)
    assert a.persistent_active() and not g.persistent_active()
    a.run_chunks(scs[0].dt, 10, 2)
    g.run_chunks(scs[0].dt, 10, 2)
    a.step(scs[0].dt, 3)
    g.step(scs[0].dt, 3)
    assert_state_equal(a, g, "persistent vs grid path")
    assert a.contact_count() == g.contact_count() > 0
    # each world alone, and the largest one against the oracle
    src = [sc.build_physics()[0] for sc in scs]
    a.writeback(src)
    for w in (0, 2, 3, 5, 6, 7, 9):
        alone = scs[w].build_physics()[0]
        alone.step(scs[w].dt, 24)
        assert_state_equal(src[w], alone, f"world {w} (n={scs[w].n_bodies})")
    o = OracleWorld(scs[6], np.float32)  # 4,096 bodies: the persistent kernel's largest
    o.step(scs[6].dt, 24)
    assert np.array_equal(bits(src[6].bodies()["position"]), bits(o.pos))


def test_persistent_user_inputs_static_bodies_and_q1():
    """Velocity requests, user accelerations and torques (consumed by the
    first substep), static bodies, and a world whose old_dt differs (Q1)."""
    rng = scenes.Rng(81)
    scs = [scenes.uneven_batch_world(n, 0.4, rng) for n in (300, 5000, 1200)]
    for sc in scs:
        sc.body_type[::7] = 1  # static
    built = [[sc.build_physics() for sc in scs] for _ in range(2)]
    for ws in built:
        ws[1][0].step(scs[1].dt * 0.5, 1)  # world 1's old_dt differs from the others'
        for p, bh in ws:
            b = p.get_rbd(int(bh[1]))
            b.apply_force((3.0, -2.0))
            p.set_rbd(int(bh[1]), b)
            b = p.get_rbd(int(bh[2]))
            b.apply_force_at_point((0.5, 4.0), (b.position[0] + 0.1, b.position[1] - 0.2))
            p.set_rbd(int(bh[2]), b)
            b = p.get_rbd(int(bh[3]))
            b.set_velocity((1.5, 0.25))
            p.set_rbd(int(bh[3]), b)
    src = [p for p, _ in built[0]]
    ref = [p for p, _ in built[1]]
    batch = Physics.batch(src)
    batch.set_persistent(True)
    batch.step(scs[0].dt, 6)
    assert batch.persistent_active()
    batch.writeback(src)
    for w, p in enumerate(ref):
        p.step(scs[w].dt, 6)
        assert_state_equal(src[w], p, f"world {w}")
    assert np.any(src[0].bodies()["rotation"] != 0.0)


def test_persistent_off_with_events_and_on_again():
    scs = mixed_batch(seed=82)[:4]
    b = Physics.batch([sc.build_physics()[0] for sc in scs])
    b.set_persistent(True)
    b.set_event_capacity(1 << 16)
    b.step(scs[0].dt, 2)
    assert not b.persistent_active()  # events need the grid path's per-substep record
    b.set_event_capacity(0)
    b.step(scs[0].dt, 2)
    assert b.persistent_active()
    q = Physics.batch([sc.build_physics()[0] for sc in scs])
    q.set_persistent(False)
    q.step(scs[0].dt, 4)
    assert_state_equal(b, q, "after switching")


def test_persistent_collisions_disabled():
    scs = mixed_batch(seed=83)[:5]
    ws = []
    for sc in scs:
        p = sc.build_physics()[0]
        p.collisions_enabled = False
        ws.append(p)
    b = Physics.batch(ws)
    b.set_persistent(True)
    b.step(scs[0].dt, 4)
    assert b.persistent_active() and b.contact_count() == 0
    for w, sc in enumerate(scs):
        o = OracleWorld(sc, np.float32, collisions_enabled=False)
        o.step(sc.dt, 4)
        lo = sum(s.n_bodies for s in scs[:w])
        assert np.array_equal(bits(b.bodies()["position"][lo:lo + sc.n_bodies]), bits(o.pos))
