"""Batched worlds, hipGraph chunks and the autotuner on the GPU.

Bar: every world of a batch evolves bit-identically to the same world
stepped on its own (which the parity tests pin bit-exactly to the oracle),
and a small batched world is compared with the oracle directly.  Chunked
graph replay and autotuned broadphase settings must not change any bit
either.
"""
import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.physics import Physics
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_same(a, b, what):
    ga, gb = bits(a), bits(b)
    bad = np.argwhere(ga != gb)
    assert bad.shape[0] == 0, f"{what}: {bad.shape[0]} of {ga.size} differ; first at {bad[:3].tolist()}"


def state(p):
    b = p.bodies()
    return {k: b[k] for k in ("position", "position_old", "velocity", "rotation")}, p.collider_positions()


def assert_state_equal(p, q, tag):
    (sa, ca), (sb, cb) = state(p), state(q)
    for k in sa:
        assert_same(sa[k], sb[k], f"{tag} {k}")
    assert_same(ca, cb, f"{tag} collider translation")


def small_batch(n_worlds=12, seed=44, cap=1500):
    return scenes.uneven_batch_scenes(n_worlds, lo=64, hi=4096, seed=seed, size_cap=cap)


def test_batch_matches_worlds_stepped_alone():
    scs = small_batch()
    steps = 30
    src = [sc.build_physics()[0] for sc in scs]
    batch = Physics.batch(src)
    assert batch.batch_size() == len(scs)
    assert batch.rbd_count() == sum(sc.n_bodies for sc in scs)
    batch.run_chunks(scs[0].dt, 10, steps // 10)
    batch.writeback(src)
    for k, sc in enumerate(scs):
        alone = sc.build_physics()[0]
        alone.step(sc.dt, steps)
        assert_state_equal(src[k], alone, f"world {k} (n={sc.n_bodies})")
        assert src[k].time == pytest.approx(alone.time, abs=0.0)


def test_batch_world_matches_oracle():
    scs = small_batch(n_worlds=3, seed=45, cap=300)
    steps = 12
    src = [sc.build_physics()[0] for sc in scs]
    batch = Physics.batch(src)
    batch.step(scs[0].dt, steps)
    batch.writeback(src)
    for k, sc in enumerate(scs):
        o = OracleWorld(sc, np.float32)
        o.step(sc.dt, steps)
        b = src[k].bodies()
        assert_same(b["position"], o.pos, f"world {k} position")
        assert_same(b["velocity"], o.vel, f"world {k} velocity")
        assert_same(src[k].collider_positions(), o.abs_t, f"world {k} collider translation")


def test_batch_sub_world_q1_old_dt_is_per_world():
    """Each world's first dynamic body sees h/old_dt with its own old_dt (Q1):
    a world whose old_dt differs must still match its lone run."""
    scs = small_batch(n_worlds=4, seed=46, cap=400)
    src = [sc.build_physics()[0] for sc in scs]
    src[2].step(scs[2].dt * 0.5, 1)  # moves world 2's old_dt away from the others
    alone2 = scs[2].build_physics()[0]
    alone2.step(scs[2].dt * 0.5, 1)
    batch = Physics.batch(src)
    batch.step(scs[0].dt, 5)
    batch.writeback(src)
    alone2.step(scs[2].dt, 5)
    assert_state_equal(src[2], alone2, "world 2")
    alone0 = scs[0].build_physics()[0]
    alone0.step(scs[0].dt, 5)
    assert_state_equal(src[0], alone0, "world 0")


@pytest.mark.parametrize("chunk,n_chunks,interval", [(10, 3, 1), (3, 5, 2), (1, 7, 3), (4, 2, 4)])
def test_chunks_equal_steps(chunk, n_chunks, interval):
    """Graph replay (cyclic or not, any remainder) leaves the same state as
    bg_step, and stepping on afterwards stays identical."""
    sc = scenes.plain_small_scene(steps=100)
    p = sc.build_physics()[0]
    q = sc.build_physics()[0]
    for w in (p, q):
        w.set_broadphase(1.0, interval, 0.5)
    p.run_chunks(sc.dt, chunk, n_chunks)
    q.step(sc.dt, chunk * n_chunks)
    assert_state_equal(p, q, "after chunks")
    assert p.contact_count() == q.contact_count()
    assert p.time == q.time
    p.step(sc.dt, 3)
    q.step(sc.dt, 3)
    assert_state_equal(p, q, "after more steps")


def test_batch_chunks_1m_scale_small_worlds():
    """A batch of many tiny worlds (several per wave) through graph chunks."""
    scs = scenes.uneven_batch_scenes(64, lo=64, hi=128, seed=47)
    src = [sc.build_physics()[0] for sc in scs]
    batch = Physics.batch(src)
    batch.run_chunks(scs[0].dt, 5, 2)
    batch.writeback(src)
    for k in (0, 17, 63):
        alone = scs[k].build_physics()[0]
        alone.step(scs[k].dt, 10)
        assert_state_equal(src[k], alone, f"world {k}")


def test_autotune_keeps_results():
    sc = scenes.plain_small_scene(steps=100)
    p = sc.build_physics()[0]
    cs, sk, iv = p.autotune(sc.dt, 2)
    assert cs >= 0.25 and 0.0 < sk <= 4.0 and iv >= 1
    n = int(round(p.time / sc.dt))
    q = sc.build_physics()[0]
    q.step(sc.dt, n)
    assert_state_equal(p, q, "after autotune")
