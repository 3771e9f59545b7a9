"""The FP64 reference adjoint (oracle/diff_reference.py) on the CPU.

It is pinned two ways: its trajectory against the oracle's FP64 twin (the
C restatement of physics.rs:78-477 with free pair decisions), and its
autograd gradients against central finite differences in FP64 with every
substep's contact pair set held fixed (eps = 1e-3)."""
import numpy as np
import pytest

from blobs_amd import scenes
from oracle.diff_reference import RefWorlds
from oracle.oracle import OracleWorld

torch = pytest.importorskip("torch")


def small(n_worlds=2, n=48, n_springs=24, seed=55, packing=0.35):
    rng = scenes.Rng(seed)
    ws = [scenes.rollout_world(n, n_springs, rng, packing=packing) for _ in range(n_worlds)]
    targets = np.array([[w.meta["box_side"] * 0.3, w.meta["box_side"] * 0.6] for w in ws], np.float32)
    return ws, targets


def test_reference_trajectory_matches_oracle_fp64_twin():
    ws, targets = small()
    ref = RefWorlds(ws, targets)
    steps = 15
    p = ref.rollout(ref.vreq0, ref.k0, ws[0].dt, steps, f64_step=True).detach().numpy()
    for w, sc in enumerate(ws):
        o = OracleWorld(sc, np.float64)
        o.step(sc.dt, steps)
        b, e = ref.b0[w], ref.b0[w + 1]
        np.testing.assert_allclose(p[b:e], o.pos, rtol=0, atol=1e-10)


def test_reference_gradients_match_central_differences():
    ws, targets = small(n_worlds=1, n=32, n_springs=16, seed=56)
    ref = RefWorlds(ws, targets)
    steps, eps = 12, 1e-3
    rec = []
    L0, gv, gk = ref.gradients(ws[0].dt, steps, record=rec)
    assert sum(int(r.shape[0]) for r in rec) > 0, "the scene should have contacts"

    def loss_at(vreq, k):
        with torch.no_grad():
            return float(ref.loss(ref.rollout(vreq, k, ws[0].dt, steps, pairs_in=rec)))

    vr0, k0 = ref.vreq0.clone(), ref.k0.clone()
    rng = np.random.default_rng(0)
    for i in rng.choice(ref.vreq0.shape[0], 4, replace=False):
        for c in (0, 1):
            vp, vm = vr0.clone(), vr0.clone()
            vp[i, c] += eps
            vm[i, c] -= eps
            fd = (loss_at(vp, k0) - loss_at(vm, k0)) / (2 * eps)
            assert fd == pytest.approx(gv[i, c], rel=1e-4, abs=1e-8 * max(1.0, abs(L0)))
    for s in rng.choice(ref.k0.shape[0], 4, replace=False):
        kp, km = k0.clone(), k0.clone()
        kp[s] += eps
        km[s] -= eps
        fd = (loss_at(vr0, kp) - loss_at(vr0, km)) / (2 * eps)
        assert fd == pytest.approx(gk[s], rel=1e-4, abs=1e-8 * max(1.0, abs(L0)))


def test_tape_adjoint_equals_autograd_on_the_fp64_trajectory():
    """The explicit reverse pass (tape_adjoint) is the derivative autograd
    computes, when run along the same FP64 trajectory and pair sets."""
    ws, targets = small(n_worlds=2, n=40, n_springs=20, seed=57)
    ref = RefWorlds(ws, targets)
    steps = 10
    rec, tape = [], []
    orig = ref.pairs

    def spy(p):
        tape.append(p.detach().numpy().copy())
        return orig(p)

    ref.pairs = spy
    L, gv, gk = ref.gradients(ws[0].dt, steps, record=rec)
    ref.pairs = orig
    with torch.no_grad():
        fp = ref.rollout(ref.vreq0, ref.k0, ws[0].dt, steps, pairs_in=rec).numpy()
    tv, tk = ref.tape_adjoint(tape, [r.numpy() for r in rec], fp, ws[0].dt)
    assert np.linalg.norm(tv - gv) <= 1e-11 * np.linalg.norm(gv)
    assert np.linalg.norm(tk - gk) <= 1e-11 * np.linalg.norm(gk)


def featured(n_worlds=2, n=48, n_springs=24, n_joints=10, seed=71, packing=0.35):
    """Worlds with damped springs and fixed joints (anchors off the centre)."""
    rng = scenes.Rng(seed)
    ws = [scenes.rollout_world(n, n_springs, rng, packing=packing, damping=5.0, n_joints=n_joints, anchor=0.3)
          for _ in range(n_worlds)]
    targets = np.array([[w.meta["box_side"] * 0.3, w.meta["box_side"] * 0.6] for w in ws], np.float32)
    return ws, targets


def test_reference_with_damping_and_joints_matches_oracle_fp64_twin():
    ws, targets = featured()
    assert all((w.spring_c > 0).all() and w.joint_a.shape[0] == 10 for w in ws)
    ref = RefWorlds(ws, targets)
    steps = 12
    p = ref.rollout(ref.vreq0, ref.k0, ws[0].dt, steps, f64_step=True).detach().numpy()
    for w, sc in enumerate(ws):
        o = OracleWorld(sc, np.float64)
        o.step(sc.dt, steps)
        b, e = ref.b0[w], ref.b0[w + 1]
        np.testing.assert_allclose(p[b:e], o.pos, rtol=0, atol=1e-10)


def test_joint_levels_are_the_arena_order_sweep():
    """The FP64 model runs joints level by level; a sequential sweep in
    arena order gives the same positions to rounding."""
    ws, targets = featured(n_worlds=1, n=40, n_joints=16, seed=72)
    ref = RefWorlds(ws, targets)
    p = ref.p0.clone() + 0.05 * torch.sin(torch.arange(ref.p0.numel(), dtype=torch.float64)).reshape(ref.p0.shape)
    got = ref._joints(p.clone())
    seq = p.clone()
    m = ref.mass
    for _ in range(ref.joint_iterations):
        for q in range(ref.ja.shape[0]):
            a, b = int(ref.ja[q]), int(ref.jb[q])
            d = (seq[b] + ref.jan_b[q]) - (seq[a] + ref.jan_a[q])
            dist = float(d.norm())
            corr = (dist - float(ref.jdist[q])) * d / dist
            ima, imb = 1.0 / float(m[a]), 1.0 / float(m[b])
            seq[a] = seq[a] + ima / (ima + imb) * corr
            seq[b] = seq[b] - (1.0 - ima / (ima + imb)) * corr
    assert not bool(ref.static.any())
    np.testing.assert_allclose(got.numpy(), seq.numpy(), rtol=0, atol=1e-12)


def test_reference_gradients_with_damping_and_joints_match_central_differences():
    ws, targets = featured(n_worlds=1, n=32, n_springs=16, n_joints=8, seed=73)
    ref = RefWorlds(ws, targets)
    steps, eps = 10, 1e-3
    rec = []
    L0, gv, gk = ref.gradients(ws[0].dt, steps, record=rec)
    assert sum(int(r.shape[0]) for r in rec) > 0, "the scene should have contacts"

    def loss_at(vreq, k):
        with torch.no_grad():
            return float(ref.loss(ref.rollout(vreq, k, ws[0].dt, steps, pairs_in=rec)))

    vr0, k0 = ref.vreq0.clone(), ref.k0.clone()
    rng = np.random.default_rng(3)
    jb = set(ref.ja.tolist()) | set(ref.jb.tolist())
    picks = [i for i in rng.permutation(vr0.shape[0]) if int(i) in jb][:3]  # bodies on joints
    for i in picks:
        for c in (0, 1):
            vp, vm = vr0.clone(), vr0.clone()
            vp[i, c] += eps
            vm[i, c] -= eps
            fd = (loss_at(vp, k0) - loss_at(vm, k0)) / (2 * eps)
            assert fd == pytest.approx(gv[i, c], rel=1e-4, abs=1e-8 * max(1.0, abs(L0)))
    for s in rng.choice(k0.shape[0], 3, replace=False):
        kp, km = k0.clone(), k0.clone()
        kp[s] += eps
        km[s] -= eps
        fd = (loss_at(vr0, kp) - loss_at(vr0, km)) / (2 * eps)
        assert fd == pytest.approx(gk[s], rel=1e-4, abs=1e-8 * max(1.0, abs(L0)))
