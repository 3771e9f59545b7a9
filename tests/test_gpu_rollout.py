"""Differentiable rollouts on the GPU: forward parity and adjoint gradients.

Bars:
  * forward: final positions bit-exact against the oracle's f32 restatement
    (the step is the same f32 arithmetic as the general path);
  * adjoint, exactness: gradients of the centroid loss w.r.t. initial
    velocity requests and spring stiffnesses within 1e-4 relative (vector
    norm) of an FP64 reverse pass along the same f32 trajectory
    (oracle/diff_reference.py tape_adjoint, pinned to autograd in
    tests/test_diff_reference.py);
  * adjoint, end to end: within 1e-3 relative of the FP64 model's
    gradients (autograd, pinned to central differences) and of FP64 central
    differences (eps 1e-3, pair sets held fixed) directly.  The f32 step
    itself limits this: position Verlet forms velocities as p - p_old in
    f32, so f32 and FP64 trajectories drift apart over contact-rich
    horizons whatever the adjoint does;
  * damped springs and fixed joints: forward bit-exact against the oracle
    and the general path; gradients within 1e-4 (velocity requests) and
    3e-3 (stiffnesses) of the FP64 model run on the f32 trajectory's pair
    sets, and within 1e-3 of its central differences;
  * checkpoint interval: gradients bit-identical for any interval;
  * autograd.Function: torch back-propagation through the rollout equals
    the direct entry point.
"""
import numpy as np
import pytest

from blobs_amd import scenes
from blobs_amd.rollout import Rollout, rollout_function
from oracle.diff_reference import RefWorlds, f32_tape
from oracle.oracle import OracleWorld

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def worlds(n_worlds=2, n=64, n_springs=32, seed=57, packing=0.3):
    rng = scenes.Rng(seed)
    ws = [scenes.rollout_world(n, n_springs, rng, packing=packing) for _ in range(n_worlds)]
    targets = np.array([[w.meta["box_side"] * 0.25, w.meta["box_side"] * 0.75] for w in ws], np.float32)
    return ws, targets


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_rollout_forward_bitexact_vs_oracle():
    ws, targets = worlds()
    steps = 25
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=10)
    loss = r.forward(ws[0].dt, steps)
    pos = r.final_positions()
    b0 = 0
    for w, sc in enumerate(ws):
        o = OracleWorld(sc, np.float32)
        o.step(sc.dt, steps)
        got = pos[b0:b0 + sc.n_bodies]
        bad = np.argwhere(bits(got) != bits(o.pos))
        assert bad.shape[0] == 0, f"world {w}: {bad.shape[0]} coordinates differ, first {bad[:3].tolist()}"
        cen = o.pos.astype(np.float64).mean(0)
        assert loss[w] == pytest.approx(float(((cen - targets[w]) ** 2).sum()), rel=1e-6)
        b0 += sc.n_bodies


def test_rollout_forward_matches_general_path():
    ws, targets = worlds(n_worlds=1, seed=58)
    steps = 20
    p = ws[0].build_physics()[0]
    r = Rollout([p], targets, checkpoint_every=20)
    r.forward(ws[0].dt, steps)
    p.step(ws[0].dt, steps)
    assert np.array_equal(bits(r.final_positions()), bits(p.bodies()["position"]))


@pytest.mark.parametrize("steps,ck", [(20, 5), (30, 30)])
def test_rollout_adjoint_exact_on_its_own_trajectory(steps, ck):
    """The GPU adjoint against an FP64 reverse pass along the GPU's own f32
    trajectory (the oracle's f32 restatement, one substep at a time, with
    the f32 pair predicate): only the adjoint's f32 arithmetic differs."""
    ws, targets = worlds(seed=59)
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=ck)
    r.forward(ws[0].dt, steps)
    gv, gk = r.backward()
    ref = RefWorlds(ws, targets)
    tape, fin = f32_tape(ws, ws[0].dt, steps)
    assert np.array_equal(bits(fin), bits(r.final_positions())), "the GPU forward is the oracle's f32 trajectory"
    pairs = [ref.pairs_f32(t) for t in tape]
    assert sum(len(p) for p in pairs) > 0, "the scene should have contacts"
    tv, tk = ref.tape_adjoint(tape, pairs, fin, ws[0].dt)
    assert rel(gv, tv) <= 1e-4, f"velocity-request gradient off by {rel(gv, tv):.2e}"
    assert rel(gk, tk) <= 1e-4, f"stiffness gradient off by {rel(gk, tk):.2e}"


def test_rollout_gradients_match_fp64_central_differences():
    """End to end: GPU gradients (f32 forward + adjoint) against the FP64
    model -- autograd over all parameters and central differences
    (eps = 1e-3, every substep's pair set held fixed) on a sample -- within
    1e-3 relative."""
    ws, targets = worlds(seed=59)
    steps = 20
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=5)
    r.forward(ws[0].dt, steps)
    gv, gk = r.backward()
    ref = RefWorlds(ws, targets)
    rec = []
    _, rv, rk = ref.gradients(ws[0].dt, steps, record=rec)
    assert rel(gv, rv) <= 1e-3, f"velocity-request gradient off by {rel(gv, rv):.2e}"
    assert rel(gk, rk) <= 1e-3, f"stiffness gradient off by {rel(gk, rk):.2e}"
    eps = 1e-3

    def loss_at(vreq, k):
        with torch.no_grad():
            return float(ref.loss(ref.rollout(vreq, k, ws[0].dt, steps, pairs_in=rec)))

    v0, k0 = ref.vreq0.clone(), ref.k0.clone()
    rng = np.random.default_rng(1)
    for i in rng.choice(v0.shape[0], 3, replace=False):
        for c in (0, 1):
            vp, vm = v0.clone(), v0.clone()
            vp[i, c] += eps
            vm[i, c] -= eps
            fd = (loss_at(vp, k0) - loss_at(vm, k0)) / (2 * eps)
            assert abs(gv[i, c] - fd) <= 1e-3 * np.abs(rv).max(), (i, c, gv[i, c], fd)
    for q in rng.choice(k0.shape[0], 4, replace=False):
        kp, km = k0.clone(), k0.clone()
        kp[q] += eps
        km[q] -= eps
        fd = (loss_at(v0, kp) - loss_at(v0, km)) / (2 * eps)
        assert abs(gk[q] - fd) <= 1e-3 * np.abs(rk).max(), (q, gk[q], fd)


def test_gradients_independent_of_checkpoint_interval():
    ws, targets = worlds(seed=60)
    steps = 24
    out = []
    for ck in (1, 7, 24):
        r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=ck)
        r.forward(ws[0].dt, steps)
        out.append(r.backward())
    for gv, gk in out[1:]:
        assert np.array_equal(bits(gv), bits(out[0][0]))
        assert np.array_equal(bits(gk), bits(out[0][1]))


def test_autograd_function_matches_entry_point():
    ws, targets = worlds(seed=61)
    steps = 16
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=8)
    f = rollout_function()
    vreq = torch.tensor(np.concatenate([w.vreq for w in ws]), requires_grad=True)
    k = torch.tensor(np.concatenate([w.spring_k for w in ws]), requires_grad=True)
    pos = f.apply(vreq, k, r, ws[0].dt, steps)
    # an arbitrary loss on the final positions
    wgt = torch.linspace(-1.0, 1.0, pos.numel()).reshape(pos.shape)
    loss = (pos * wgt).sum() + (pos ** 2).mean()
    loss.backward()
    dl = (wgt + 2.0 * pos.detach() / pos.numel()).numpy()
    gv, gk = r.backward(dl)
    assert np.allclose(vreq.grad.numpy(), gv, rtol=1e-6, atol=1e-9)
    assert np.allclose(k.grad.numpy(), gk, rtol=1e-6, atol=1e-9)


def test_set_params_changes_the_rollout():
    ws, targets = worlds(n_worlds=1, seed=62)
    r = Rollout([ws[0].build_physics()[0]], targets, checkpoint_every=10)
    l0 = r.forward(ws[0].dt, 10)
    v = ws[0].vreq * 1.5
    r.set_params(velocity_request=v)
    l1 = r.forward(ws[0].dt, 10)
    assert l1[0] != l0[0]
    sc = ws[0]
    sc.vreq = v.astype(np.float32)
    o = OracleWorld(sc, np.float32)
    o.step(sc.dt, 10)
    assert np.array_equal(bits(r.final_positions()), bits(o.pos))


def featured(n_worlds=2, n=64, n_springs=32, n_joints=16, seed=74, packing=0.3):
    """Worlds with damped springs (c ~ U[0, 5]) and fixed joints (anchors
    U[-0.3, 0.3]^2)."""
    rng = scenes.Rng(seed)
    ws = [scenes.rollout_world(n, n_springs, rng, packing=packing, damping=5.0, n_joints=n_joints, anchor=0.3)
          for _ in range(n_worlds)]
    targets = np.array([[w.meta["box_side"] * 0.25, w.meta["box_side"] * 0.75] for w in ws], np.float32)
    return ws, targets


def test_rollout_with_damping_and_joints_forward_bitexact():
    """Forward of worlds with damped springs and joints: bit-exact against
    the oracle's f32 restatement and against the world's own general path."""
    ws, targets = featured()
    steps = 25
    phys = [w.build_physics()[0] for w in ws]
    r = Rollout(phys, targets, checkpoint_every=10)
    r.forward(ws[0].dt, steps)
    pos = r.final_positions()
    b0 = 0
    for w, sc in enumerate(ws):
        o = OracleWorld(sc, np.float32)
        o.step(sc.dt, steps)
        got = pos[b0:b0 + sc.n_bodies]
        bad = np.argwhere(bits(got) != bits(o.pos))
        assert bad.shape[0] == 0, f"world {w}: {bad.shape[0]} coordinates differ, first {bad[:3].tolist()}"
        phys[w].step(sc.dt, steps)
        assert phys[w].step_path() == "general"
        assert np.array_equal(bits(got), bits(phys[w].bodies()["position"]))
        b0 += sc.n_bodies


def test_rollout_with_damping_and_joints_gradients_match_fp64():
    """Adjoint of damped springs and joint positions: GPU gradients against
    the FP64 model (autograd, pinned to central differences on the CPU) run
    with the pair sets of the f32 trajectory the GPU follows (the FP64
    trajectory's own pair sets differ in 2 of the 128 substeps here, which
    alone moves the gradients by 0.4%), and against FP64 central differences
    with those pair sets on bodies that carry joints.  Bars: 1e-4 relative
    for velocity requests; 3e-3 for stiffnesses, whose gradient holds
    L - rest with rest the initial length, a difference the f32 trajectory
    resolves to a few ulps of L."""
    ws, targets = featured(seed=75)
    steps = 16
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=5)
    r.forward(ws[0].dt, steps)
    gv, gk = r.backward()
    ref = RefWorlds(ws, targets)
    tape, fin = f32_tape(ws, ws[0].dt, steps)
    assert np.array_equal(bits(fin), bits(r.final_positions()))
    rec = [torch.as_tensor(ref.pairs_f32(t)) for t in tape]
    assert sum(int(p.shape[0]) for p in rec) > 0, "the scene should have contacts"
    vr = ref.vreq0.clone().requires_grad_(True)
    kr = ref.k0.clone().requires_grad_(True)
    L = ref.loss(ref.rollout(vr, kr, ws[0].dt, steps, pairs_in=rec))
    rv, rk = [g.numpy() for g in torch.autograd.grad(L, [vr, kr])]
    assert rel(gv, rv) <= 1e-4, f"velocity-request gradient off by {rel(gv, rv):.2e}"
    assert rel(gk, rk) <= 3e-3, f"stiffness gradient off by {rel(gk, rk):.2e}"
    # the joints and the damping each matter: without the joints the FP64 gradients move by
    # 16% (v) and 73% (k), without the damping by 0.75% (v) and 13% (k) -- beyond the bars
    for drop, bv, bk in (("joints", 1e-2, 1e-1), ("damping", 2e-3, 3e-2)):
        cut = RefWorlds(ws, targets)
        if drop == "joints":
            cut.ja = cut.ja[:0]
        else:
            cut.c = torch.zeros_like(cut.c)
        vc = cut.vreq0.clone().requires_grad_(True)
        kc = cut.k0.clone().requires_grad_(True)
        Lc = cut.loss(cut.rollout(vc, kc, ws[0].dt, steps, pairs_in=rec))
        cv, ck = [g.numpy() for g in torch.autograd.grad(Lc, [vc, kc])]
        assert rel(cv, rv) > bv and rel(ck, rk) > bk, drop
    eps = 1e-3

    def loss_at(vreq, k):
        with torch.no_grad():
            return float(ref.loss(ref.rollout(vreq, k, ws[0].dt, steps, pairs_in=rec)))

    v0, k0 = ref.vreq0.clone(), ref.k0.clone()
    rng = np.random.default_rng(2)
    on_joint = sorted(set(ref.ja.tolist()) | set(ref.jb.tolist()))
    for i in rng.choice(on_joint, 3, replace=False):
        c = int(rng.integers(2))
        vp, vm = v0.clone(), v0.clone()
        vp[i, c] += eps
        vm[i, c] -= eps
        fd = (loss_at(vp, k0) - loss_at(vm, k0)) / (2 * eps)
        assert abs(gv[i, c] - fd) <= 1e-3 * np.abs(rv).max(), (i, c, gv[i, c], fd)
    for q in rng.choice(k0.shape[0], 3, replace=False):
        kp, km = k0.clone(), k0.clone()
        kp[q] += eps
        km[q] -= eps
        fd = (loss_at(v0, kp) - loss_at(v0, km)) / (2 * eps)
        assert abs(gk[q] - fd) <= 1e-3 * np.abs(rk).max(), (q, gk[q], fd)


def test_rollout_with_damping_and_joints_independent_of_checkpoint_interval():
    ws, targets = featured(seed=76)
    steps = 12
    out = []
    for ck in (1, 5, 12):
        r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=ck)
        r.forward(ws[0].dt, steps)
        out.append(r.backward())
    for gv, gk in out[1:]:
        assert np.array_equal(bits(gv), bits(out[0][0]))
        assert np.array_equal(bits(gk), bits(out[0][1]))
