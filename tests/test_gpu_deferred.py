"""Deferred checks (bg_set_deferred_checks): bg_step returns once its kernels
are queued, the step's panic flags come back behind them and are reported
by a later call.  Same results as the synchronous mode, bit for bit; a
reference panic (here physics.rs:447-448, a joint on a body whose mass is
not positive) surfaces at bg_synchronize instead of at the step."""
import numpy as np
import pytest

from blobs_amd import _lib as L
from blobs_amd import scenes
from blobs_amd.physics import ColliderBuilder, Physics, RigidBodyBuilder

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("general", [False, True])
def test_deferred_checks_give_the_same_trajectory(general):
    sc = scenes.plain_small_scene(steps=40) if not general else scenes.multi_collider_scene(
        n=1024, n_springs=256, n_joints=0, steps=40, seed=31, phi=0.3)
    p, _ = sc.build_physics()
    q, _ = sc.build_physics()
    q.set_deferred_checks(True)
    for _ in range(40):
        p.step(sc.dt)
        q.step(sc.dt)
    q.synchronize()
    assert np.array_equal(bits(p.bodies()["position"]), bits(q.bodies()["position"]))
    assert p.contact_count() == q.contact_count()


def _joint_on_massless_body(deferred):
    p = Physics((0.0, -9.81), False)
    a = p.insert_rbd(RigidBodyBuilder().position((0.0, 0.0)).build())
    p.insert_collider_with_parent(ColliderBuilder().radius(0.5).mass_override(-1.0)
                                  .absolute_transform_translation((0.0, 0.0)).build(), a)
    b = p.insert_rbd(RigidBodyBuilder().position((3.0, 0.0)).build())
    p.insert_collider_with_parent(ColliderBuilder().radius(0.5).absolute_transform_translation((3.0, 0.0)).build(), b)
    p.create_fixed_joint(a, b, (0.0, 0.0), (0.0, 0.0))
    if deferred:
        p.set_deferred_checks(True)
    return p


def test_panic_is_reported_by_the_step_in_synchronous_mode():
    p = _joint_on_massless_body(False)
    with pytest.raises(L.BlobsError) as e:
        p.step(1 / 60)
    assert e.value.code == L.BG_ERR_PANIC and "447" in str(e.value)


def test_panic_is_reported_later_in_deferred_mode():
    p = _joint_on_massless_body(True)
    p.step(1 / 60)  # queued; the flags are still on their way
    with pytest.raises(L.BlobsError) as e:
        p.synchronize()
    assert e.value.code == L.BG_ERR_PANIC and "447" in str(e.value)
    p.synchronize()  # reported once
