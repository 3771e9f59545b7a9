"""The square-root-free contact test (sqrt_lt in blobs_amd/csrc/common.hpp)
restated in numpy: sqrtf(d2) < r exactly when float64(d2) < m*m with m the
midpoint of r and its float32 predecessor.  Checked here against numpy's
correctly rounded float32 square root on adversarial pairs; the device copy
is checked on the GPU by test_gpu_primitives.py."""
import numpy as np

from tests.sqrt_cases import cases, reference


def midpoint_test(d2, r):
    d2 = d2.astype(np.float32)
    r = r.astype(np.float32)
    out = np.zeros(d2.shape, bool)
    pos = r > 0
    inf = pos & ~np.isfinite(r)
    out[inf] = d2[inf] < r[inf]
    fin = pos & np.isfinite(r)
    p = (r[fin].view(np.uint32) - 1).view(np.float32)
    m = (p.astype(np.float64) + r[fin].astype(np.float64)) * 0.5
    with np.errstate(invalid="ignore"):
        out[fin] = d2[fin].astype(np.float64) < m * m
    return out


def test_midpoint_form_equals_sqrtf_comparison():
    d2, r = cases()
    ref = reference(d2, r)
    got = midpoint_test(d2, r)
    bad = np.flatnonzero(ref != got)
    assert bad.size == 0, [(float(d2[i]), float(r[i])) for i in bad[:5]]
    # the cases straddle the boundary: both answers occur near r*r
    assert 0.2 < ref.mean() < 0.8


def test_a_dropped_midpoint_fails():
    """Comparing against r*r instead of the midpoint square disagrees with
    sqrtf on the boundary cases (so the cases can tell the forms apart)."""
    d2, r = cases(n_random=0)
    fin = (r > 0) & np.isfinite(r)
    naive = d2[fin].astype(np.float64) < r[fin].astype(np.float64) ** 2
    assert np.any(naive != reference(d2[fin], r[fin]))
