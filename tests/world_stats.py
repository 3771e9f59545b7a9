"""Penetration and energy statistics of a world state (test helper).

Shared by the GPU statistics tests and by the script that writes their FP64
fixtures (tests/golden/make_multi_collider_stats.py), so both sides measure
with the same code.
"""
import hashlib

import numpy as np
from scipy.spatial import cKDTree

NAMES = ["max penetration", "mean penetration", "overlapping pairs", "kinetic energy", "potential energy"]


def world_stats(col_xy, col_r, col_parent, body_y, body_vel, mass, dyn, gy):
    """[max penetration, mean penetration of overlapping pairs, overlapping
    pairs, kinetic energy, potential energy] (float64)."""
    x = np.asarray(col_xy, np.float64)
    r = np.asarray(col_r, np.float64)
    ok = np.isfinite(x).all(1)
    idx = np.flatnonzero(ok)
    tree = cKDTree(x[idx])
    pairs = tree.query_pairs(2.0 * float(r.max()), output_type="ndarray")
    a, b = idx[pairs[:, 0]], idx[pairs[:, 1]]
    keep = col_parent[a] != col_parent[b]
    a, b = a[keep], b[keep]
    d = np.sqrt(((x[a] - x[b]) ** 2).sum(1))
    pen = (r[a] + r[b]) - d
    pen = pen[pen > 0]
    v = np.asarray(body_vel, np.float64)
    m = np.asarray(mass, np.float64)
    ke = 0.5 * (m[dyn] * (v[dyn] ** 2).sum(1)).sum()
    pe = -(m[dyn] * gy * np.asarray(body_y, np.float64)[dyn]).sum()
    return np.array([pen.max() if pen.size else 0.0, pen.mean() if pen.size else 0.0, float(pen.size), ke, pe])


def state_digest(pos, rot, col_xy):
    """SHA-256 of the f32 bytes of positions, rotations and collider translations."""
    h = hashlib.sha256()
    for a in (pos, rot, col_xy):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    return h.hexdigest()
