"""Writes rollout_grads.npz: FP64 gradients of the differentiable rollout's
centroid loss (256 worlds x 2,048 spheres + 1,024 springs, step(1/120),
rollout_scenes() defaults) for sampled worlds, computed by the FP64 reverse
pass (oracle/diff_reference.py tape_adjoint) along the f32 trajectory that
the oracle's f32 restatement -- and the GPU, bit for bit -- follows.

* worlds 0 and 1 at 50 and 200 steps: d loss / d velocity_request [2048, 2]
  and d loss / d stiffness [1024], stored in float32, and the relative
  distance (vector norm) from them of the same reverse pass run in float32
  (`_f32rel_v`, `_f32rel_k`): how far an f32 adjoint stands from FP64 on
  this trajectory, whatever its summation order;
* worlds 3 and 97 at 200 steps: only the largest |gradient| entry, which
  lies beyond the f32 range (the fixed-pair derivative of the contact
  projection grows about 1e8-fold every 50 steps).
Each (world, horizon) costs about a minute of CPU, too long for a GPU test.
Run from the repository root:  python tests/golden/make_rollout_grads.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from blobs_amd import scenes  # noqa: E402
from oracle.diff_reference import RefWorlds, f32_tape  # noqa: E402


def grads(ws, targets, w, horizon, f32=False):
    ref = RefWorlds([ws[w]], targets[w:w + 1])
    tape, final = f32_tape([ws[w]], ws[w].dt, horizon)
    pairs = [ref.pairs_f32(t) for t in tape]
    g64 = ref.tape_adjoint(tape, pairs, final, ws[w].dt)
    if not f32:
        return g64
    return g64, ref.tape_adjoint(tape, pairs, final, ws[w].dt, dtype=np.float32)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ws, targets = scenes.rollout_scenes()
    out = {}
    for w in (0, 1):
        for horizon in (50, 200):
            (gv, gk), (fv, fk) = grads(ws, targets, w, horizon, f32=True)
            assert np.abs(gv).max() < 3e38 and np.abs(gk).max() < 3e38
            out[f"w{w}_h{horizon}_gv"] = gv.astype(np.float32)
            out[f"w{w}_h{horizon}_gk"] = gk.astype(np.float32)
            out[f"w{w}_h{horizon}_f32rel_v"] = np.array(rel(fv, gv))
            out[f"w{w}_h{horizon}_f32rel_k"] = np.array(rel(fk, gk))
            print(w, horizon, np.abs(gv).max(), rel(fv, gv), rel(fk, gk), flush=True)
    for w in (3, 97):
        gv, gk = grads(ws, targets, w, 200)
        out[f"w{w}_h200_maxabs"] = np.array(max(np.abs(gv).max(), np.abs(gk).max()), np.float64)
        print(w, 200, out[f"w{w}_h200_maxabs"], flush=True)
    np.savez_compressed(Path(__file__).resolve().parent / "rollout_grads.npz", **out)


if __name__ == "__main__":
    main()
