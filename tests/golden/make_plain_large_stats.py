"""Writes plain_large_stats.json for the plain large scene (1,048,576
single spheres, r U[0.5, 1.0], packing 0.4, step(1/120) x 100), every 10
steps:
* SHA-256 digests of the f32 state (positions, rotations, collider
  translations) along the oracle's f32 restatement: the GPU trajectory must
  reproduce them bit for bit over the scene's whole horizon;
* penetration and energy statistics along that f32 run and along the
  oracle's FP64 twin (the error model of the f32 step).
Each run takes about 12 minutes of one CPU core -- too long to repeat inside
a GPU test -- so the outputs are kept here as a fixture.  Run from the
repository root:  python tests/golden/make_plain_large_stats.py
"""
import json
import multiprocessing as mp
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from blobs_amd import scenes  # noqa: E402
from oracle.oracle import OracleWorld, mass_setup  # noqa: E402
from tests.world_stats import NAMES, state_digest, world_stats  # noqa: E402

STEPS, EVERY = 100, 10


def run(kind):
    sc = scenes.plain_large_scene(steps=STEPS)
    o = OracleWorld(sc, np.float32 if kind == "f32" else np.float64)
    mass = mass_setup(sc)[0]
    dyn = sc.body_type == 0
    stats, digests = [], []
    t0 = time.time()
    for k in range(STEPS // EVERY):
        o.step(sc.dt, EVERY)
        if kind == "f32":
            digests.append(state_digest(o.pos, o.rot, o.abs_t))
        stats.append(world_stats(o.abs_t, sc.col_r, sc.col_parent, o.pos[:, 1], o.vel, mass, dyn,
                                 sc.gravity[1]).tolist())
        print(f"{kind}: step {(k + 1) * EVERY} ({time.time() - t0:.0f} s)", flush=True)
    return stats, digests


def main():
    with mp.Pool(2) as pool:
        (f32, dig), (f64, _) = pool.map(run, ["f32", "fp64"])
    sc = scenes.plain_large_scene(steps=STEPS)
    doc = {
        "scene": "plain_large_scene(steps=100) defaults: 1048576 single spheres, r U[0.5,1.0], packing 0.4, "
                 "step(1/120)",
        "n_bodies": int(sc.n_bodies),
        "statistics": NAMES,
        "steps": [(k + 1) * EVERY for k in range(STEPS // EVERY)],
        "f32_sha256": dig,
        "f32": f32,
        "fp64": f64,
    }
    (Path(__file__).resolve().parent / "plain_large_stats.json").write_text(json.dumps(doc, indent=1))


if __name__ == "__main__":
    main()
