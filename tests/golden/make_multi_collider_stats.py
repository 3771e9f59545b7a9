"""Writes multi_collider_stats.json for the full multi-collider scene
(65,536 bodies with 1-8 colliders, 32,768 springs, 16,384 fixed joints,
step(1/60) x 300), every 25 steps:
* penetration and energy statistics along the oracle's FP64 twin (the error
  model the GPU trajectory's statistics are held to) and along its f32
  restatement on glibc's libm;
* SHA-256 digests of the f32 state (positions, rotations, collider
  translations) along the f32 restatement whose sinf/cosf/atan2f are
  evaluated in double and rounded, as the device does: the GPU trajectory
  must reproduce them bit for bit over the whole horizon.
Each run takes about a quarter of an hour of one CPU core -- too long to
repeat inside a GPU test -- so the outputs are kept here as a fixture.  Run
from the repository root:  python tests/golden/make_multi_collider_stats.py
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

STEPS, EVERY = 300, 25


def run(kind):
    dtype_name, libm = {"fp64": ("float64", "glibc"), "f32": ("float32", "glibc"),
                        "f32_cr": ("float32", "cr")}[kind]
    sc = scenes.multi_collider_scene()
    o = OracleWorld(sc, np.dtype(dtype_name), libm=libm)
    mass = mass_setup(sc)[0]
    dyn = sc.body_type == 0
    out = []
    t0 = time.time()
    for k in range(STEPS // EVERY):
        o.step(sc.dt, EVERY)
        if kind == "f32_cr":
            out.append(state_digest(o.pos, o.rot, o.abs_t))
        else:
            out.append(world_stats(o.abs_t, sc.col_r, sc.col_parent, o.pos[:, 1], o.vel, mass, dyn,
                                   sc.gravity[1]).tolist())
        print(f"{kind}: step {(k + 1) * EVERY} ({time.time() - t0:.0f} s)", flush=True)
    return out


def main():
    with mp.Pool(3) as pool:
        f64, f32, cr = pool.map(run, ["fp64", "f32", "f32_cr"])
    sc = scenes.multi_collider_scene(steps=STEPS)
    doc = {
        "scene": "multi_collider_scene() defaults: 65536 bodies, U{1..8} colliders, 32768 springs, "
                 "16384 fixed joints, step(1/60)",
        "n_colliders": int(sc.n_colliders),
        "statistics": NAMES,
        "steps": [(k + 1) * EVERY for k in range(STEPS // EVERY)],
        "fp64": f64,
        "f32": f32,
        "f32_cr_sha256": cr,
    }
    (Path(__file__).resolve().parent / "multi_collider_stats.json").write_text(json.dumps(doc, indent=1))


if __name__ == "__main__":
    main()
