#!/usr/bin/env python3
"""Benchmark of the MI355X `Physics::step` drop-in.

Workload: the plain large scene -- 1,048,576 single-sphere bodies, f32,
radii U[0.5,1.0], packing fraction 0.4, centres U with overlap rejection,
v ~ N(0,2), gravity on, step(1/120) with the default 8 substeps.  Inputs
are resident in HBM before the timed region; a "step" is one
`Physics::step(1/120)` over the whole world.

Prints ONE JSON line: sphere-steps/s, the roofline record of the dominant
kernel (the fused substep; algorithmic bytes = 64 B per body per substep,
timed with HIP events on the world's stream) and the CPU baseline (the C
restatement of the reference O(C^2) step, timed on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_BODY_SUBSTEP = 64  # pos 8 + pos_old 8 + radius 4 + mass 4 + gravity_mod 4 + rot 4 + omega 4 (read)
#                              pos 8 + pos_old 8 + vel 8 + rot 4 (write)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, budget_s: float) -> dict:
    """Reference algorithm on the host: the oracle (C restatement of
    physics.rs:78-477, single thread).  The O(C^2) contact pass is timed on a
    bounded sample of its rows over the same scene and extrapolated to
    C(C-1)/2 pair tests per substep; the O(N) stages are timed on a full
    step with collisions disabled."""
    from oracle.oracle import OracleWorld
    n = scene.n_bodies
    c = scene.n_colliders
    ow = OracleWorld(scene, np.float32, use_grid=False, collisions_enabled=False)
    t0 = time.perf_counter()
    ow.step(scene.dt, 1)
    t_rest = time.perf_counter() - t0
    ow2 = OracleWorld(scene, np.float32, use_grid=False)
    rows = 16
    t0 = time.perf_counter()
    tests = ow2.bruteforce_rows(c - rows, c)
    t = time.perf_counter() - t0
    rows = max(16, min(c, int(rows * max(budget_s - t_rest, 1.0) / max(t, 1e-6))))
    t0 = time.perf_counter()
    tests = ow2.bruteforce_rows(c - rows, c)
    t_rows = time.perf_counter() - t0
    per_test = t_rows / max(tests, 1)
    substeps = 8
    t_step = substeps * per_test * (c * (c - 1) / 2.0) + t_rest
    return {
        "value": n / t_step,
        "unit": "sphere-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"C restatement of blobs Physics::step (oracle/, f32, O(C^2) brute-force contacts as in "
                   f"physics.rs:248-249) on the same {n}-sphere scene: {rows} contact-loop rows "
                   f"({tests:.3g} pair tests, {t_rows:.1f} s) extrapolated to {substeps} x C(C-1)/2 pair tests per "
                   f"step, plus one full step with collisions off ({t_rest:.2f} s); single thread; extrapolated. "
                   f"Not included: the reference's per-body SpatialHash move (physics.rs:341) and per-contact "
                   f"channel send (physics.rs:304-311), which would only add to its time"),
        "s_per_step": t_step,
    }


TRAFFIC_FILE = ROOT / "profiles" / "k_sphere_substep_plain_1M_traffic.json"


def traffic_per_launch(n_bodies: int):
    """HBM bytes per substep launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each calibrated by a 256 MiB copy kernel), from the committed
    profile summary of the same kernel and workload; None if absent."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if rec.get("bodies") != n_bodies:
        return None
    return rec.get("traffic_bytes_per_launch")


def batched_leg(chunk_steps: int, n_chunks: int, n_worlds: int, autotune: bool, segments: int = 1) -> dict:
    """The uneven batch: n_worlds single-sphere worlds (N_w log-uniform in
    [64, 16384]), step(1/60) in chunk_steps-step hipGraph chunks; whole-job
    sphere-steps/s over the timed chunks (inputs resident in HBM)."""
    from blobs_amd import scenes
    from blobs_amd.physics import Physics
    t0 = time.perf_counter()
    scs = scenes.uneven_batch_scenes(n_worlds)
    worlds = [sc.build_physics()[0] for sc in scs]
    batch = Physics.batch(worlds)
    n_bodies = batch.rbd_count()
    del worlds
    log(f"batched: {n_worlds} worlds, {n_bodies} spheres built in {time.perf_counter() - t0:.1f} s")
    dt = scs[0].dt
    batch.run_chunks(dt, chunk_steps, 1)  # upload, first grid, one-off inputs
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        cs, sk, iv = batch.autotune(dt, 4)
        tuned = {"cell_scale": cs, "skin": sk, "rebuild_interval": iv}
        log(f"batched: autotune -> {tuned} ({time.perf_counter() - t0:.1f} s)")
    broadphase = tuned or {"cell_scale": 1.5, "skin": 0.25, "rebuild_interval": 3, "source": "the batch's defaults"}
    def timed(n, segments=1):
        batch.run_chunks(dt, chunk_steps, 1)  # graph capture warm-up
        batch._lib.bg_synchronize(batch._w)
        total, left = 0.0, n
        for k in range(segments):
            m = left // (segments - k)
            t0 = time.perf_counter()
            batch.run_chunks(dt, chunk_steps, m)  # synchronizes at the end
            el = time.perf_counter() - t0
            total, left = total + el, left - m
            if segments > 1:
                log(f"batched: segment {k}: {m * chunk_steps} steps in {el:.3f} s -> "
                    f"{n_bodies * m * chunk_steps / el:.4g} sphere-steps/s")
        return total

    elapsed = timed(n_chunks, segments)
    persistent = batch.persistent_active()
    steps = chunk_steps * n_chunks
    if not np.isfinite(batch.bodies()["position"]).all():
        raise RuntimeError("non-finite positions in the batched run")
    rec = {
        "workload": f"{n_worlds} uneven worlds (N_w logU[64,16384], r U[0.1,0.5], packing U[0.2,0.5]), step(1/60)",
        "worlds": n_worlds, "bodies": n_bodies, "steps": steps, "chunk_steps": chunk_steps,
        "value": n_bodies * steps / elapsed, "unit": "sphere-steps/s", "ms_per_step": 1000.0 * elapsed / steps,
        "broadphase": broadphase,
        "schedule": "persistent (one workgroup per world, one launch per chunk)" if persistent
        else "grid path (one launch per substep)",
    }
    log(f"batched: {steps} steps in {elapsed:.2f} s -> {rec['value']:.4g} sphere-steps/s ({rec['schedule']})")
    if persistent:
        # the grid path on the same batch, for comparison (same results, other schedule)
        batch.set_persistent(False)
        n_cmp = max(1, n_chunks // 4)
        e2 = timed(n_cmp)
        rec["grid_path_value"] = n_bodies * chunk_steps * n_cmp / e2
        log(f"batched: grid path {chunk_steps * n_cmp} steps in {e2:.2f} s -> {rec['grid_path_value']:.4g} "
            "sphere-steps/s")
    return rec


def differentiable_leg(n_worlds: int, n_steps: int, ck_every: int) -> dict:
    """Forward and forward+adjoint of the differentiable rollout: n_worlds
    worlds of 2,048 spheres + 1,024 springs, step(1/120), checkpoint every
    ck_every steps; gradients of the centroid loss w.r.t. initial velocity
    requests and stiffnesses."""
    from blobs_amd import scenes
    from blobs_amd.rollout import Rollout
    t0 = time.perf_counter()
    ws, targets = scenes.rollout_scenes(n_worlds)
    r = Rollout([w.build_physics()[0] for w in ws], targets, checkpoint_every=ck_every)
    log(f"differentiable: {n_worlds} worlds built in {time.perf_counter() - t0:.1f} s")
    dt = ws[0].dt
    r.forward(dt, n_steps)
    r.backward()  # warm-up (buffers, code objects)
    t0 = time.perf_counter()
    loss = r.forward(dt, n_steps)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    gv, gk = r.backward()
    t_bwd = time.perf_counter() - t0
    fin_v, fin_k = np.isfinite(gv), np.isfinite(gk)
    rec = {
        "workload": f"{n_worlds} worlds x (2048 spheres + 1024 springs), step(1/120) x {n_steps}, "
                    f"checkpoint every {ck_every}",
        "bodies": r.n_bodies, "springs": r.n_springs, "steps": n_steps,
        "forward_s": t_fwd, "forward_adjoint_s": t_fwd + t_bwd,
        "fwd_adj_over_fwd": (t_fwd + t_bwd) / t_fwd,
        "forward_steps_per_s": n_steps / t_fwd, "fwd_adj_steps_per_s": n_steps / (t_fwd + t_bwd),
        "fwd_adj_sphere_steps_per_s": r.n_bodies * n_steps / (t_fwd + t_bwd),
        "loss": float(loss.sum()),
        # the fixed-pair-set derivative of the contact projection grows
        # exponentially with the horizon (1e4-1e33 after 200 steps in the
        # FP64 model), so long rollouts overflow f32 in some worlds
        "grad_finite_fraction": float((fin_v.sum() + fin_k.sum()) / (fin_v.size + fin_k.size)),
        "grad_max_abs_finite": float(max(np.abs(gv[fin_v]).max(initial=0.0), np.abs(gk[fin_k]).max(initial=0.0))),
    }
    log(f"differentiable: forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms, "
        f"ratio {rec['fwd_adj_over_fwd']:.2f}")
    return rec


def plain_small_leg(steps: int = 600) -> dict:
    """The plain CPU-runnable scene (1,024 single spheres + static walls,
    step(1/60) x 600) as a drop-in caller drives it: one Physics::step per
    call.  Synchronous mode waits for each step and checks its panic flags;
    deferred mode (bg_set_deferred_checks) returns once the step is queued."""
    from blobs_amd import scenes
    sc = scenes.plain_small_scene(steps=steps)
    rec = {"workload": "plain_small: 1024 spheres + static walls, step(1/60) x 600, one step per call",
           "bodies": sc.n_bodies, "steps": steps, "unit": "ms/step"}
    for mode in ("synchronous", "deferred"):
        p, _ = sc.build_physics()
        if mode == "deferred":
            p.set_deferred_checks(True)
        p.step(sc.dt)
        p.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            p.step(sc.dt)
        p.synchronize()
        rec[f"ms_per_step_{mode}"] = 1000.0 * (time.perf_counter() - t0) / (steps - 1)
    log(f"plain small: {rec['ms_per_step_synchronous']:.3f} ms/step synchronous, "
        f"{rec['ms_per_step_deferred']:.3f} ms/step deferred")
    return rec


def multi_collider_leg(steps: int = 20, warmup: int = 3) -> dict:
    """The multi-collider scene on the general path: 65,536 bodies with 1-8
    sphere colliders each, 32,768 springs, 16,384 fixed joints, step(1/60)."""
    from blobs_amd import scenes
    sc = scenes.multi_collider_scene()
    p, _ = sc.build_physics()
    p.step(sc.dt, warmup)
    p._lib.bg_synchronize(p._w)
    t0 = time.perf_counter()
    p.step(sc.dt, steps)
    p._lib.bg_synchronize(p._w)
    el = time.perf_counter() - t0
    rec = {
        "workload": "65536 bodies x U{1..8} colliders, 32768 springs, 16384 fixed joints, step(1/60), general path",
        "bodies": sc.n_bodies, "colliders": sc.n_colliders, "steps": steps,
        "ms_per_step": 1000.0 * el / steps, "value": sc.n_bodies * steps / el, "unit": "body-steps/s",
        "path": p.step_path(),
    }
    log(f"multi-collider: {rec['ms_per_step']:.2f} ms/step")
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline sampling")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rebuild-interval", type=int, default=None)
    ap.add_argument("--cell-scale", type=float, default=None)
    ap.add_argument("--skin", type=float, default=None)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 256 MiB device copy (hardware byte-counter calibration)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-differentiable", action="store_true")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--batch-worlds", type=int, default=4096)
    ap.add_argument("--batch-chunks", type=int, default=85,
                    help="timed 10-step chunks (with warm-up and tuning the run stays within 1000 steps)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-autotune", action="store_true",
                    help="autotune the batch instead of its default broadphase (its trials, at the start, "
                         "miss the worlds' spreading over the timed steps)")
    ap.add_argument("--batch-segments", type=int, default=1,
                    help="time the batched chunks in this many consecutive segments, each logged")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the plain step of one world runs on one device and the "
              "multi-GPU decomposition (strip halo exchange) is a design, not built. Run with --gpus 1.",
              file=sys.stderr)
        return 2

    from blobs_amd import scenes

    t0 = time.perf_counter()
    sc = scenes.plain_large_scene(n=args.n, steps=args.steps)
    log(f"scene: {sc.n_bodies} spheres, box {sc.meta['box_side']:.1f}, forced overlaps "
        f"{sc.meta['forced_overlaps']} ({time.perf_counter() - t0:.1f} s)")
    tuned = None
    if args.rebuild_interval or args.cell_scale or args.skin:
        chosen = (args.cell_scale or 1.0, args.rebuild_interval or 1, args.skin or 0.5)
    elif args.no_autotune:
        chosen = None
    else:
        # the autotuner times real steps under candidate broadphase settings on a
        # throwaway copy of the scene (results never depend on them), so the timed
        # world below still starts at step 0 and is timed inside the scene's horizon
        t0 = time.perf_counter()
        q, _ = sc.build_physics()
        q.step(sc.dt, args.warmup)
        cs_, sk_, iv_ = q.autotune(sc.dt, 4)
        del q
        tuned = {"cell_scale": cs_, "skin": sk_, "rebuild_interval": iv_}
        chosen = (cs_, iv_, sk_)
        log(f"autotune on a copy of the scene: {tuned} ({time.perf_counter() - t0:.1f} s)")
    p, _ = sc.build_physics()
    if chosen is not None:
        p.set_broadphase(*chosen)

    # warmup (uploads the world, builds the grid)
    t0 = time.perf_counter()
    p.step(sc.dt, args.warmup)
    p._lib.bg_synchronize(p._w)
    log(f"warmup {args.warmup} steps: {time.perf_counter() - t0:.2f} s, path {p.step_path()}")

    # the timed region: K steps, nothing recorded between the kernels
    p._lib.bg_synchronize(p._w)
    t0 = time.perf_counter()
    p.step(sc.dt, args.steps)  # synchronizes and checks device error flags at the end
    p._lib.bg_synchronize(p._w)
    elapsed = time.perf_counter() - t0

    # The substep kernel's average duration (roofline): the same K steps of a fresh copy
    # of the world, with a HIP event pair around every launch on the world's stream.
    # The events add gaps between launches, so this region is not the one timed above.
    q, _ = sc.build_physics()
    if chosen is not None:
        q.set_broadphase(*chosen)
    q.step(sc.dt, args.warmup)
    q.set_profiling(True)
    q._lib.bg_synchronize(q._w)
    t1 = time.perf_counter()
    q.step(sc.dt, args.steps)
    q._lib.bg_synchronize(q._w)
    elapsed_instrumented = time.perf_counter() - t1
    n_sub, ms_sub = q.stage_stats(0)
    n_reb, ms_reb = q.stage_stats(1)
    q.set_profiling(False)
    del q

    if args.copy_probe:
        import ctypes as C
        p._lib.bgx_copy_probe.argtypes = [C.c_uint64, C.c_uint32]
        if p._lib.bgx_copy_probe(256 << 20, 3) != 0:
            log("copy probe failed")
            return 1

    pos = p.bodies()["position"]
    if not np.isfinite(pos).all():
        log("non-finite positions after the timed region")
        return 1

    ms_per_step = 1000.0 * elapsed / args.steps
    value = sc.n_bodies * args.steps / elapsed
    avg_sub_ms = ms_sub / max(n_sub, 1)
    achieved = BYTES_PER_BODY_SUBSTEP * sc.n_bodies / (avg_sub_ms * 1e-3) / 1e9
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step; instrumented copy: "
        f"{1000.0 * elapsed_instrumented / args.steps:.3f} ms/step, substep kernel avg {avg_sub_ms * 1e3:.1f} us over "
        f"{n_sub} launches, rebuild avg {ms_reb / max(n_reb, 1) * 1e3:.1f} us over {n_reb}")

    line = {
        "metric": "sphere-steps/s @1M (plain vs batched worlds); fwd+adjoint steps/s; % HBM roofline",
        "value": value,
        "unit": "sphere-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain_large: Physics::step(1/120), 8 substeps, single-sphere bodies; "
                        "timed steps [warmup, warmup+steps) of the scene",
            "bodies": sc.n_bodies,
            "radius": [0.5, 1.0],
            "packing_fraction": 0.4,
            "box_side": round(sc.meta["box_side"], 3),
            "dt": sc.dt,
            "substeps": 8,
            "path": p.step_path(),
            "broadphase_autotuned": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_sphere_substep (fused gravity+contacts+Verlet+transforms+constraints)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_launch(sc.n_bodies),
            "traffic_source": "committed rocprofv3 --pmc summary of this kernel and workload "
                              "(profiles/k_sphere_substep_plain_1M_traffic.json: FETCH_SIZE and WRITE_SIZE per "
                              "dispatch, each calibrated by a 256 MiB copy kernel), not measured in this run",
            "bytes_per_launch": BYTES_PER_BODY_SUBSTEP * sc.n_bodies,
            "avg_launch_us": avg_sub_ms * 1e3,
            "rebuild_avg_us": ms_reb / max(n_reb, 1) * 1e3,
            "rebuilds": n_reb,
            "launch_timing": "the kernel's own start and end timestamps (hipExtLaunchKernelGGL events on the "
                             "world's stream) for every substep launch over the same steps of a fresh copy of the "
                             "world (the timed region itself runs without them); rocprofv3 --kernel-trace on the same "
                             "command gives the same per-launch average",
            "instrumented_ms_per_step": 1000.0 * elapsed_instrumented / args.steps,
        },
        "cpu_baseline": None,
    }
    if not args.no_batched:
        b = batched_leg(10, args.batch_chunks, args.batch_worlds, args.batch_autotune, max(1, args.batch_segments))
        b["vs_plain"] = b["value"] / value
        line["batched"] = b
    if not args.no_differentiable:
        line["differentiable"] = differentiable_leg(256, 200, 20)
    if not args.no_multi:
        line["multi_collider"] = multi_collider_leg()
    if not args.no_small:
        line["plain_small"] = plain_small_leg()
    if not args.no_cpu:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_budget)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.4g} sphere-steps/s ({time.perf_counter() - t0:.1f} s)")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
