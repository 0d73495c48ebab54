#!/usr/bin/env python3
"""Throughput of the PBF substep on one MI355X.

Workload: dam break, 1,048,576 particles (128x128x64 block at 0.5h, h=0.05,
box 12.8x6.4x6.4), 4 solver iterations, dt = 1/120 s, fp32.  A step is one
Simulation::Step substep.  W warm-up substeps from the seeded initial state,
then EXACTLY K timed substeps bracketed by stream synchronisation; the
inputs are resident in HBM before timing starts.

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP events on the library's stream over the timed region) and a CPU
baseline (the oracle restatement on this host, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "position-based-fluids_amd"))

METRIC = "particle-substeps/s at 1M particles, 4 iters"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle for each stage (compulsory streaming traffic
# with cell-sorted SoA float4 arrays; neighbour gathers served from cache and
# neighbour-list traffic excluded).  Fused kernels sum the stages they cover.
STAGE_BYTES = {
    "predict_hash": 56, "sort_pass": 16, "gather": 68, "cells": 6, "lambda": 20, "delta": 52, "final_rho": 4,
    "velocity": 64, "xsph": 68, "vorticity": 68,
}


def kernel_bytes_per_particle(name: str, iters: int, passes: int) -> float | None:
    base = name.rsplit("_", 1)[0] if name.rsplit("_", 1)[-1].isdigit() else name
    if base == "predictPositions":
        return STAGE_BYTES["predict_hash"]
    if base == "onesweep":
        return (4 + 16 * passes) / passes
    if base == "sortParticles":
        return STAGE_BYTES["gather"]
    if base == "updateCells":
        return STAGE_BYTES["cells"]
    if base == "computeScaling":
        return STAGE_BYTES["lambda"]
    if base == "computeDelta":
        last = name == f"computeDelta_{iters - 1}"
        return STAGE_BYTES["delta"] + (STAGE_BYTES["final_rho"] + STAGE_BYTES["velocity"] if last else 0)
    if base == "applyViscosity":
        return STAGE_BYTES["xsph"]
    if base == "applyVorticity":
        return STAGE_BYTES["vorticity"]
    return None


def family(name: str) -> str:
    tail = name.rsplit("_", 1)
    return tail[0] if len(tail) == 2 and tail[1].isdigit() else name


def cpu_baseline(scene, budget_s: float = 15.0) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import pbf_oracle
    o = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(scene.params, scene.options))
    o.set_state(scene.x, scene.v)
    o.step()  # warm-up (first-touch of the host buffers)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 20:
            break
    o.close()
    return {
        "value": scene.n * steps / el,
        "unit": "particle-substeps/s",
        "cores": int(pbf_oracle.lib().pbfo_threads()),
        "kind": "port",
        "sample": f"dam1m substeps 2..{steps + 1} from the initial state ({steps} timed, {el:.1f} s), "
                  f"C/OpenMP restatement of the reference kernels",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    import pbf_amd
    scene = pbf_amd.config("dam1m", seed=args.seed)
    n = scene.n
    iters = int(scene.params.simIterations)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    sim.run(args.warmup)
    sim.sync()
    sim.set_profiling(True)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    timings = sim.timings()
    sim.set_profiling(False)
    flags, dropped = sim.error_flags()
    sim.close()

    value = n * args.steps / el
    passes = 0
    while (1 << (8 * passes)) < int(scene.params.gridBufSize):
        passes += 1
    passes = max(passes, 1)

    fam_ms: dict[str, float] = {}
    for name, t in timings.items():
        fam_ms[family(name)] = fam_ms.get(family(name), 0.0) + t["sum_ms"]
    dom = max(fam_ms, key=fam_ms.get)
    launches = [(name, t) for name, t in timings.items() if family(name) == dom]
    tot_ms = sum(t["sum_ms"] for _, t in launches)
    tot_calls = sum(t["calls"] for _, t in launches)
    bytes_per_launch = sum(kernel_bytes_per_particle(name, iters, passes) * n * t["calls"]
                           for name, t in launches) / max(tot_calls, 1)
    avg_ms = tot_ms / max(tot_calls, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    step_ms = {k: round(v / args.steps, 4) for k, v in sorted(fam_ms.items(), key=lambda kv: -kv[1])}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "dam1m: PBF dam break, 1,048,576 particles (128x128x64 block at 0.5h, h=0.05, "
                        "box 12.8x6.4x6.4), 4 solver iterations, dt=1/120 s, seeded jitter",
            "particles": n,
            "solver_iterations": iters,
            "grid_buf_size": int(scene.params.gridBufSize),
            "sort_passes": passes,
            "substeps_before_timing": args.warmup,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
        },
        "stage_ms_per_step": step_ms,
        "error_flags": flags,
        "dropped_friends": dropped,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
