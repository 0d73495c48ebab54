#!/usr/bin/env python3
"""Throughput of the PBF substep on one MI355X.

Workload: dam break, 1,048,576 particles (128x128x64 block at 0.5h, h=0.05,
box 12.8x6.4x6.4), 4 solver iterations, dt = 1/120 s, fp32.  A step is one
Simulation::Step substep.  W warm-up substeps from the seeded initial state,
then EXACTLY K timed substeps bracketed by stream synchronisation; the
inputs are resident in HBM before timing starts.

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP events that the kernels themselves stamp at their start and end, over a
profiled replay of the timed substeps) and a CPU baseline (the oracle
restatement on this host, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "position-based-fluids_amd"))

METRIC = "particle-substeps/s at 1M particles, 4 iters; batched gain; adjoint/forward ratio"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle for each stage (compulsory streaming traffic
# with cell-sorted SoA float4 arrays; neighbour gathers served from cache and
# neighbour-list traffic excluded).  Fused kernels sum the stages they cover.
STAGE_BYTES = {
    "predict_hash": 56, "sort_pass": 16, "gather": 68, "cells": 6, "lambda": 20, "delta": 52, "final_rho": 4,
    "velocity": 64, "xsph": 68, "vorticity": 68,
}


# HBM bytes per launch measured with rocprofv3 PMC counters on the
# 1,048,576-particle dam break, over the launches of this script's profiled
# pass (FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors a plain 256 MiB copy
# gives the two counters), averaged over a family's launches (computeDelta:
# three per substep plus the last, which also writes v and x);
# profiles/pmc_traffic_dam1m.md holds the counts and commands (the
# --pmc-summary mode of this script prints this table).  They belong to the
# kernel sources whose code hash is stamped here (KERNEL_SOURCES, comments
# and indentation left out): with any other code the line reports traffic as
# null.
PMC_TRAFFIC = {
    "source_sha16": "f5d1cd2fda632585",
    "bytes_per_launch": {
        "applyViscosity": 220147697,
        "applyVorticity": 221264322,
        "buildFriendsList": 944571338,
        "computeDelta": 206793692,
        "computeScaling": 205410462,
        "onesweep": 17673322,
        "predictPositions": 43423758,
        "sortParticles": 97364246,
    },
}

KERNEL_SOURCES = ("pbf_forward.hip", "pbf_device.h", "pbf_sort.hip", "pbf_internal.h")


def source_code_tokens(text: str) -> str:
    """C/C++ source with its comments removed and each whitespace run made a
    line break (if it holds one; preprocessor lines end there) or a space:
    what the compiler sees, so an edit to comments or indentation alone
    keeps the stamp.  String and character literals are kept whole."""
    out, i, n = [], 0, len(text)
    while i < n:
        ch = text[i]
        if ch in "\"'":
            j = i + 1
            while j < n and text[j] != ch:
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
            out.append(" ")
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
            out.append(" ")
        else:
            out.append(ch)
            i += 1
    return re.sub(r"\s+", lambda m: "\n" if "\n" in m.group(0) else " ", "".join(out)).strip()


def kernel_source_sha16() -> str:
    """Hash of the kernel sources' code (comments and layout left out)."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        text = (ROOT / "position-based-fluids_amd" / "csrc" / name).read_text()
        h.update(source_code_tokens(text).encode())
        h.update(b"\0")
    return h.hexdigest()[:16]


def pmc_traffic(family_name: str, n: int) -> float | None:
    """Measured HBM bytes per launch of a kernel family at 1,048,576
    particles, or None when the table was measured on other kernel sources."""
    if n != 1048576 or PMC_TRAFFIC["source_sha16"] != kernel_source_sha16():
        return None
    return PMC_TRAFFIC["bytes_per_launch"].get(family_name)


def kernel_bytes_per_particle(name: str, iters: int, passes: int, friends: bool = False) -> float | None:
    """Algorithmic bytes per particle of one launch.  The friends build has a
    figure only for the first computeScaling it runs (its lists are index
    structure, which the algorithmic figure excludes); friends=False leaves
    it out, as the roofline kernel is chosen among the streaming stages."""
    base = name.rsplit("_", 1)[0] if name.rsplit("_", 1)[-1].isdigit() else name
    if base == "buildFriendsList":
        return STAGE_BYTES["lambda"] if friends else None
    if base == "predictPositions":
        return STAGE_BYTES["predict_hash"]
    if base == "onesweep":
        return (4 + 16 * passes) / passes
    if base == "sortParticles":   # k_gather_cells: sortParticles + updateCells
        return STAGE_BYTES["gather"] + STAGE_BYTES["cells"]
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


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(scene, max_steps: int, budget_s: float) -> tuple[float, int, float]:
    """particle-substeps/s of the CPU oracle from the scene's initial state
    (one untimed warm-up substep, then up to max_steps or budget_s)."""
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
        if el >= budget_s or steps >= max_steps:
            break
    o.close()
    return scene.n * steps / el, steps, el


def cpu_baseline(scene, budget_s: float = 15.0, runs: int = 3) -> dict:
    """The oracle on this host's cores: the median of `runs` runs of the
    1M-particle dam break (substeps 2..21 from the initial state) and of the
    16,384-particle dam break over its full 200 substeps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pbf_oracle
    import pbf_amd
    big = sorted((_oracle_rate(scene, 20, budget_s) for _ in range(runs)), key=lambda r: r[0])
    rate, steps, el = big[len(big) // 2]
    small = pbf_amd.config("dam16k", seed=0)
    sm = sorted((_oracle_rate(small, 200, budget_s) for _ in range(runs)), key=lambda r: r[0])
    rate16k, steps16k, el16k = sm[len(sm) // 2]
    return {
        "value": rate,
        "unit": "particle-substeps/s",
        "cores": int(pbf_oracle.lib().pbfo_threads()),
        "kind": "port",
        "sample": f"dam1m substeps 2..{steps + 1} from the initial state ({steps} timed, {el:.1f} s), median of "
                  f"{runs} runs, C/OpenMP restatement of the reference kernels",
        "runs": [r[0] for r in big],
        "host": {"nproc": os.cpu_count(), "model": _cpu_model()},
        "dam16k": {"value": rate16k, "steps": steps16k, "seconds": round(el16k, 2),
                   "ms_per_substep": round(el16k / steps16k * 1e3, 3), "runs": [r[0] for r in sm]},
    }


CAPTURE_STEPS = 6


def capture_graphs(sim) -> None:
    """Runs CAPTURE_STEPS substeps that capture every graph a later run
    replays from the state they leave: one-substep graphs of both buffer
    rotations (pbf_run's remainder), then a group of four (pbf_run groups
    four substeps to a graph), so that no capture falls inside a timed run
    of any length.  Both rotations recur every two substeps."""
    sim.run(2)
    sim.run(4)


def full_run(scene, sizes: dict | None) -> dict:
    """All 500 substeps of the workload from its initial state, replayed as
    HIP graphs (the default launch shapes, or the tuned sizes), with the
    error flags and dropped friends they raise: the late substeps where
    circles reach the 50-friend cap are timed too."""
    import pbf_amd
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    if sizes:
        sim.set_workgroup_sizes(sizes)
    sim.set_graph(True)
    # capture the graphs the timed run replays (pbf_run groups its substeps;
    # the first substep after a new state is a variant of its own), so that
    # no capture falls inside it
    for _ in range(2):
        sim.run(4)
        sim.set_state(scene.x, scene.v)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(500)
    sim.sync()
    el = time.perf_counter() - t0
    flags, dropped = sim.error_flags()
    sim.close()
    return {"substeps": 500, "seconds": el, "particle_substeps_per_s": scene.n * 500 / el,
            "ms_per_substep": el * 1e3 / 500, "error_flags": flags, "dropped_friends": dropped}


def batched_gain(steps: int, warmup: int, seed: int, tune: bool) -> dict:
    """64-scene ensemble: one launch per stage (graph-replayed substeps)
    against the single-scene path over the same scenes.

    gain: against the sum of the single-scene times (each scene stepped
    alone, graph-replayed, one after another).  gain_vs_overlapped_loop:
    against a loop that steps every scene's own context in turn without
    waiting, so the GPU overlaps the scenes' streams (the faster of eager
    launches and graph replay)."""
    import pbf_amd
    scenes = pbf_amd.ensemble(64, seed=seed)
    total = sum(s.n for s in scenes)
    x0 = np.concatenate([s.x for s in scenes])
    v0 = np.concatenate([s.v for s in scenes])

    batch = pbf_amd.BatchSimulation([s.params for s in scenes], [s.options for s in scenes])
    batch.set_state(x0, v0)
    batch.run(warmup)
    sizes = batch.autotune(2) if tune else None
    batch.set_graph(True)
    capture_graphs(batch)
    batch.sync()
    t0 = time.perf_counter()
    batch.run(steps)
    batch.sync()
    t_batch = time.perf_counter() - t0
    batch.close()

    singles = []
    for s in scenes:
        sim = pbf_amd.Simulation(s.params, s.options)
        sim.set_state(s.x, s.v)
        if sizes:
            sim.set_workgroup_sizes(sizes)
        singles.append(sim)

    # each scene alone, after the same warm-up as the batch
    t_single = 0.0
    for sim in singles:
        sim.set_graph(True)
        sim.run(warmup)
        capture_graphs(sim)         # as the batch
        sim.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.sync()
        t_single += time.perf_counter() - t0
    for s, sim in zip(scenes, singles):
        sim.set_state(s.x, s.v)

    def loop_time(graph: bool) -> float:
        for sim in singles:
            sim.set_graph(graph)
        for _ in range(warmup + CAPTURE_STEPS):   # the batch's timed substeps
            for sim in singles:
                sim.Step()
        for sim in singles:
            sim.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            for sim in singles:
                sim.Step()
        for sim in singles:
            sim.sync()
        return time.perf_counter() - t0

    # the loop over the single-scene path, with eager launches and with each
    # context's substeps replayed as graphs; the gain is against the faster
    t_eager = loop_time(False)
    t_graph = loop_time(True)
    t_loop = min(t_eager, t_graph)
    for sim in singles:
        sim.close()
    return {
        "scenes": len(scenes),
        "particles": total,
        "steps": steps,
        "batched_particle_substeps_per_s": total * steps / t_batch,
        "single_scenes_particle_substeps_per_s": total * steps / t_single,
        "looped_particle_substeps_per_s": total * steps / t_loop,
        "looped_eager_particle_substeps_per_s": total * steps / t_eager,
        "looped_graph_particle_substeps_per_s": total * steps / t_graph,
        "gain": t_single / t_batch,
        "gain_vs_overlapped_loop": t_loop / t_batch,
        "workgroup_sizes": sizes,
    }


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _dev4(a):
    import torch
    t = torch.zeros((a.shape[0], 4), dtype=torch.float32)
    t[:, :3] = torch.from_numpy(np.ascontiguousarray(a[:, :3], np.float32))
    return t.cuda()


def _plain_forward_s(sim, x, v, steps: int, reps: int = 3) -> float:
    """Best wall time of `steps` graph-replayed substeps from (x, v)."""
    sim.set_graph(True)
    sim.set_state(x, v)
    sim.run(steps)
    sim.sync()
    best = float("inf")
    for _ in range(reps):
        sim.set_state(x, v)
        sim.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.sync()
        best = min(best, time.perf_counter() - t0)
    return best


def adjoint_ratio(seed: int, steps: int = 60, interval: int = 10, reps: int = 3) -> dict:
    """Inverse problem: 65,536 particles (64x32x32 block, h=0.1), 60
    substeps, checkpoint every 10.  Targets are the rollout from Gaussian
    initial velocities (sigma 0.5 m/s); the rollout differentiated starts at
    v0 = 0 with loss mean |xT - target|^2.  Ratio = backward (replays
    included) / plain forward of the same substeps."""
    import torch

    import pbf_amd
    from pbf_amd.adjoint import Rollout
    scene = pbf_amd.scenes.inverse_problem(64, 32, 32, seed=seed)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    t_fwd = _plain_forward_s(sim, scene.x, np.zeros_like(scene.v), steps, reps)
    sim.set_graph(False)
    roll = Rollout(sim, steps, interval)
    target, _ = roll.forward_raw(_dev4(scene.x), _dev4(scene.v))
    x0, v0 = _dev4(scene.x), _dev4(np.zeros_like(scene.v))
    fwd, bwd, rep = [], [], []
    for r in range(reps + 1):
        xT, _ = roll.forward_raw(x0, v0)
        g = 2.0 * (xT - target) / scene.n
        g[:, 3] = 0.0
        _, gv0, _ = roll.backward_raw(g, None)
        if r:
            t = roll.times()
            fwd.append(t["forward_ms"])
            bwd.append(t["backward_ms"])
            rep.append(t["replay_ms"])
    loss = float(((xT - target)[:, :3] ** 2).sum(1).mean())
    gnorm = float(gv0.norm())
    roll.close()
    sim.close()
    b = min(bwd)
    return {
        "workload": "inverse problem: 65,536 particles (64x32x32 block, h=0.1), 60 substeps, 4 iterations, "
                    "checkpoint every 10, loss mean |xT - target|^2 w.r.t. v0, c, eps",
        "particles": scene.n,
        "substeps": steps,
        "interval": interval,
        "plain_forward_ms": t_fwd * 1e3,
        "adjoint_forward_ms": min(fwd),
        "adjoint_backward_ms": b,
        "replay_ms": rep[bwd.index(b)],
        "ratio": b / (t_fwd * 1e3),
        "loss": loss,
        "grad_v0_norm": gnorm,
    }


def batched_adjoint(seed: int, steps: int = 100, budget_gb: float = 64.0, reps: int = 2) -> dict:
    """Batched differentiation: 32 scenes (counts log-uniform in [8192,
    65536], h=0.1, per-scene gravity of 9.8 m/s^2 on the lower hemisphere),
    100 substeps; per-scene loss |centroid - target|^2 w.r.t. the scene's
    gravity vector and initial velocities; interval planned under the
    budget."""
    import torch

    import pbf_amd
    from pbf_amd.adjoint import Rollout, plan_interval, tape_bytes
    scenes, targets = pbf_amd.batch.gravity_ensemble(32, seed=seed)
    batch = pbf_amd.BatchSimulation([s.params for s in scenes], [s.options for s in scenes])
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    t_fwd = _plain_forward_s(batch, x, v, steps, reps)
    batch.set_graph(False)
    interval = plan_interval(batch, steps, int(budget_gb * (1 << 30)))
    slot, ck = tape_bytes(batch)
    roll = Rollout(batch, steps, interval)
    off = batch.offsets.astype(np.int64)
    sid = torch.from_numpy(np.repeat(np.arange(len(scenes)), np.diff(off))).cuda()
    cnt = torch.from_numpy(np.diff(off).astype(np.float32)).cuda()
    tgt = torch.from_numpy(targets).cuda()
    x0, v0 = _dev4(x), _dev4(v)
    fwd, bwd, rep = [], [], []
    for r in range(reps + 1):
        xT, _ = roll.forward_raw(x0, v0)
        cen = torch.zeros((len(scenes), 3), device="cuda").index_add_(0, sid, xT[:, :3]) / cnt[:, None]
        g = torch.zeros_like(xT)
        g[:, :3] = (2.0 * (cen - tgt) / cnt[:, None])[sid]
        _, gv0, gc = roll.backward_raw(g, None)
        if r:
            t = roll.times()
            fwd.append(t["forward_ms"])
            bwd.append(t["backward_ms"])
            rep.append(t["replay_ms"])
    loss = ((cen - tgt) ** 2).sum(1)
    gg = gc[:, 2:5].double().cpu().numpy()
    fin = np.isfinite(gg).all(axis=1)
    gn = np.linalg.norm(np.where(fin[:, None], gg, 0.0), axis=1)
    roll.close()
    batch.close()
    b = min(bwd)
    total = int(off[-1])
    return {
        "workload": "32 scenes, counts log-uniform in [8192, 65536] (h=0.1), 100 substeps, per-scene gravity "
                    "on the lower hemisphere, loss |centroid - target|^2 per scene",
        "scenes": len(scenes),
        "particles": total,
        "substeps": steps,
        "budget_gb": budget_gb,
        "interval": interval,
        "tape_gb": (slot * interval + ck * ((steps + interval - 1) // interval)) / (1 << 30),
        "plain_forward_ms": t_fwd * 1e3,
        "adjoint_forward_ms": min(fwd),
        "adjoint_backward_ms": b,
        "replay_ms": rep[bwd.index(b)],
        "ratio": b / (t_fwd * 1e3),
        "differentiated_particle_substeps_per_s": total * steps / ((min(fwd) + b) * 1e-3),
        "mean_loss": float(loss.mean()),
        # sensitivities of these flows grow ~3x per substep after ~25
        # substeps; scenes whose true gradient exceeds binary32 report inf
        "grad_gravity_finite_scenes": int(fin.sum()),
        "grad_gravity_norm_median": float(np.median(gn[fin])) if fin.any() else None,
    }


# ---------------------------------------------------------------------------
# rocprofv3 summaries.  The bench brackets its profiled pass (the launches its
# roofline and stage times come from) with two tiny k_copy_probe dispatches,
# launched after a device synchronisation; nothing else in a bench run
# launches that kernel.  A kernel trace or counter run of the bench is cut to
# the dispatches between the two groups, which are the line's launches.
# ---------------------------------------------------------------------------
WINDOW_MARK_KERNEL = "k_copy_probe"
TARE_TRACKER, TARE_KERNEL = "eventTare", "k_tare"   # the empty kernel of a profiled substep

KERNEL_FAMILIES = (
    ("k_friends_direct<", "buildFriendsList"), ("k_friends<", "buildFriendsList"),
    ("k_lambda<", "computeScaling"), ("k_delta<", "computeDelta"), ("k_viscosity<", "applyViscosity"),
    ("k_vorticity<", "applyVorticity"), ("k_onesweep", "onesweep"), ("k_predict_hash", "predictPositions"),
    ("k_predict_only", "predictPositions"), ("k_gather_cells", "sortParticles"),
)


def _trace_mark() -> None:
    from pbf_amd._lib import copy_probe
    copy_probe(1 << 16, 1)


def _kernel_short(name: str) -> str:
    """'void pbf::(anonymous namespace)::k_delta<256, false, false, false>(...)' -> 'k_delta<256, false, false, false>'."""
    base = name[5:] if name.startswith("void ") else name
    for ns in ("pbf::(anonymous namespace)::", "(anonymous namespace)::", "pbf::"):
        base = base.replace(ns, "")
    return base.split("(")[0].strip()


def _family_of_kernel(short: str) -> str | None:
    for prefix, fam in KERNEL_FAMILIES:
        if short.startswith(prefix):
            return fam
    return None


def _read_dispatches(run_dir: str, counter: str | None = None) -> list[dict]:
    """Dispatches of a rocprofv3 CSV run (--output-format csv): kernel,
    start, end (ns) and, for a counter run, the counter's value."""
    import csv
    import glob
    pat = "*counter_collection.csv" if counter else "*kernel_trace.csv"
    files = sorted(glob.glob(os.path.join(run_dir, "**", pat), recursive=True))
    if not files:
        raise SystemExit(f"no {pat} under {run_dir}")
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if counter and r.get("Counter_Name") != counter:
                    continue
                rows.append({"kernel": _kernel_short(r["Kernel_Name"]), "start": int(r["Start_Timestamp"]),
                             "end": int(r["End_Timestamp"]),
                             "value": float(r["Counter_Value"]) if counter else None})
    rows.sort(key=lambda d: d["start"])
    return rows


def _window(rows: list[dict]) -> tuple[int, int]:
    """(end of the first marker group, start of the second)."""
    marks = [d for d in rows if d["kernel"].startswith(WINDOW_MARK_KERNEL)]
    if len(marks) < 2:
        raise SystemExit("the run holds no window markers (was it bench.py's forward line?)")
    gaps = [(marks[q + 1]["start"] - marks[q]["end"], q) for q in range(len(marks) - 1)]
    _, q = max(gaps)
    return marks[q]["end"], marks[q + 1]["start"]


def _in_window(rows: list[dict], keep: str | None = None) -> list[dict]:
    """The pass's dispatches of the stage kernels (or of kernel `keep`)."""
    lo, hi = _window(rows)
    sel = (lambda k: k.startswith(keep)) if keep else _family_of_kernel
    return [d for d in rows if lo <= d["start"] and d["end"] <= hi and sel(d["kernel"])]


def trace_summary(run_dir: str, line_path: str | None) -> str:
    import csv
    import glob
    rows = _read_dispatches(run_dir)
    out = []
    stats = sorted(glob.glob(os.path.join(run_dir, "**", "*kernel_stats.csv"), recursive=True))
    if stats:   # rocprofv3 --stats, as written
        out.append("Whole run, every dispatch (`--stats` summary, microseconds):\n")
        out.append("| Kernel | Calls | Total (us) | Average (us) | % |\n|---|---:|---:|---:|---:|")
        with open(stats[0], newline="") as fh:
            for r in list(csv.DictReader(fh))[:24]:
                out.append(f"| `{_kernel_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
                           f"{float(r['AverageNs']) / 1e3:.3f} | {float(r['Percentage']):.2f} |")
    win = _in_window(rows)
    fam: dict[str, list[float]] = {}
    per_k: dict[str, list[float]] = {}
    for d in win:
        us = (d["end"] - d["start"]) / 1e3
        fam.setdefault(_family_of_kernel(d["kernel"]), []).append(us)
        per_k.setdefault(d["kernel"], []).append(us)
    out.append("\nThe bench line's profiled pass (dispatches between the window markers):\n")
    out.append("| Kernel | Calls | Average (us) |\n|---|---:|---:|")
    for k, v in sorted(per_k.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| `{k}` | {len(v)} | {sum(v) / len(v):.3f} |")
    line = None
    if line_path:
        with open(line_path) as fh:
            for ln in fh:
                ln = ln.strip()
                if ln.startswith("{") and '"metric"' in ln:
                    line = json.loads(ln)
    tare_tr = [(d["end"] - d["start"]) / 1e3 for d in _in_window(rows, keep=TARE_KERNEL)]
    if tare_tr:
        ev = line["roofline"].get("empty_kernel_ms") if line else None
        out.append(f"\nThe empty `{TARE_KERNEL}` launches of the pass: {len(tare_tr)}, {np.mean(tare_tr):.3f} us each "
                   "in the trace" + (f", {ev * 1e3:.3f} us by the line's events." if ev else "."))
    out.append("\n| Family | Launches | Trace average (us) | Line average (us) | Line - trace (us) | Trace / line |\n"
               "|---|---:|---:|---:|---:|---:|")
    for f, v in sorted(fam.items(), key=lambda kv: -sum(kv[1])):
        avg = sum(v) / len(v)
        la = None
        if line:
            bk = line["roofline"]["by_kernel"].get(f)
            la = bk["avg_launch_ms"] * 1e3 if bk else (line["roofline"]["avg_launch_ms"] * 1e3
                                                         if line["roofline"]["kernel"] == f else None)
        out.append(f"| {f} | {len(v)} | {avg:.3f} | {la:.3f} | {la - avg:.3f} | {avg / la:.4f} |" if la else
                   f"| {f} | {len(v)} | {avg:.3f} | - | - | - |")
    if line:
        r = line["roofline"]
        v = fam.get(r["kernel"])
        if v:
            avg_s = sum(v) / len(v) * 1e-6
            frac = r["algorithmic_bytes_per_launch"] / avg_s / 1e9 / r["peak"]
            out.append(f"\nRoofline kernel {r['kernel']}: {r['algorithmic_bytes_per_launch'] / 1e6:.3f} MB "
                       f"algorithmic per launch / {avg_s * 1e6:.3f} us (trace) = {frac:.5f} of {r['peak']:.0f} GB/s; "
                       f"the line reports frac {r['frac']:.5f} (ratio {frac / r['frac']:.4f}).")
    return "\n".join(out)


def pmc_summary(fetch_dir: str, write_dir: str, copy_fetch_dir: str, copy_write_dir: str, copy_mib: int,
                line_path: str | None) -> str:
    """Per-launch HBM bytes of each family over the bench line's profiled
    pass, with FETCH_SIZE and WRITE_SIZE scaled by what they read for a plain
    copy of copy_mib MiB (copy_mib MiB read and as many written)."""
    kib = copy_mib * 1024.0
    cf = [d["value"] for d in _read_dispatches(copy_fetch_dir, "FETCH_SIZE") if d["kernel"].startswith("k_copy_probe")]
    cw = [d["value"] for d in _read_dispatches(copy_write_dir, "WRITE_SIZE") if d["kernel"].startswith("k_copy_probe")]
    # the probe's warm-up dispatch runs first; the timed ones follow
    f_cal, w_cal = kib / float(np.median(cf[1:] or cf)), kib / float(np.median(cw[1:] or cw))
    out = [f"Calibration: a plain copy of {copy_mib} MiB ({kib:.0f} KiB read and written) reads FETCH_SIZE = "
           f"{float(np.median(cf[1:] or cf)):.0f} KiB and WRITE_SIZE = {float(np.median(cw[1:] or cw)):.0f} KiB per "
           f"dispatch: HBM bytes = (FETCH_SIZE x {f_cal:.3f} + WRITE_SIZE x {w_cal:.3f}) x 1024.\n"]
    fw = {"FETCH_SIZE": _in_window(_read_dispatches(fetch_dir, "FETCH_SIZE")),
          "WRITE_SIZE": _in_window(_read_dispatches(write_dir, "WRITE_SIZE"))}
    per_k: dict[str, dict[str, list[float]]] = {}
    fam: dict[str, dict[str, list[float]]] = {}
    for cname, rows in fw.items():
        for d in rows:
            per_k.setdefault(d["kernel"], {}).setdefault(cname, []).append(d["value"])
            fam.setdefault(_family_of_kernel(d["kernel"]), {}).setdefault(cname, []).append(d["value"])
    # each pass autotunes on its own, so a family's kernel may differ between
    # the two passes (another workgroup size): kernels are listed per pass
    # and the figures bench.py uses are per family
    out.append("| Kernel | FETCH_SIZE dispatches | FETCH_SIZE (KiB) | WRITE_SIZE dispatches | WRITE_SIZE (KiB) |\n"
               "|---|---:|---:|---:|---:|")
    for k, v in sorted(per_k.items(), key=lambda kv: kv[0]):
        f, w = v.get("FETCH_SIZE"), v.get("WRITE_SIZE")
        out.append(f"| `{k}` | {len(f) if f else 0} | {np.mean(f):.1f} | " if f else f"| `{k}` | 0 | - | ")
        out[-1] += f"{len(w)} | {np.mean(w):.1f} |" if w else "0 | - |"

    def mb(v: dict) -> float:
        return (np.mean(v["FETCH_SIZE"]) * f_cal + np.mean(v["WRITE_SIZE"]) * w_cal) * 1024 / 1e6

    out.append("\n| Family | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM MB per launch |\n|---|---:|---:|---:|")
    for f, v in sorted(fam.items(), key=lambda kv: -mb(kv[1]) if len(kv[1]) == 2 else 0):
        if len(v) == 2:
            out.append(f"| {f} | {np.mean(v['FETCH_SIZE']):.1f} | {np.mean(v['WRITE_SIZE']):.1f} | {mb(v):.2f} |")
    sha = kernel_source_sha16()
    if line_path:
        with open(line_path) as fh:
            for ln in fh:
                if ln.strip().startswith("{") and '"metric"' in ln:
                    sha = json.loads(ln)["roofline"]["kernel_source_sha16"]
    table = {f: round(mb(v) * 1e6) for f, v in sorted(fam.items()) if "FETCH_SIZE" in v and "WRITE_SIZE" in v}
    out.append("\nPer launch by family (averaged over the pass's launches), the table bench.py carries:\n")
    out.append("```\nPMC_TRAFFIC = " + json.dumps({"source_sha16": sha, "bytes_per_launch": table}, indent=4)
               + "\n```")
    return "\n".join(out)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-full-run", action="store_true", help="skip the 500-substep run from the initial state")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default workgroup sizes")
    ap.add_argument("--batched-steps", type=int, default=30)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint and batched-differentiation lines")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain HBM copy of MIB MiB (calibration under rocprofv3) and print GB/s")
    ap.add_argument("--trace-summary", metavar="DIR",
                    help="only summarise a rocprofv3 --kernel-trace CSV run of this script (no GPU needed)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="only summarise FETCH_SIZE / WRITE_SIZE runs of this script and of --copy-probe")
    ap.add_argument("--copy-mib", type=int, default=256, help="the --copy-probe size of the --pmc-summary runs")
    ap.add_argument("--line", metavar="FILE", help="output of the profiled run (its JSON line), for the summaries")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary, args.line))
        return 0
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary, args.copy_mib, args.line))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    import pbf_amd
    if args.copy_probe:
        from pbf_amd._lib import copy_probe
        print(json.dumps({"copy_probe_mib": args.copy_probe, "reps": 10,
                          "gbps": copy_probe(args.copy_probe << 20, 10)}))
        return 0
    scene = pbf_amd.config("dam1m", seed=args.seed)
    n = scene.n
    iters = int(scene.params.simIterations)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    sim.run(args.warmup)
    # tuned on the warmed-up state, which autotune restores afterwards
    sizes = sim.autotune(4) if not args.no_autotune else sim.stats()["workgroup_size"]
    # the timed region: substeps replayed as HIP graphs, no events between launches
    sim.set_graph(True)
    capture_graphs(sim)
    sim.sync()
    start = sim.get_state()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    # the same substeps again from the same state (bitwise the same work), each
    # launch stamping a pair of HIP events at its kernel's start and end: the
    # per-stage breakdown and the roofline kernel's launch durations.  The
    # window markers let a rocprofv3 run of this script find these dispatches.
    sim.set_state(start["x"], start["v"])
    sim.set_profiling(True)
    sim.reset_timings()
    sim.sync()
    _trace_mark()
    sim.run(args.steps)
    sim.sync()
    _trace_mark()
    timings = sim.timings()
    sim.set_profiling(False)
    flags, dropped = sim.error_flags()
    sim.close()
    prof_steps = args.steps

    value = n * args.steps / el
    _progress(f"forward: {value:.4g} particle-substeps/s")
    passes = 0
    while (1 << (8 * passes)) < int(scene.params.gridBufSize):
        passes += 1
    passes = max(passes, 1)

    # an empty kernel stamped the same way, once per profiled substep: the
    # fixed part of every launch's event duration (reported, not subtracted)
    tare = timings.pop(TARE_TRACKER, None)
    tare_ms = tare["sum_ms"] / tare["calls"] if tare and tare["calls"] else None
    fam_ms: dict[str, float] = {}
    for name, t in timings.items():
        fam_ms[family(name)] = fam_ms.get(family(name), 0.0) + t["sum_ms"]
    # dominant kernel: the family with the most time among those with an
    # algorithmic byte figure (the friends build moves index structure only)
    measured = {k: v for k, v in fam_ms.items() if kernel_bytes_per_particle(k, iters, passes) is not None}
    dom = max(measured, key=measured.get)
    top = max(fam_ms, key=fam_ms.get)
    launches = [(name, t) for name, t in timings.items() if family(name) == dom]
    tot_ms = sum(t["sum_ms"] for _, t in launches)
    tot_calls = sum(t["calls"] for _, t in launches)
    bytes_per_launch = sum(kernel_bytes_per_particle(name, iters, passes) * n * t["calls"]
                           for name, t in launches) / max(tot_calls, 1)
    avg_ms = tot_ms / max(tot_calls, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    step_ms = {k: round(v / prof_steps, 4) for k, v in sorted(fam_ms.items(), key=lambda kv: -kv[1])}

    # the same figures for every family with a byte model (the friends build
    # with its first computeScaling included), with the HBM rate that the PMC
    # traffic gives at the measured launch time
    by_kernel = {}
    for fam_name in [f for f in fam_ms if kernel_bytes_per_particle(f, iters, passes, friends=True) is not None]:
        ls = [(name, t) for name, t in timings.items() if family(name) == fam_name]
        calls = sum(t["calls"] for _, t in ls)
        ms = sum(t["sum_ms"] for _, t in ls) / max(calls, 1)
        alg = sum(kernel_bytes_per_particle(name, iters, passes, friends=True) * n * t["calls"]
                  for name, t in ls) / max(calls, 1)
        traffic = pmc_traffic(fam_name, n)
        by_kernel[fam_name] = {
            "avg_launch_ms": round(ms, 5),
            "algorithmic_bytes_per_launch": alg,
            "achieved": alg / (ms * 1e-3) / 1e9,
            "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_gbs": traffic / (ms * 1e-3) / 1e9 if traffic else None,
            "traffic_frac": traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
        }

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
            "substeps_before_timing": args.warmup + CAPTURE_STEPS,
            "workgroup_sizes": sizes,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pmc_traffic(dom, n),
            "traffic_source_sha16": PMC_TRAFFIC["source_sha16"] or None,
            "kernel_source_sha16": kernel_source_sha16(),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
            "empty_kernel_ms": tare_ms,
            "top_kernel_overall": top,
            "by_kernel": by_kernel,
        },
        "stage_ms_per_step": step_ms,
        "error_flags": flags,
        "dropped_friends": dropped,
    }
    if not args.no_full_run:
        out["full_run"] = full_run(scene, sizes if not args.no_autotune else None)
        _progress(f"full run: {out['full_run']['particle_substeps_per_s']:.4g} particle-substeps/s over 500 "
                  f"substeps, {out['full_run']['dropped_friends']} friends dropped")
    if not args.no_batched:
        out["batched"] = batched_gain(args.batched_steps, 5, args.seed, not args.no_autotune)
        out["batched_gain"] = out["batched"]["gain"]
        _progress(f"batched gain {out['batched_gain']:.3f} "
                  f"(against the overlapped loop {out['batched']['gain_vs_overlapped_loop']:.3f})")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_ratio(args.seed)
        out["adjoint_ratio"] = out["adjoint"]["ratio"]
        _progress(f"adjoint/forward {out['adjoint_ratio']:.3f}")
        out["batched_adjoint"] = batched_adjoint(args.seed)
        _progress(f"batched adjoint/forward {out['batched_adjoint']['ratio']:.3f}")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
        _progress("cpu baseline done")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
