"""Benchmark of the Langevin step engine on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload condensate_ab]
                [--no-cpu] [--no-fp64]

Workload (default): the 1,048,576-particle two-species patchy condensate
(packing fraction 0.05, A: 3 patches, B: 2, A-B eps_p=8 + WCA, kT=1,
dt=0.005) in mixed precision.  One step = one Langevin step of every particle
(binning, 27-cell pair loop with WCA + patches + torques, rotor update), state
resident in HBM.  Prints one JSON line.  Also reported: the plain FP64 GPU
form on the same system, the mixed-vs-FP64 force error on the same frozen
configuration, the step's HBM roofline fraction (kernel times from HIP events
on the engine's stream; measured HBM traffic per step from the rocprofv3
counter summary in profiles/ when present), and the reference's own CPU path
(oracle/_ref) timed on a bounded sample of the same workload on this host's
cores.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mtcode_amd import mdx  # noqa: E402
from mtcode_amd import workloads as wl  # noqa: E402

METRIC = "particle-steps/s at 1M patchy particles (mixed vs FP64 GPU); force rel err vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle-step: one read and one write of every
# per-particle state array the step updates, at the stored precision, plus one
# read of the static per-particle type word.
#   mixed: x FP64 24 + p FP32 12 + Q FP32 36 + L FP32 12 = 84 B state -> 2*84 + 4
#   fp64 : x 24 + p 24 + Q 72 + L 24 = 144 B state -> 2*144 + 4
ALGO_BYTES = {"mixed": 2 * 84 + 4, "fp64": 2 * 144 + 4}
# kernels of one step and the phases of mdx_profile_step that time them
STEP_PHASES = ("bins", "walk", "patch", "finish")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_condensate_ab_1M_mixed.json")


def measured_traffic():
    """HBM bytes per step of each step kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-probe scaled), or None"""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(name):
    if name == "condensate_ab":
        return wl.condensate_ab()
    if name == "patchy_tetra":
        return wl.patchy_tetra()
    raise SystemExit(f"unknown workload {name}")


def force_error(s):
    """mixed vs FP64 conservative forces/torques on the same frozen
    configuration (orientations rounded to FP32, as the mixed engine holds them)"""
    s = dict(s)
    s["Q"] = s["Q"].astype(np.float32).astype(np.float64)
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64) as e:
        F64, T64 = e.compute_forces()
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED) as e:
        F32, T32 = e.compute_forces()
    return {
        "force_rel_err_max": float(np.abs(F32 - F64).max() / np.abs(F64).max()),
        "force_rel_err_rms": float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64)),
        "torque_rel_err_max": float(np.abs(T32 - T64).max() / max(np.abs(T64).max(), 1e-300)),
    }


def time_engine(s, precision, steps, warmup, seed=1):
    with mdx.engine_from_system(s, precision=precision, seed=seed) as e:
        e.prime()
        if warmup:
            e.step(warmup)
        t0 = time.perf_counter()
        e.step(steps)
        t1 = time.perf_counter()
        prof = e.profile(min(steps, 20))
        info = e.info()
    return (t1 - t0), prof, info


def cpu_baseline(s, nsteps, threads):
    """the reference's own Condensate::run loop (oracle/_ref/mtref) on a
    bounded sample: full system, nsteps steps after its pair-list set-up"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    if not po.have_ref():
        return None
    ps = s["patches"]
    job = {"N": s["N"], "L": repr(float(s["box"][0])), "periodic": 1,
           "n": int(s["box"][0] // 4), "cut": 3.5, "wca_eps": 1.0, "wca_sigma": 1.0,
           "dt": s["dt"], "gamma": s["gamma"], "gammar": repr(float(s["gammar"])),
           "kT": s["kT"], "nsteps": nsteps, "maxd2": 2.4, "patch_kind": "general",
           "ntypes": 2, "npatch_0": 3, "npatch_1": 2,
           "numpertype_0": int(np.sum(ps["types"] == 0)), "numpertype_1": s["N"]}
    with tempfile.TemporaryDirectory() as d:
        out = po.run_ref("time_rotor", d, job, {"x": s["x"], "Q": s["Q"],
                                                "params": ps["params"], "dirs": ps["dirs"]},
                         threads=threads, timeout=900)
    r = json.loads(out.strip().splitlines()[-1])
    return {
        "value": r["N"] * r["steps"] / r["loop_s"],
        "unit": "particle-steps/s",
        "cores": int(r["threads"]),
        "kind": "reference",
        "sample": (f"{r['N']} particles, {r['steps']} steps of the reference's Condensate::run "
                   f"loop (3.5 Verlet list, OpenMP, -O3) after its {r['setup_s']:.1f} s pair-list "
                   f"set-up"),
    }


def copy_probe(nbytes=1 << 28):
    """a plain device copy of known size, for scaling rocprof byte counters"""
    try:
        import torch

        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    except Exception as exc:  # pragma: no cover - diagnostics only
        print(f"copy probe skipped: {exc}", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="condensate_ab")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--copy-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path in this build "
              "(see DESIGN.md for how replicas and large systems would shard)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()

    log("building workload", args.workload)
    s = build_workload(args.workload)
    N = s["N"]
    log("timing mixed precision", args.steps, "steps after", args.warmup, "warm-up")
    t_mixed, prof, info = time_engine(s, mdx.PREC_MIXED, args.steps, args.warmup)
    log("mixed ms/step", 1e3 * t_mixed / args.steps, "kernels", prof)
    value = N * args.steps / t_mixed
    step_ms = sum(prof.get(k, 0.0) for k in STEP_PHASES)
    achieved = ALGO_BYTES["mixed"] * N / (step_ms * 1e-3) / 1e9
    tr = measured_traffic()
    traffic = None
    per_kernel = {k: {"ms": prof[k]} for k in STEP_PHASES if k in prof}
    if tr and tr.get("N") == N:
        traffic = tr["bytes_per_step"]
        for k, v in tr.get("per_phase_bytes", {}).items():
            if k in per_kernel:
                per_kernel[k]["traffic_bytes"] = v
                per_kernel[k]["achieved_gbs"] = v / (per_kernel[k]["ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 positions and decisions)",
        "data": "synthetic (seeded uniform positions with overlap rejection, Arvo random rotations)",
        "config": {
            "workload": f"{args.workload}: N={N}, box={float(s['box'][0]):.3f}, "
                        "A 3 planar patches / B 2 patches, A-B eps_p=8, d=1.2, cos(theta_m)=0.95, "
                        "WCA eps=sigma=1 rc=1.12246, kT=1, gamma=1, gamma_r=1/3, dt=0.005",
            "precision": "mixed",
            "cells": info["cells"],
            "cell_size": round(info["cell_size"], 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "one step: k_walk + k_patch + k_integrate + list upkeep (k_bin, k_scan1-3, "
                      "k_fill, k_cellsort, k_nbuild, gated; amortised)",
            "algorithmic_bytes_per_particle_step": ALGO_BYTES["mixed"],
            "step_kernel_ms": step_ms,
            "per_kernel": per_kernel,
        },
        "kernel_ms": prof,
    }
    if not args.no_fp64:
        fsteps = max(10, args.steps // 10)
        log("timing fp64", fsteps, "steps")
        t64, prof64, _ = time_engine(s, mdx.PREC_FP64, fsteps, 2)
        log("fp64 ms/step", 1e3 * t64 / fsteps)
        v64 = N * fsteps / t64
        out["fp64_gpu_value"] = v64
        out["mixed_over_fp64"] = value / v64
        out["fp64_kernel_ms"] = prof64
        out.update(force_error(s))
        log("force error", {k: out[k] for k in out if "err" in k})
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline (reference harness),", threads, "threads")
        cb = cpu_baseline(s, args.cpu_steps, threads)
        log("cpu baseline", cb)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
