"""Benchmark of the Langevin step engine on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload condensate_ab]
                [--no-cpu] [--no-fp64]

Workload (default): the 1,048,576-particle two-species patchy condensate
(packing fraction 0.05, A: 3 patches, B: 2, A-B eps_p=8 + WCA, kT=1,
dt=0.005) in mixed precision.  One step = one Langevin step of every particle
(binning, 27-cell pair loop with WCA + patches + torques, rotor update), state
resident in HBM.  Prints one JSON line.  Also reported: the same engine in
FP64 (fp64_gpu_value), the plain FP64 GPU form (plain_fp64_gpu_value: one
kernel per step walking the 27 cells of every particle, no neighbour list),
the mixed-vs-FP64 force error on the same frozen
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
# the engine's re-sort period (list steps between re-sorts; MDX_RESORT)
RESORT_EVERY = int(os.environ.get("MDX_RESORT", "200"))

# Algorithmic bytes per particle-step of the workload (a fixed figure of the
# problem, not of this engine's storage): one read and one write of every
# per-particle state array the step updates, at the precision the workload
# specifies, plus one read of the static per-particle data.
#   patchy, mixed: x split FP64+FP32 36 + p 12 + Q (3x3 FP32) 36 + L 12 = 96 B
#                  state -> 2*96 + type 4 = 196
#   patchy, fp64 : x 24 + p 24 + Q 72 + L 24 = 144 B state -> 2*144 + 4 = 292
#   beads (bundle): x 36 (mixed split) + p 12 = 48 B state -> 2*48 + type 4 +
#                  bond partners 16 + angle partners 8 = 124
#   active polymer: bead state 48 B (2*48), its share of end-rotor state
#                  (2/256 of 96 B), type 4 + chain neighbours 8 = 109
# The engine stores less (x 24, p 12, Q as a 16 B quaternion, L 12: 64 B per
# rotor), so its own traffic can sit below these figures.
#   binding (condensate_bind): the patchy figure plus the bound state, 8 B
#                  per patch (isbound, boundto) read and written, 2.5 patches
#                  per particle -> +40
ALGO_BYTES = {"mixed": 196, "fp64": 292}
BIND_BYTES = {"mixed": 236, "fp64": 332}
BEAD_BYTES = {"mixed": 124, "fp64": 124}
POLYMER_BYTES = {"mixed": 109, "fp64": 109}
# kernels of one step and the phases of mdx_profile_step that time them
STEP_PHASES = ("bins", "walk", "patch", "finish")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_condensate_ab_1M_mixed.json")


def measured_traffic():
    """HBM bytes per step of each step kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-probe scaled; see
    mtcode_amd/profiling.py), or None -- also when it was measured on other
    kernel sources than these (its source stamp differs)"""
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        return None
    from mtcode_amd import profiling

    if tr.get("source") != profiling.source_stamp():
        log("recorded traffic is stale (measured on other kernel sources): not reported")
        return None
    return tr


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


WORKLOADS = {
    "condensate_ab": ("1,048,576 two-species patchy condensate: A 3 planar patches / B 2 "
                      "patches, A-B eps_p=8, d=1.2, cos(theta_m)=0.95, WCA eps=sigma=1 "
                      "rc=1.12246, packing fraction 0.05, kT=1, gamma=1, gamma_r=1/3, dt=0.005"),
    "patchy_tetra": ("4096 tetrahedral patchy particles, eps_p=6, d=1.2, cos(theta_m)=0.95, WCA, "
                     "packing fraction 0.1, kT=1, gamma=1, gamma_r=1/3, dt=0.005"),
    "replicas": ("256 independent tetrahedral patchy replicas, N log-uniform in [512, 16384] "
                 "(1.17M total), kT ~ U[0.8,1.2], eps_p ~ U[4,10], packing fraction 0.1, "
                 "one engine"),
    "mt_bundle": ("416,000-bead microtubule bundle: 64 tubes x 13 protofilaments x 500 beads, "
                  "harmonic bonds k=100 r0=1, cosine bending kappa=50, WCA between non-bonded "
                  "beads, box 200x200x600, dt=0.002"),
    "condensate_bind": ("the condensate_ab system under the one-bond-per-patch binding model: "
                        "BindingModelSingle(on 0.99999, off 1e-5) (mainPatchy.cpp, "
                        "run_singlebond), WCA plus the bound pairs' Kern-Frenkel terms, bound "
                        "state updated every step"),
    "active_polymer": ("2,097,152-bead active polymer melt: 8192 chains x 256 beads, harmonic "
                       "k=200, cosine bending kappa=20, tangential f_a=1, patchy rotor end "
                       "beads (2 patches, eps_p=5), packing fraction 0.15"),
}


def build_workload(name):
    """(system dict or None, list of replica dicts or None), with momenta
    drawn at each system's temperature, so that the timed steps run at the
    steady temperature (and list-rebuild rate) from the start"""
    if name == "replicas":
        reps = [wl.thermal_momenta(r, seed=99 + k) for k, r in
                enumerate(wl.replicas(progress=log))]
        return reps[0], reps
    if name == "condensate_ab":
        s = wl.condensate_ab()
    elif name == "condensate_bind":
        s = wl.condensate_bind()
    elif name == "patchy_tetra":
        s = wl.patchy_tetra()
    elif name == "mt_bundle":
        s = wl.mt_bundle()
    elif name == "active_polymer":
        s = wl.active_polymer(progress=log)
    else:
        raise SystemExit(f"unknown workload {name}")
    return wl.thermal_momenta(s), None


def force_error(s, reps=None):
    """mixed vs FP64 conservative forces/torques on the same frozen
    configuration: the FP64 engine gets the orientations the mixed engine's
    kernels use (its FP32 unit quaternions as matrices)"""
    def rounded(x):
        x = dict(x)
        if x.get("Q") is not None:
            x["Q"] = wl.mixed_orientations(x["Q"])
        return x
    s64 = rounded(s)
    reps64 = [rounded(r) for r in reps] if reps else None
    with mdx.engine_from_system(s64, precision=mdx.PREC_FP64, replicas=reps64) as e:
        F64, T64 = e.compute_forces()
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, replicas=reps) as e:
        F32, T32 = e.compute_forces()
    out = {
        "force_rel_err_max": float(np.abs(F32 - F64).max() / np.abs(F64).max()),
        "force_rel_err_rms": float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64)),
    }
    if np.abs(T64).max() > 0:
        out["torque_rel_err_max"] = float(np.abs(T32 - T64).max() / np.abs(T64).max())
    return out


def time_engine(s, precision, steps, warmup, seed=1, reps=None, autotune=False):
    with mdx.engine_from_system(s, precision=precision, seed=seed, replicas=reps) as e:
        if s["integrator"] == "rotor":
            e.prime()
        if autotune:
            log("autotune", e.autotune(100))
        if warmup:
            e.step(warmup)
        t0 = time.perf_counter()
        e.step(steps)
        t1 = time.perf_counter()
        prof = e.profile(min(steps, 20))
        info = e.info()
        info["resort_ms"] = resort_cost(e, info) if info["skin"] > 0 else None
    return (t1 - t0), prof, info


def resort_cost(e, info, reps=3):
    """extra wall time of a re-sort step (state gathered into cell order,
    list rebuilt; every RESORT_EVERY list steps) over a list step: two-step
    calls with and without an invalidated list, best of reps"""
    def two():
        t = time.perf_counter()
        e.step(2)
        return time.perf_counter() - t
    plain, resort = [], []
    for _ in range(reps):
        plain.append(two())
        e.set_neighbor_list(info["skin"], info["kmax"], RESORT_EVERY)
        resort.append(two())
    return 1e3 * max(0.0, min(resort) - min(plain))


def time_plain_fp64(s, steps, warmup, seed=1, reps=None):
    """the plain FP64 GPU form of the step: every step bins the particles and
    walks the 27 cells in one kernel (walk, patches and integrator together),
    no neighbour list, no walk / patch / integrator split, FP64 throughout"""
    old = os.environ.get("MDX_SPLIT")
    os.environ["MDX_SPLIT"] = "0"
    try:
        with mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=seed, replicas=reps) as e:
            e.set_neighbor_list(0.0)
            if s["integrator"] == "rotor":
                e.prime()
            e.step(warmup)
            t0 = time.perf_counter()
            e.step(steps)
            t1 = time.perf_counter()
    finally:
        if old is None:
            del os.environ["MDX_SPLIT"]
        else:
            os.environ["MDX_SPLIT"] = old
    return t1 - t0


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
           "kT": s["kT"], "nsteps": nsteps, "maxd2": 2.4}
    if list(ps["npatch"]) == [4]:
        e, d, th = ps["params"][0]
        job.update({"patch_kind": "tetra", "eps_p": repr(float(e)), "d": repr(float(d)),
                    "thetam": repr(float(th))})
    else:
        job.update({"patch_kind": "general", "ntypes": len(ps["npatch"])})
        cum = 0
        for t, npt in enumerate(ps["npatch"]):
            cum += int(np.sum(ps["types"] == t))
            job[f"npatch_{t}"] = int(npt)
            job[f"numpertype_{t}"] = cum
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


def copy_probe(nbytes=1 << 28, reps=3):
    """a plain device copy of known size (for scaling rocprof byte counters);
    returns the copy's rate in GB/s (bytes read + written over the mean
    duration, HIP events), or None"""
    try:
        import torch

        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps):
            b.copy_(a)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / reps
        del a, b
        return 2 * nbytes / (ms * 1e-3) / 1e9
    except Exception as exc:  # pragma: no cover - diagnostics only
        print(f"copy probe skipped: {exc}", file=sys.stderr)
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="condensate_ab")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-steps-1core", type=int, default=8,
                    help="steps of the one-thread CPU sample (0 skips it)")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="let the engine's autotuner pick skin and walk form before timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path "
              "(replicas would shard whole per GPU, a large system by z-slabs)", file=sys.stderr)
        sys.exit(2)
    # a plain device copy: the practical bandwidth beside the 8 TB/s peak
    copy_gbs = copy_probe(reps=3 if args.copy_probe else 5)

    log("building workload", args.workload)
    s, reps = build_workload(args.workload)
    N = sum(r["N"] for r in reps) if reps else s["N"]
    rotor = s["integrator"] == "rotor"
    algo = (BIND_BYTES if s.get("binding") else ALGO_BYTES if rotor else
            POLYMER_BYTES if args.workload == "active_polymer" else BEAD_BYTES)
    log("timing mixed precision", args.steps, "steps after", args.warmup, "warm-up")
    t_mixed, prof, info = time_engine(s, mdx.PREC_MIXED, args.steps, args.warmup, reps=reps,
                                      autotune=args.autotune)
    log("mixed ms/step", 1e3 * t_mixed / args.steps, "kernels", prof)
    value = N * args.steps / t_mixed
    step_ms = sum(prof.get(k, 0.0) for k in STEP_PHASES + ("gather",) if k in prof
                  and not (k == "gather" and "walk" in prof))
    achieved = algo["mixed"] * N / (step_ms * 1e-3) / 1e9
    tr = measured_traffic()
    traffic = None
    per_kernel = {k: {"ms": prof[k]} for k in STEP_PHASES if k in prof}
    if tr and tr.get("N") == N and args.workload == "condensate_ab":
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
        "data": ("synthetic (seeded uniform positions with overlap rejection, Arvo random rotations, "
                 "Maxwell-Boltzmann momenta at kT)"),
        "config": {
            "workload": f"{args.workload}: N={N}, " + WORKLOADS[args.workload],
            "precision": "mixed",
            "neighbour_list_skin": info["skin"],
            "cells": info["cells"],
            "cell_size": round(info["cell_size"], 4),
            "resort_every": RESORT_EVERY,
        },
        # the state is gathered into cell order every RESORT_EVERY list steps;
        # a timed window of that many steps or more contains one such step
        "resort": {
            "extra_ms": info.get("resort_ms"),
            "amortised_ms_per_step": (info["resort_ms"] / RESORT_EVERY
                                      if info.get("resort_ms") is not None else None),
            # (re-sort steps: step 1 after prime, then every RESORT_EVERY + 1)
            "in_timed_window": (args.warmup + args.steps - 1) // (RESORT_EVERY + 1)
                               - (args.warmup - 1) // (RESORT_EVERY + 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": ("one step: k_walkc (WCA) + k_bind_edges + k_bind_decide + k_bind_forces "
                       "+ k_integrate + list upkeep (gated; amortised) + k_decide"
                       if s.get("binding") else
                       "one step: k_walkc + k_patch + k_integrate + list upkeep (k_bin, k_scan1, "
                       "k_scan3, k_fill, k_cellsort, k_nbuild, gated; amortised) + k_decide"
                       if "walk" in prof else
                       "one step: k_gather + list upkeep (gated; amortised) + k_decide"),
            "algorithmic_bytes_per_particle_step": algo["mixed"],
            "step_kernel_ms": step_ms,
            "per_kernel": per_kernel,
            "copy_probe_gbs": copy_gbs,
        },
        "kernel_ms": prof,
    }
    if not args.no_fp64:
        fsteps = max(10, args.steps // 10)
        log("timing fp64", fsteps, "steps")
        t64, prof64, _ = time_engine(s, mdx.PREC_FP64, fsteps, 2, reps=reps)
        log("fp64 ms/step", 1e3 * t64 / fsteps)
        v64 = N * fsteps / t64
        out["fp64_gpu_value"] = v64
        out["mixed_over_fp64"] = value / v64
        out["fp64_kernel_ms"] = prof64
        tp = time_plain_fp64(s, fsteps, 2, reps=reps)
        log("plain fp64 ms/step", 1e3 * tp / fsteps)
        out["plain_fp64_gpu_value"] = N * fsteps / tp
        out["mixed_over_plain_fp64"] = value * tp / (N * fsteps)
        out.update(force_error(s, reps))
        log("force error", {k: out[k] for k in out if "err" in k})
    if not args.no_cpu and args.workload in ("condensate_ab", "patchy_tetra"):
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline (reference harness),", threads, "threads")
        cb = cpu_baseline(s, args.cpu_steps, threads)
        log("cpu baseline", cb)
        if cb is not None and args.cpu_steps_1core > 0 and threads != 1:
            log("cpu baseline (reference harness), 1 thread")
            c1 = cpu_baseline(s, args.cpu_steps_1core, 1)
            log("cpu baseline 1 thread", c1)
            cb["one_core"] = {k: c1[k] for k in ("value", "cores", "sample")}
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
