"""Summaries of rocprofv3 output for the step kernels, and the kernel-source
stamp that ties a recorded HBM-traffic figure to the kernels it was measured
on.

    python -m mtcode_amd.profiling trace  <kernel_trace.csv> [out.md]
    python -m mtcode_amd.profiling traffic <fetch_counter_collection.csv>
           <write_counter_collection.csv> <N> [out.json]

`trace` tabulates a `rocprofv3 --kernel-trace` CSV (calls, mean / min / max
duration per kernel).  `traffic` turns two `rocprofv3 --pmc FETCH_SIZE` and
`--pmc WRITE_SIZE` runs of `bench.py --copy-probe` into HBM bytes per step:
each counter (KiB per dispatch) is scaled by what it reads for the probe's
256 MiB device copy (256 MiB read, 256 MiB written), averaged over the
dispatches of each step kernel (the gated list kernels over all their
dispatches, so that rebuilds are amortised), and summed per phase.  The JSON
carries `source`, the stamp of the kernel sources it was measured on;
bench.py reports the traffic only while that stamp matches the tree.
"""
from __future__ import annotations

import csv
import hashlib
import json
import os
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
KERNEL_SOURCES = ("mdx_kernels.h", "mdx_device.h", "mdx_bind.h", "mdx_engine.hip")
COPY_BYTES = 256 << 20

# the kernels of one list step of the rotor integrator (mode 3: M_FUSED) and
# the phase each belongs to
PHASES = {
    "bins": ("k_bin<float>", "k_scan1", "k_scan3", "k_fill", "k_cellsort", "k_nbuild<float, 800>",
             "k_nbuild<float, 1152>"),
    "walk": ("k_walkc<float, 3, false>",),
    "patch": ("k_patch<float, 3, false>", "k_patch_ovf<float, 3>"),
    "finish": ("k_integrate<float, 3>", "k_decide"),
}


def source_stamp() -> str:
    """sha256 (first 16 hex digits) of the kernel sources"""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(HERE, "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def short_name(name: str) -> str:
    n = name.split("(")[0]
    for p in ("void mdx::", "mdx::", "void "):
        if n.startswith(p):
            n = n[len(p):]
    return n


def trace_summary(path: str) -> list:
    """[(kernel, calls, mean us, min us, max us, total ms)] by total time"""
    d = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            d[short_name(r["Kernel_Name"])].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    rows = []
    for k, v in d.items():
        rows.append((k, len(v), sum(v) / len(v), min(v), max(v), sum(v) / 1000.0))
    rows.sort(key=lambda r: -r[5])
    return rows


def trace_markdown(rows: list) -> str:
    tot = sum(r[5] for r in rows)
    out = ["| kernel | calls | avg us | min us | max us | total ms | % |",
           "|---|---|---|---|---|---|---|"]
    for k, n, avg, lo, hi, ms in rows:
        out.append(f"| `{k}` | {n} | {avg:.1f} | {lo:.1f} | {hi:.1f} | {ms:.3f} | "
                   f"{100 * ms / tot:.1f} |")
    return "\n".join(out)


def _counter(path: str, counter: str) -> dict:
    """{kernel: [KiB per dispatch]} of one counter"""
    d = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                d[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return d


def _copy_scale(d: dict) -> float:
    """true bytes over counted bytes, from the probe's device copy (the copy
    kernel dispatch with the largest count)"""
    best = 0.0
    for k, v in d.items():
        if "copy" in k.lower():
            best = max(best, max(v))
    if best <= 0:
        raise SystemExit("no copy-kernel dispatch found: run bench.py with --copy-probe")
    return COPY_BYTES / (best * 1024.0)


def traffic(fetch_csv: str, write_csv: str, n: int) -> dict:
    fetch, write = _counter(fetch_csv, "FETCH_SIZE"), _counter(write_csv, "WRITE_SIZE")
    fs, ws = _copy_scale(fetch), _copy_scale(write)
    per_kernel, per_phase = {}, {}
    for ph, ks in PHASES.items():
        tot = 0.0
        for k in ks:
            if k not in fetch and k not in write:
                continue
            fb = 1024.0 * fs * (sum(fetch.get(k, [0.0])) / max(1, len(fetch.get(k, []))))
            wb = 1024.0 * ws * (sum(write.get(k, [0.0])) / max(1, len(write.get(k, []))))
            per_kernel[k] = {"fetch_bytes": fb, "write_bytes": wb,
                             "dispatches": len(fetch.get(k, []))}
            tot += fb + wb
        per_phase[ph] = tot
    return {
        "N": n,
        "source": source_stamp(),
        "bytes_per_step": sum(per_phase.values()),
        "per_phase_bytes": per_phase,
        "per_kernel": per_kernel,
        "fetch_scale": fs,
        "write_scale": ws,
        "note": ("rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch), averaged over each "
                 "step kernel's dispatches (gated list kernels over all of theirs) and summed "
                 "per list step, scaled by the probe's 256 MiB device copy"),
    }


def main(argv: list) -> None:
    if len(argv) >= 2 and argv[0] == "trace":
        md = trace_markdown(trace_summary(argv[1]))
        if len(argv) > 2:
            with open(argv[2], "w") as f:
                f.write(md + "\n")
        print(md)
    elif len(argv) >= 4 and argv[0] == "traffic":
        t = traffic(argv[1], argv[2], int(argv[3]))
        js = json.dumps(t, indent=1)
        if len(argv) > 4:
            with open(argv[4], "w") as f:
                f.write(js + "\n")
        print(js)
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
