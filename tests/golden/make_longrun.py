"""Writes tests/golden/longrun_patchy4096.npz: statistics of a long run of the
reference's own Condensate::run loop (oracle/_ref/mtref long_rotor: rand()
noise, 3.5 Verlet list rebuilt every 20 steps) on the 4096-particle
tetrahedral patchy fluid (packing fraction 0.1, eps_p 6, kT 1, dt 0.005),
single-threaded (the reference's OpenMP force scatter races).

Run from the repository root:  python tests/golden/make_longrun.py
"""
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle as po  # noqa: E402
import longrun_stats as ls  # noqa: E402
from mtcode_amd import workloads as wl  # noqa: E402

NSTEPS, EVERY, WINDOW = 20000, 250, 10000


def main(nsteps=NSTEPS):
    s = wl.patchy_tetra(4096)
    N, L = s["N"], float(s["box"][0])
    e, d, th = s["patches"]["params"][0]
    job = {"N": N, "L": repr(L), "periodic": 1, "n": int(L // 4), "cut": 3.5,
           "wca_eps": 1.0, "wca_sigma": 1.0, "dt": s["dt"], "gamma": s["gamma"],
           "gammar": repr(float(s["gammar"])), "kT": s["kT"], "nsteps": nsteps, "every": EVERY,
           "seed": 12345, "patch_kind": "tetra", "eps_p": repr(float(e)), "d": repr(float(d)),
           "thetam": repr(float(th)), "maxd2": 2.4}
    with tempfile.TemporaryDirectory() as wd:
        t0 = time.time()
        po.run_ref("long_rotor", wd, job, {"x": s["x"], "Q": s["Q"]}, threads=1, timeout=7200)
        print(f"reference run {time.time() - t0:.0f} s", flush=True)
        nf = nsteps // EVERY
        x = po.read_ref(wd, "snap_x").reshape(nf, N, 3)
        p = po.read_ref(wd, "snap_p").reshape(nf, N, 3)
        Q = po.read_ref(wd, "snap_Q").reshape(nf, N, 9)
        Lm = po.read_ref(wd, "snap_L").reshape(nf, N, 3)
    first = WINDOW // EVERY  # frames after step WINDOW
    ser = ls.series(zip(x[first:], p[first:], Q[first:], Lm[first:]), L, s["patches"]["dirs"])
    out = {"nsteps": nsteps, "every": EVERY, "window": WINDOW, "edges": ls.EDGES}
    for k, v in ser.items():
        out[f"series_{k}"] = v
    # the whole time course of the mean bond count (equilibration)
    out["bond_mean_course"] = np.array(
        [ls.bonds(x[f], Q[f], L, s["patches"]["dirs"]).mean() for f in range(nf)])
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "longrun_patchy4096.npz"), **out)
    for k in ("bond_mean", "ke_t", "ke_r"):
        m, err = ls.mean_err(ser[k])
        print(k, m, "+-", err)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else NSTEPS)
