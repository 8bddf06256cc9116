"""Writes the golden fixtures of tests/golden/ by running the reference's own
C++ classes (oracle/_ref/mtref, built by ``make -C oracle ref`` from the
reference sources) on seeded inputs.  Single-threaded (OMP_NUM_THREADS=1):
the reference's OpenMP force scatter is racy.

Run from the repository root:  python tests/golden/make_golden.py
Each fixture is an .npz of inputs and the reference's outputs.
"""
from __future__ import annotations

import math
import os
import shutil
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from mtcode_amd import workloads as wl  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
WORK = "/tmp/mtcode_golden"


def job_dir(name):
    d = os.path.join(WORK, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    return d


def patch_job(s):
    ps = s["patches"]
    if ps is None:
        return {"patch_kind": "none"}
    if len(ps["npatch"]) == 1 and ps["npatch"][0] == 4:
        e, d, th = ps["params"][0]
        return {"patch_kind": "tetra", "eps_p": repr(float(e)), "d": repr(float(d)),
                "thetam": repr(float(th)), "maxd2": repr(float(ps["maxd2"]))}
    job = {"patch_kind": "general", "ntypes": len(ps["npatch"]),
           "maxd2": repr(float(ps["maxd2"]))}
    types = ps["types"]
    cum = 0
    for t, npt in enumerate(ps["npatch"]):
        cnt = int(np.sum(types == t))
        # GeneralPatch needs ascending contiguous type ranges (GeneralPatch.cpp:268-282)
        assert np.all(types[cum:cum + cnt] == t)
        cum += cnt
        job[f"npatch_{t}"] = npt
        job[f"numpertype_{t}"] = cum
    return job


def fixture_cells():
    s = wl.patchy_tetra()
    L = float(s["box"][0])
    x = s["x"].copy()
    # edge cases: 0, L, L - ulp, exact cell boundaries, tiny negatives
    n = 16
    extra = [
        [0.0, 0.0, 0.0],
        [L, L, L],
        [np.nextafter(L, 0), np.nextafter(L, 0), 0.5],
        [L / n, 2 * L / n, 15 * L / n],
        [-1e-12, -0.0, L * 0.5],
        [np.nextafter(L / n, 0), np.nextafter(L / n, 2 * L), 3 * L / n],
    ]
    x = np.vstack([x, np.array(extra)])
    out = {"x": x, "L": L}
    for nn in [1, 2, 3, 6, 16]:
        d = job_dir(f"cells{nn}")
        po.run_ref("cells", d, {"N": len(x), "L": repr(L), "periodic": 1, "n": nn}, {"x": x})
        out[f"cells_n{nn}"] = po.read_ref(d, "cells", np.int32)
        shape = po.read_ref(d, "boxes_shape", np.int32)
        boxes = po.read_ref(d, "boxes", np.int32).reshape(shape)
        if nn <= 6:
            out[f"boxes_n{nn}"] = boxes
        else:
            out[f"boxes_n{nn}_sum"] = np.int64(boxes.astype(np.int64).sum())
            out[f"boxes_n{nn}_row_hash"] = (boxes.astype(np.int64) *
                                            np.arange(1, boxes.shape[1] + 1)).sum(1)
    np.savez_compressed(os.path.join(OUT, "cells_patchy4096.npz"), **out)


def _pairs(x, L, n, cut, periodic=1):
    d = job_dir(f"pairs_{n}_{cut}")
    po.run_ref("pairs", d, {"N": len(x), "L": repr(L), "periodic": periodic, "n": n,
                            "cut": repr(cut)}, {"x": x})
    return po.read_ref(d, "pairs", np.int32).reshape(-1, 2)


def fixture_pairs():
    out = {}
    s = wl.patchy_tetra(512)
    L = float(s["box"][0])
    out["x512"] = s["x"]
    out["L512"] = L
    for (n, cut) in [(3, 3.5), (8, math.sqrt(2.4)), (2, 3.5), (1, 3.5), (9, 1.12246)]:
        out[f"pairs512_n{n}_cut{cut:.6f}"] = _pairs(s["x"], L, n, cut)
    # reflecting (non-periodic) box
    out["pairs512_n3_cut3.500000_nonperiodic"] = _pairs(s["x"], L, 3, 3.5, periodic=0)
    s = wl.patchy_tetra()
    L = float(s["box"][0])
    pr = _pairs(s["x"], L, 16, 1.68)
    norm = np.sort(np.sort(pr, axis=1).astype(np.int64) @ np.array([1, 1 << 20]))
    out["pairs4096_n16_cut1.68_count"] = np.int64(len(pr))
    out["pairs4096_n16_cut1.68_keys"] = norm
    np.savez_compressed(os.path.join(OUT, "pairs.npz"), **out)


def _forces(s, name, n, cut):
    d = job_dir(name)
    L = float(s["box"][0])
    job = {"N": s["N"], "L": repr(L), "periodic": 1, "n": n, "cut": repr(cut),
           "wca_eps": repr(float(s["wca"][0])), "wca_sigma": repr(float(s["wca"][1]))}
    job.update(patch_job(s))
    arrays = {"x": s["x"], "Q": s["Q"]}
    ps = s["patches"]
    if job["patch_kind"] == "general":
        arrays["params"] = ps["params"]
        arrays["dirs"] = ps["dirs"]
    po.run_ref("forces", d, job, arrays)
    N = s["N"]
    return {
        "x": s["x"], "Q": s["Q"], "L": L, "n": n, "cut": cut,
        "F_wca": po.read_ref(d, "F_wca").reshape(N, 3),
        "F": po.read_ref(d, "F").reshape(N, 3),
        "T": po.read_ref(d, "T").reshape(N, 3),
        "epot": po.read_ref(d, "epot")[0],
        "npairs": np.int64(len(po.read_ref(d, "pairs", np.int32)) // 2),
    }


def fixture_forces():
    s = wl.patchy_tetra()
    np.savez_compressed(os.path.join(OUT, "forces_patchy4096.npz"),
                        **_forces(s, "forces_tetra", int(s["box"][0] // 4), 3.5))
    s = wl.condensate_ab(4096)
    out = _forces(s, "forces_ab", int(s["box"][0] // 4), 3.5)
    out["types"] = s["patches"]["types"]
    np.savez_compressed(os.path.join(OUT, "forces_condensate4096.npz"), **out)


def _align(a, b):
    """rotation matrix R with R a = b (unit vectors)"""
    v = np.cross(a, b)
    c = float(np.dot(a, b))
    if c < -1 + 1e-12:
        axis = np.cross(a, [1.0, 0, 0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, [0, 1.0, 0])
        axis /= np.linalg.norm(axis)
        return 2 * np.outer(axis, axis) - np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx / (1 + c)


def _axis_rot(axis, ang):
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _tilt(u, ang, rng):
    """unit vector at angle ang from u, random azimuth"""
    w = rng.normal(size=3)
    w -= w.dot(u) * u
    w /= np.linalg.norm(w)
    return math.cos(ang) * u + math.sin(ang) * w


def pair_sweep_system(dirs, eps_p, npairs=512, seed=11):
    """Isolated particle pairs on a grid (spacing 8): pair k has its chosen
    patches pointing at each other within prescribed angles and distance,
    covering theta=0, the d1=0 zone, the cone edge, the maxd2 and 1.4d edges
    and the WCA cut-off."""
    rng = np.random.default_rng(seed)
    side = int(math.ceil(npairs ** (1 / 3)))
    L = 8.0 * side
    thm = wl.THETAM_095
    x = np.zeros((2 * npairs, 3))
    Q = np.zeros((2 * npairs, 9))
    for k in range(npairs):
        g = np.array([k % side, (k // side) % side, k // (side * side)], dtype=float) * 8.0 + 2.0
        mode = k % 8
        if mode == 0:
            ti, tj = 0.0, 0.0
        elif mode == 1:
            ti, tj = rng.uniform(0, 3e-3, 2)
        elif mode == 2:
            ti, tj = thm * (1 - 1e-7), rng.uniform(0, thm)
        elif mode == 3:
            ti, tj = rng.uniform(0, thm), thm * (1 + 1e-7)
        else:
            ti, tj = rng.uniform(0, 1.15 * thm, 2)
        rmode = (k // 8) % 6
        if rmode == 0:
            r = 1.12246 * (1 - 1e-9)
        elif rmode == 1:
            r = math.sqrt(2.4) * (1 - 1e-9)
        elif rmode == 2:
            r = math.sqrt(2.4) * (1 + 1e-9)
        elif rmode == 3:
            r = rng.uniform(0.92, 1.12)
        else:
            r = rng.uniform(0.95, 1.62)
        ia = rng.integers(len(dirs))
        ib = rng.integers(len(dirs))
        RA = _axis_rot(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
        na = RA @ dirs[ia]
        dvec = _tilt(na, ti, rng)  # direction A -> B
        xa = g
        xb = g + r * dvec
        target = _tilt(-dvec, tj, rng)
        RB = _align(dirs[ib], target)
        RB = _axis_rot(target, rng.uniform(0, 2 * math.pi)) @ RB
        x[2 * k] = xa
        x[2 * k + 1] = xb
        # Q is lab -> body, i.e. the transpose of body -> lab
        Q[2 * k] = RA.T.ravel()
        Q[2 * k + 1] = RB.T.ravel()
    return x, Q, L


def fixture_pairsweep():
    out = {}
    for name, dirs, kind in [("tetra", wl.TETRA_DIRS, "tetra"),
                             ("bivalent", wl.BIVALENT_DIRS, "bivalent")]:
        x, Q, L = pair_sweep_system(dirs, 6.0)
        N = len(x)
        d = job_dir("sweep_" + name)
        job = {"N": N, "L": repr(L), "periodic": 1, "n": int(L // 4), "cut": 3.5,
               "wca_eps": 1.0, "wca_sigma": 1.0, "patch_kind": kind, "eps_p": 6.0,
               "d": 1.2, "thetam": repr(wl.THETAM_095), "maxd2": 2.4}
        po.run_ref("forces", d, job, {"x": x, "Q": Q})
        out[f"{name}_x"] = x
        out[f"{name}_Q"] = Q
        out[f"{name}_L"] = L
        out[f"{name}_F"] = po.read_ref(d, "F").reshape(N, 3)
        out[f"{name}_T"] = po.read_ref(d, "T").reshape(N, 3)
        out[f"{name}_F_wca"] = po.read_ref(d, "F_wca").reshape(N, 3)
    np.savez_compressed(os.path.join(OUT, "pairsweep.npz"), **out)


def chain_positions(nchains, nbeads, L, rng, step=1.0, jitter=0.1):
    x = np.zeros((nchains * nbeads, 3))
    for c in range(nchains):
        cur = rng.uniform(0, L, 3)
        for b in range(nbeads):
            if b > 0:
                for _ in range(1000):
                    v = rng.normal(size=3)
                    v /= np.linalg.norm(v)
                    cand = cur + v * (step + rng.uniform(-jitter, jitter))
                    placed = x[: c * nbeads + b]
                    dd = placed - cand
                    dd -= L * np.round(dd / L)
                    if len(placed) == 0 or np.min(np.sum(dd * dd, 1)) > 0.81:
                        cur = cand
                        break
            x[c * nbeads + b] = np.mod(cur, L)
    return x


def fixture_bonded():
    rng = np.random.default_rng(21)
    nch, nb = 3, 30
    L = 20.0
    x = chain_positions(nch, nb, L, rng)
    idx = np.arange(nch * nb).reshape(nch, nb)
    bonds = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1).astype(np.int32)
    angles = np.stack([idx[:, :-2].ravel(), idx[:, 1:-1].ravel(), idx[:, 2:].ravel()],
                      1).astype(np.int32)
    r0s = rng.uniform(0.9, 1.1, len(bonds))
    out = {"x": x, "L": L, "bonds": bonds, "angles": angles, "r0s": r0s}
    for kind, k, r0 in [("harmonic", 100.0, 1.0), ("fene", 30.0, 1.5), ("perbond", 100.0, 0.0)]:
        for th0 in [0.3, 0.0]:
            d = job_dir(f"bonded_{kind}_{th0}")
            job = {"N": len(x), "L": L, "periodic": 1, "nbonds": len(bonds),
                   "nangles": len(angles), "bond_kind": kind, "bond_k": k, "bond_r0": r0,
                   "angle_k": 50.0, "angle_theta0": th0}
            po.run_ref("bonded", d, job, {"x": x, "bonds": bonds, "angles": angles,
                                          "bond_r0s": r0s})
            out[f"Fb_{kind}"] = po.read_ref(d, "Fb").reshape(-1, 3)
            out[f"Fa_theta0_{th0}"] = po.read_ref(d, "Fa").reshape(-1, 3)
    np.savez_compressed(os.path.join(OUT, "bonded.npz"), **out)


def fixture_calcg1():
    """calcg1 histograms of the reference (Analysis/PairCorrelation.cpp) at
    4, 2 and 1 cells per side"""
    out = {}
    cases = [("c4096", wl.patchy_tetra(4096), 4.0, 0.05),
             ("c512", wl.patchy_tetra(512), 6.0, 0.1),
             ("c512n1", wl.patchy_tetra(512), 13.0, 0.25)]
    for name, s, mx, dx in cases:
        d = job_dir(f"calcg1_{name}")
        L = float(s["box"][0])
        po.run_ref("calcg1", d, {"N": s["N"], "L": repr(L), "periodic": 1, "max": repr(mx),
                                 "dx": repr(dx)}, {"x": s["x"]})
        out[f"{name}_x"] = s["x"]
        out[f"{name}_L"] = L
        out[f"{name}_max"] = mx
        out[f"{name}_dx"] = dx
        out[f"{name}_counts"] = po.read_ref(d, "counts", dtype=np.int64)
    np.savez_compressed(os.path.join(OUT, "calcg1.npz"), **out)


def fixture_dgraph():
    """distance_graph labels of the reference (Analysis/AnalyzeGrowth.cpp:
    136-177: calculatepairs_parallel + ConnectedComponentsParallel_Old), at the
    1.25 binding distance of Condensate::run_singlebond, with and without the
    excludea index range, on 2 cells per side, and on bead chains"""
    out = {}
    u = wl.patchy_tetra(4096, phi=0.15, seed_pos=21)
    p = wl.active_polymer(16, 48, phi=0.15, seed=9)
    cases = [("u4096", u["x"], float(u["box"][0]), 1.25, 0, -1),
             ("u4096ex", u["x"], float(u["box"][0]), 1.25, 1000, 2999),
             ("u4096c16", u["x"], float(u["box"][0]), 1.6, 0, -1),
             ("n2", wl.patchy_tetra(150, phi=0.15, seed_pos=22)["x"],
              float(wl.box_for_packing(150, 0.15)), 1.25, 0, -1),
             ("chains", p["x"], float(p["box"][0]), 1.05, 0, -1)]
    for name, x, L, cut, lo, hi in cases:
        n = max(int(np.floor(L / 3.5)), 1) if name != "n2" else 2
        d = job_dir(f"dgraph_{name}")
        po.run_ref("dgraph", d, {"N": len(x), "L": repr(L), "periodic": 1, "n": n,
                                 "cut": repr(cut), "ex_lo": lo, "ex_hi": hi}, {"x": x})
        out[f"{name}_x"] = x
        out[f"{name}_L"] = L
        out[f"{name}_n"] = n
        out[f"{name}_cut"] = cut
        out[f"{name}_ex"] = np.array([lo, hi])
        out[f"{name}_labels"] = po.read_ref(d, "labels", dtype=np.int32)
    np.savez_compressed(os.path.join(OUT, "dgraph.npz"), **out)


def fixture_csv():
    """the reference's CSV text of a matrix (operator<<=, matrix2.cpp:898-912)"""
    rng = np.random.default_rng(17)
    x = np.concatenate([rng.uniform(-300, 300, (40, 3)),
                        np.array([[0.0, 1.0, -0.5], [1e-7, 123456789.0, 2.0 / 3.0],
                                  [1e20, -1e-300, 5e-324]])])
    d = job_dir("csv")
    po.run_ref("csv", d, {"N": len(x), "M": 3}, {"x": x})
    with open(os.path.join(d, "ref.csv")) as f:
        text = f.read()
    np.savez_compressed(os.path.join(OUT, "csv.npz"), x=x, text=np.array(text))


def fixture_traj_rotor(N, nsteps=1000, snap=100, seed=77, periodic=1):
    s = wl.patchy_tetra(N)
    L = float(s["box"][0])
    n = max(int(L // 4), 1)
    d = job_dir(f"traj_rotor_{N}_{periodic}")
    job = {"N": N, "L": repr(L), "periodic": periodic, "n": n, "cut": 3.5, "rebuild": 20,
           "wca_eps": 1.0, "wca_sigma": 1.0, "dt": s["dt"], "gamma": s["gamma"],
           "gammar": repr(float(s["gammar"])), "kT": s["kT"], "nsteps": nsteps, "snap_every": snap,
           "noise_seed": seed}
    job.update(patch_job(s))
    po.run_ref("traj_rotor", d, job, {"x": s["x"], "p": s["p"], "Q": s["Q"], "L": s["L"]})
    out = {"x": s["x"], "Q": s["Q"], "box": L, "n": n, "nsteps": nsteps, "snap_every": snap,
           "noise_seed": seed,
           "x_out": po.read_ref(d, "x_out").reshape(N, 3),
           "p_out": po.read_ref(d, "p_out").reshape(N, 3),
           "Q_out": po.read_ref(d, "Q_out").reshape(N, 9),
           "L_out": po.read_ref(d, "L_out").reshape(N, 3),
           "snaps": po.read_ref(d, "snaps").reshape(-1, N, 3)}
    name = f"traj_rotor{N}.npz" if periodic else f"traj_rotor_walls{N}.npz"
    np.savez_compressed(os.path.join(OUT, name), **out)


def fixture_traj_walls():
    """1000 steps of the reference's Condensate::run loop with reflecting
    walls (cube with pb = false, the reference's default:
    Condensate.cpp:17-18; geometry.h:433-444)"""
    fixture_traj_rotor(256, seed=79, periodic=0)


def fixture_observables():
    """LangevinNVTR::measured_temperature (LangevinR.cpp:170-193),
    WCAPotential::energy / HSPotential::force / force2mdx
    (potential.h:79-200) at given distances, and MD::calculatestress
    (MD.cpp:2381-2424) with an HSPotential interaction"""
    out = {}
    rng = np.random.default_rng(23)
    for tag, m in (("m1", 1.0), ("m2", 2.0)):
        N = 777
        p = rng.normal(0, 1.3, (N, 3))
        L = rng.normal(0, 0.4, (N, 3))
        d = job_dir(f"temperature_{tag}")
        po.run_ref("temperature", d, {"N": N, "L": 10.0, "periodic": 1, "m": repr(m)},
                   {"p": p, "L": L})
        out[f"temp_{tag}_p"] = p
        out[f"temp_{tag}_L"] = L
        out[f"temp_{tag}_m"] = m
        out[f"temp_{tag}_out"] = po.read_ref(d, "temperature")
    for tag, eps, sig in (("e1s1", 1.0, 1.0), ("e25s13", 2.5, 1.3)):
        pl = 1.12246 * sig
        r = np.concatenate([np.linspace(0.75 * sig, 1.3 * sig, 200),
                            [pl, np.nextafter(pl, 0), np.nextafter(pl, 10), sig]])
        d = job_dir(f"pairpot_{tag}")
        po.run_ref("pairpot", d, {"n": len(r), "eps": repr(eps), "sigma": repr(sig)}, {"r": r})
        out[f"pot_{tag}_r"] = r
        out[f"pot_{tag}_eps"] = eps
        out[f"pot_{tag}_sigma"] = sig
        for k in ("energy", "force", "force2mdx"):
            out[f"pot_{tag}_{k}"] = po.read_ref(d, k)
    for tag, s in (("c1", wl.patchy_tetra()), ("c3", wl.condensate_ab(4096))):
        L = float(s["box"][0])
        n = int(L // 4)
        d = job_dir(f"stress_{tag}")
        po.run_ref("stress", d, {"N": s["N"], "L": repr(L), "periodic": 1, "n": n, "cut": 3.5,
                                 "wca_eps": 1.0, "wca_sigma": 1.0}, {"x": s["x"]})
        out[f"stress_{tag}_x"] = s["x"]
        out[f"stress_{tag}_L"] = L
        out[f"stress_{tag}_n"] = n
        out[f"stress_{tag}_out"] = po.read_ref(d, "stress").reshape(3, 3)
        out[f"stress_{tag}_npairs"] = len(po.read_ref(d, "pairs", np.int32)) // 2
    np.savez_compressed(os.path.join(OUT, "observables.npz"), **out)


def fixture_traj_bead(nsteps=1000, snap=100, seed=88):
    rng = np.random.default_rng(31)
    nch, nb, L = 4, 32, 12.0
    x = chain_positions(nch, nb, L, rng, jitter=0.05)
    N = len(x)
    idx = np.arange(N).reshape(nch, nb)
    bonds = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1).astype(np.int32)
    n = int(L // 4)
    d = job_dir("traj_bead")
    job = {"N": N, "L": L, "periodic": 1, "n": n, "cut": 3.5, "rebuild": 20, "wca_eps": 1.0,
           "wca_sigma": 1.0, "dt": 0.002, "gamma": 1.0, "kT": 1.0, "nsteps": nsteps,
           "snap_every": snap, "noise_seed": seed, "nbonds": len(bonds), "bond_k": 100.0,
           "bond_r0": 1.0}
    po.run_ref("traj_bead", d, job, {"x": x, "p": np.zeros((N, 3)), "bonds": bonds})
    out = {"x": x, "box": L, "n": n, "bonds": bonds, "nsteps": nsteps, "snap_every": snap,
           "noise_seed": seed, "x_out": po.read_ref(d, "x_out").reshape(N, 3),
           "p_out": po.read_ref(d, "p_out").reshape(N, 3),
           "snaps": po.read_ref(d, "snaps").reshape(-1, N, 3)}
    np.savez_compressed(os.path.join(OUT, "traj_bead128.npz"), **out)


def main():
    if not po.have_ref():
        sys.exit("oracle/_ref/mtref not built (make -C oracle ref)")
    fixture_cells()
    fixture_pairs()
    fixture_forces()
    fixture_pairsweep()
    fixture_bonded()
    fixture_traj_rotor(64)
    fixture_traj_rotor(512)
    fixture_traj_bead()
    fixture_calcg1()
    fixture_csv()
    fixture_dgraph()
    fixture_traj_walls()
    fixture_observables()
    print("fixtures written to", OUT)


if __name__ == "__main__":
    # python tests/golden/make_golden.py [fixture ...]: only the named ones
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            globals()[f"fixture_{name}"]()
    else:
        main()
