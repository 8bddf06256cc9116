"""The C++ host classes (mtcode_amd/host/mtcode_amd.hpp) used the way the
reference's drivers use MDBase: tests/cpp/dropin_driver runs the loop of
Condensate::run call by call (calculatepairs, calculateforces,
calculate_forces_and_torques3D, create_forces_and_torques_sphere,
advancemom_halfstep, advance_pos, rotate) or as device-resident steps
(run()), and the bead loop likewise; final states are compared with the
oracle driven by the same noise."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from mtcode_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")



def say(*a):
    print(*a, flush=True)


def run_driver(d, job, arrays):
    with open(os.path.join(d, "job.txt"), "w") as f:
        for k, v in job.items():
            f.write(f"{k} {v}\n")
    for k, v in arrays.items():
        v.tofile(os.path.join(d, k + ".bin"))
    if not os.path.exists(DRIVER):
        raise RuntimeError("tests/cpp/dropin_driver is not built (run __graft_entry__.build())")
    r = subprocess.run([DRIVER, d], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"dropin_driver failed: {r.stderr}")


def rd(d, name, shape):
    return np.fromfile(os.path.join(d, name), dtype=np.float64).reshape(shape)


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n", [1, 2, 3, 6, 16])
def test_stencil_matches_reference_fixture(tmp_path, golden, n):
    """cube::generate_boxes_relationships of the C++ classes (CPU only)"""
    g = golden("cells_patchy4096.npz")
    run_driver(str(tmp_path), {"integ": "stencil", "n": n}, {})
    out = np.fromfile(os.path.join(str(tmp_path), "out_stencil.bin"), dtype=np.int32)
    boxes = out.reshape(n ** 3, -1)
    if f"boxes_n{n}" in g:
        assert np.array_equal(boxes, g[f"boxes_n{n}"])
    else:
        h = (boxes.astype(np.int64) * np.arange(1, boxes.shape[1] + 1)).sum(1)
        assert np.array_equal(h, g[f"boxes_n{n}_row_hash"])
        assert boxes.sum() == g[f"boxes_n{n}_sum"]


def test_csv_output_matches_reference_bytes(tmp_path, golden):
    """outfunc writes the reference's CSV text byte for byte; importcsv reads
    it back as the reference's reader does (atof)"""
    g = golden("csv.npz")
    x = g["x"]
    run_driver(str(tmp_path), {"integ": "csv", "N": x.shape[0], "M": x.shape[1]}, {"x": x})
    with open(os.path.join(str(tmp_path), "out.csv")) as f:
        assert f.read() == str(g["text"])
    back = np.fromfile(os.path.join(str(tmp_path), "out_import.bin")).reshape(x.shape)
    ref = np.array([[float(v) for v in line.split(",")] for line in str(g["text"]).splitlines()])
    assert np.array_equal(back, ref)


@pytest.mark.gpu
def test_condensate_run_writes_reference_files(tmp_path):
    """Condensate::run(runtime, every): the output files' names (steps 0, 10,
    20 of 25; index zero-padded to the digits of ceil(25 / 10)) and shapes,
    positions inside the box and orthonormal orientations"""
    s = wl.patchy_tetra(512)
    N, nsteps, every = s["N"], 25, 10
    e, d, th = s["patches"]["params"][0]
    job = {"integ": "rotor", "mode": "output", "N": N, "L": repr(float(s["box"][0])),
           "nsteps": nsteps, "every": every, "dt": s["dt"], "gamma": s["gamma"],
           "gammar": repr(float(s["gammar"])), "kT": s["kT"], "precision": 0,
           "eps_p": repr(float(e)), "d": repr(float(d)), "thetam": repr(float(th))}
    nz = po.noise(1, 0, (nsteps + 1) * N * 6)
    run_driver(str(tmp_path), job, {"x": s["x"], "Q": s["Q"], "noise": nz})
    names = sorted(f for f in os.listdir(str(tmp_path)) if f.endswith(".csv"))
    assert names == ["orientation_test_i=0.csv", "orientation_test_i=1.csv",
                     "orientation_test_i=2.csv", "pos_test_i=0.csv", "pos_test_i=1.csv",
                     "pos_test_i=2.csv"]
    pos = np.loadtxt(os.path.join(str(tmp_path), "pos_test_i=2.csv"), delimiter=",")
    ori = np.loadtxt(os.path.join(str(tmp_path), "orientation_test_i=2.csv"), delimiter=",")
    assert pos.shape == (N, 3) and ori.shape == (N, 9)
    assert np.all((pos >= 0) & (pos <= s["box"][0]))
    assert np.allclose(np.einsum("nij,nkj->nik", ori.reshape(N, 3, 3), ori.reshape(N, 3, 3)),
                       np.eye(3), atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["stages", "run"])
def test_rotor_condensate_loop_vs_oracle(tmp_path, mode):
    s = wl.patchy_tetra(512)
    N, nsteps, seed = s["N"], 30, 9
    say("drop-in rotor", mode, N, nsteps)
    nz = po.noise(seed, 0, (nsteps + 1) * N * 6)
    e, d, th = s["patches"]["params"][0]
    job = {"integ": "rotor", "mode": mode, "N": N, "L": repr(float(s["box"][0])),
           "nsteps": nsteps, "dt": s["dt"], "gamma": s["gamma"], "gammar": repr(float(s["gammar"])),
           "kT": s["kT"], "precision": 0, "eps_p": repr(float(e)), "d": repr(float(d)),
           "thetam": repr(float(th))}
    run_driver(str(tmp_path), job, {"x": s["x"], "Q": s["Q"], "noise": nz})
    ref = po.run_rotors(s, nsteps, seed, 0, list_cut=3.5, rebuild_every=20)
    for k, shape in (("x", (N, 3)), ("p", (N, 3)), ("Q", (N, 9)), ("L", (N, 3))):
        err = relmax(rd(str(tmp_path), f"out_{k}.bin", shape), ref[k])
        say(k, err)
        assert err < 1e-10, k


@pytest.mark.gpu
@pytest.mark.parametrize("mode,model", [("bind", "single"), ("bindrun", "single"),
                                        ("bind", "binary"), ("bindrun", "ternary")])
def test_singlebond_loop_vs_oracle(tmp_path, mode, model):
    """Condensate::run_singlebond through the C++ classes (BinaryBindStore,
    BindingModelSingle / BindingModelBinary / BindingModelTernary,
    calculate_forces_and_torques3D_onlyone_nonlets), call by call and on the
    device, against the oracle's loop with the same injected noise and
    binding draws (0, 1, 2, ...: the call-by-call counter and a new engine's
    force-evaluation count): bound state equal, state within 1e-10.  The
    binary model's tables come from the classes' own setup_equilibrium, the
    oracle's from the reference's (tests/golden/bind.npz)"""
    import test_gpu_bind as tb
    s = wl.thermal_momenta(tb.lattice_system(8, 1.0, 1.1, 81), seed=3)
    N, nsteps, seed = s["N"], 25, 12
    say("drop-in singlebond", mode, N, nsteps)
    nz = po.noise(seed, 0, (nsteps + 1) * N * 6)
    e, d, th = s["patches"]["params"][0]
    job = {"integ": "rotor", "mode": mode, "N": N, "L": repr(float(s["box"][0])),
           "nsteps": nsteps, "dt": s["dt"], "gamma": s["gamma"], "gammar": repr(float(s["gammar"])),
           "kT": s["kT"], "precision": 0, "eps_p": repr(float(e)), "d": repr(float(d)),
           "thetam": repr(float(th)), "on_rate": 0.7, "off_rate": 0.3}
    s0 = dict(s, p=np.zeros((N, 3)), L=np.zeros((N, 3)))  # the driver starts from rest
    nh = 4 * N
    arrays = {"x": s["x"], "Q": s["Q"], "noise": nz}
    tables = None
    if model == "binary":
        job.update({"model": "binary", "div": nh // 2})
        tables = {q: tb.Z[f"bin_{q}"] for q in ("dtab", "ttab", "tr4")}
        tables["species"] = np.where(np.arange(nh) < nh // 2, 1, 2).astype(np.int32)
    elif model == "ternary":
        job.update({"model": "ternary", "div1": 400, "div2": 1200, "np": 4})
        tables = {q: tb.Z[f"ter_{q}"] for q in ("dtab", "ttab", "tr4")}
        i = np.arange(nh)
        tables["species"] = np.where(i < 400, 1, np.where(i < 1200, 2,
                                     np.where((i - 1200) % 4 == 3, 3, 1))).astype(np.int32)
        arrays.update({"dtab": tables["dtab"], "ttab": tables["ttab"]})
    run_driver(str(tmp_path), job, arrays)
    key = (1, 0xB1D0B1D0)   # the drop-in classes' default seed 1
    tb_key = tb.KEY
    tb.KEY = key
    try:
        ref, ibo, bto = tb.oracle_bind_run(s0, nsteps, nz.reshape(nsteps + 1, N, 6), 0.7, 0.3,
                                           list(range(nsteps + 1)), tables=tables)
    finally:
        tb.KEY = tb_key
    nh = len(ibo)
    assert ibo.sum() > 0
    assert np.array_equal(np.fromfile(os.path.join(str(tmp_path), "out_isbound.bin"), np.int32), ibo)
    assert np.array_equal(np.fromfile(os.path.join(str(tmp_path), "out_boundto.bin"), np.int32), bto)
    for k, shape in (("x", (N, 3)), ("p", (N, 3)), ("Q", (N, 9)), ("L", (N, 3))):
        err = relmax(rd(str(tmp_path), f"out_{k}.bin", shape), ref[k])
        say(k, err)
        assert err < 1e-10, k


@pytest.mark.gpu
def test_singlebond_run_writes_reference_files(tmp_path):
    """Condensate::run_singlebond(runtime, every, strbase) through the C++
    classes: the reference's file set for the steps i % every == 0 with
    i > 0.9 runtime (pos, orientation at precision 10, bindings: isbound then
    boundto, connectedcomponents), bonds mutual, and the components equal to
    the oracle's distance_graph labels of the written positions"""
    import test_gpu_bind as tb
    s = tb.lattice_system(8, 1.0, 1.1, 91)
    N, runtime, every = s["N"], 60, 5
    e, d, th = s["patches"]["params"][0]
    job = {"integ": "rotor", "mode": "bindout", "N": N, "L": repr(float(s["box"][0])),
           "nsteps": runtime, "every": every, "dt": s["dt"], "gamma": s["gamma"],
           "gammar": repr(float(s["gammar"])), "kT": s["kT"], "precision": 0,
           "eps_p": repr(float(e)), "d": repr(float(d)), "thetam": repr(float(th)),
           "on_rate": 0.7, "off_rate": 0.3}
    say("drop-in singlebond output")
    run_driver(str(tmp_path), job, {"x": s["x"], "Q": s["Q"],
                                    "noise": np.zeros((runtime + 1) * N * 6)})
    steps = [i for i in range(runtime) if i % every == 0 and i > int(0.9 * runtime)]
    names = sorted(os.listdir(str(tmp_path)))
    L = float(s["box"][0])
    for i in steps:
        k = f"_test_i={i // every:02d}.csv"
        for base in ("pos", "orientation", "bindings", "connectedcomponents"):
            assert base + k in names, base + k
        x = np.loadtxt(os.path.join(str(tmp_path), "pos" + k), delimiter=",")
        q = open(os.path.join(str(tmp_path), "orientation" + k)).read().splitlines()
        assert max(len(v) for ln in q for v in ln.split(",")) > 9   # precision 10
        ib_s, bt_s = open(os.path.join(str(tmp_path), "bindings" + k)).read().split("\n")
        ib = np.array([int(v) for v in ib_s.split(",")])
        bt = np.array([int(v) for v in bt_s.split(",")])
        assert len(ib) == 4 * N and set(np.unique(ib)) <= {0, 1} and ib.sum() > 0
        b = np.nonzero(ib)[0]
        assert np.all(ib[bt[b]] == 1) and np.all(bt[bt[b]] == b)
        cc = np.array([int(v) for v in
                       open(os.path.join(str(tmp_path), "connectedcomponents" + k)).read().split(",")])
        # labels: each the smallest index of its component; pairs within the
        # distance share one.  The CSV positions carry 6 digits: the oracle's
        # labels of them must agree unless a pair sits within 1e-4 of 1.25
        assert np.all(cc <= np.arange(N)) and np.all(cc[cc] == cc)
        dx = x[:, None, :] - x[None, :, :]
        dx -= L * np.round(dx / L)
        r = np.sqrt((dx ** 2).sum(-1))
        close = np.argwhere((r < 1.25 - 1e-4) & ~np.eye(N, dtype=bool))
        assert np.all(cc[close[:, 0]] == cc[close[:, 1]])
        if not np.any(np.abs(r - 1.25) < 1e-4):
            n = max(1, int(np.floor(L / 3.0)))
            assert np.array_equal(cc, po.dgraph_labels(L, x, n, 1.25))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["stages", "run"])
def test_bead_loop_vs_oracle(tmp_path, mode):
    s = wl.mt_bundle(2, 13, 40, spacing=6.0, box=(40.0, 40.0, 40.0))
    s = dict(s)
    s["exclusions"] = None  # the reference's bead loops have no exclusion list
    N, nsteps, seed = s["N"], 30, 4
    say("drop-in bead", mode, N, nsteps)
    nz = po.noise(seed, 0, nsteps * N * 3)
    assert np.all(s["bond_k"] == s["bond_k"][0]) and np.all(s["bond_r0"] == s["bond_r0"][0])
    job = {"integ": "bead", "mode": mode, "N": N, "L": repr(float(s["box"][0])),
           "nsteps": nsteps, "dt": s["dt"], "gamma": s["gamma"], "kT": s["kT"], "precision": 0,
           "nbonds": len(s["bonds"]), "nangles": len(s["angles"]),
           "kbond": repr(float(s["bond_k"][0])), "r0": repr(float(s["bond_r0"][0])),
           "kappa": repr(float(s["angle_k"][0]))}
    assert s["angle_form"] == 1
    run_driver(str(tmp_path), job, {"x": s["x"], "noise": nz,
                                    "bonds": np.asarray(s["bonds"], dtype=np.int32),
                                    "angles": np.asarray(s["angles"], dtype=np.int32)})
    ref = po.run_beads(s, nsteps, seed, 0, list_cut=3.5, rebuild_every=20)
    for k in ("x", "p"):
        err = relmax(rd(str(tmp_path), f"out_{k}.bin", (N, 3)), ref[k])
        say(k, err)
        assert err < 1e-10, k


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["u4096", "u4096ex", "n2", "chains"])
def test_distance_graph_matches_reference(tmp_path, golden, case):
    """distance_graph(pos, boxes, geo, d, i1) of the C++ classes (index_test or
    excludea) against the reference's labels"""
    g = golden("dgraph.npz")
    x = g[f"{case}_x"]
    lo, hi = (int(v) for v in g[f"{case}_ex"])
    say("distance_graph", case, len(x))
    run_driver(str(tmp_path), {"integ": "dgraph", "N": len(x), "L": repr(float(g[f"{case}_L"])),
                               "n": int(g[f"{case}_n"]), "cut": repr(float(g[f"{case}_cut"])),
                               "ex_lo": lo, "ex_hi": hi}, {"x": x})
    lab = np.fromfile(os.path.join(str(tmp_path), "out_labels.bin"), dtype=np.int32)
    assert np.array_equal(lab, g[f"{case}_labels"])


@pytest.mark.gpu
def test_patch_list_through_dropin(tmp_path):
    """LangevinNVTR::calculate_patch_list, adjacency and
    calculate_bound_pairs through the C++ classes (TetrahedralPatch) on the
    reference's own pairs of the wide-cone fixture: rows, their order and
    the run starts equal to the compiled reference's
    (tests/golden/plist.npz)"""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_bind_golden as g
    PL = np.load(os.path.join(ROOT, "tests", "golden", "plist.npz"))
    _, s = g.systems()
    N = s["N"]
    pairs = np.ascontiguousarray(PL["wide_pairs"], dtype=np.int32)
    e, d, th = s["patches"]["params"][0]
    say("drop-in patch list", N, len(pairs))
    job = {"integ": "rotor", "mode": "plist", "N": N, "L": repr(float(s["box"][0])),
           "nsteps": 0, "dt": s["dt"], "gamma": s["gamma"], "gammar": repr(float(s["gammar"])),
           "kT": s["kT"], "precision": 0, "eps_p": repr(float(e)), "d": repr(float(d)),
           "thetam": repr(float(th)), "npairs": len(pairs)}
    run_driver(str(tmp_path), job, {"x": s["x"], "Q": s["Q"], "noise": np.zeros(N * 6),
                                    "pairs": pairs})
    rows = np.fromfile(os.path.join(tmp_path, "out_plist.bin"), dtype=np.int32).reshape(-1, 5)
    runs = np.fromfile(os.path.join(tmp_path, "out_adjacency.bin"), dtype=np.int32)
    assert np.array_equal(rows, PL["wide_plist"])
    assert np.array_equal(runs, PL["wide_adjacency"])
    brows = np.fromfile(os.path.join(tmp_path, "out_blist.bin"), dtype=np.int32).reshape(-1, 5)
    assert np.array_equal(brows, PL["wide_blist"])
