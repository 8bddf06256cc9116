"""Parity of the HIP engine (through the C ABI) with the reference's golden
vectors and with the oracle, on an MI355X.

Bars: bit-exact for cell ids and neighbour sets; forces and torques within
1e-12 (FP64) / 1e-5 (mixed) of the FP64 value, relative to the largest
force (torque) of the configuration; 1000-step FP64 trajectories with
injected noise within the chaos-amplified rounding floor measured between
the oracle and the reference themselves."""
import sys

import numpy as np
import pytest

import pyoracle as po
from mtcode_amd import mdx
from mtcode_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def engine_for(s, precision=mdx.PREC_FP64, **kw):
    return mdx.engine_from_system(s, precision=precision, **kw)


# ------------------------------------------------------------ binning, pairs
@pytest.mark.parametrize("n", [1, 2, 3, 6, 16])
def test_cell_assign_bit_exact(golden, n):
    g = golden("cells_patchy4096.npz")
    say("cell_assign n", n)
    x = g["x"]
    with mdx.Engine(len(x), g["L"], 1.0, 0.005) as e:
        e.upload(x)
        assert np.array_equal(e.cell_assign(n), g[f"cells_n{n}"])


def test_neighbor_pairs_bit_exact(golden):
    g = golden("pairs.npz")
    for k in [k for k in g if k.startswith("pairs512")]:
        parts = k.split("_")
        n, cut = int(parts[1][1:]), float(parts[2][3:])
        periodic = not k.endswith("nonperiodic")
        say("neighbor_pairs", k)
        with mdx.Engine(len(g["x512"]), g["L512"], 1.0, 0.005, periodic=periodic) as e:
            e.upload(g["x512"])
            got = e.neighbor_pairs(n, cut)
        ref = np.sort(np.sort(g[k], axis=1).astype(np.int64) @ np.array([1, 1 << 20]))
        key = np.sort(got.astype(np.int64) @ np.array([1, 1 << 20]))
        assert np.all(got[:, 0] < got[:, 1])
        assert np.array_equal(key, ref), k


def test_neighbor_pairs_full_size(golden):
    g = golden("pairs.npz")
    s = wl.patchy_tetra()
    say("neighbor_pairs 4096")
    with mdx.Engine(s["N"], s["box"], 1.0, 0.005) as e:
        e.upload(s["x"])
        got = e.neighbor_pairs(16, 1.68)
    key = np.sort(got.astype(np.int64) @ np.array([1, 1 << 20]))
    assert np.array_equal(key, g["pairs4096_n16_cut1.68_keys"])


# ------------------------------------------------------------------- forces
@pytest.mark.parametrize("name,builder", [("forces_patchy4096.npz", lambda: wl.patchy_tetra()),
                                          ("forces_condensate4096.npz",
                                           lambda: wl.condensate_ab(4096))])
def test_forces_fp64_vs_reference(golden, name, builder):
    g = golden(name)
    s = builder()
    say("forces fp64", name)
    with engine_for(s) as e:
        F, T = e.compute_forces()
        obs = e.observables()
    assert relmax(F, g["F"]) < 1e-12
    assert relmax(T, g["T"]) < 1e-12
    # patch + WCA energy: the reference fixture holds the patch part only
    wca_e = 0.0
    pr = po.pairs_reference(s["box"][0], s["x"], 16, 1.12246)
    for a, b in pr:
        d = s["x"][a] - s["x"][b]
        d -= s["box"][0] * np.round(d / s["box"][0])
        wca_e += po.lib().orc_wca_energy(float(np.linalg.norm(d)), 1.0, 1.0, 0)
    assert abs(obs["epot"] - (g["epot"] + wca_e)) < 1e-9 * max(1.0, abs(g["epot"]))


@pytest.mark.parametrize("name,dirs", [("tetra", wl.TETRA_DIRS), ("bivalent", wl.BIVALENT_DIRS)])
def test_two_particle_sweeps_fp64(golden, name, dirs):
    g = golden("pairsweep.npz")
    x = g[name + "_x"]
    N = len(x)
    L = float(g[name + "_L"])
    say("sweep fp64", name)
    with mdx.Engine(N, L, 1.0, 0.005) as e:
        e.upload(x, None, g[name + "_Q"], None)
        e.set_wca(1.0, 1.0, 0.0)
        e.set_patches(np.zeros(N), [len(dirs)], dirs,
                      np.tile([6.0, 1.2, wl.THETAM_095], (len(dirs) ** 2, 1)), 2.4)
        F, T = e.compute_forces()
    assert relmax(F, g[name + "_F"]) < 1e-12
    assert relmax(T, g[name + "_T"]) < 1e-12


def test_bonded_fp64(golden):
    g = golden("bonded.npz")
    x, L, b, a = g["x"], float(g["L"]), g["bonds"], g["angles"]
    N = len(x)
    cases = [("Fb_harmonic", dict(k=100.0, r0=np.ones(len(b)), kind=1)),
             ("Fb_fene", dict(k=30.0, r0=np.full(len(b), 1.5), kind=2)),
             ("Fb_perbond", dict(k=100.0, r0=g["r0s"], kind=1))]
    for key, c in cases:
        say("bonded", key)
        with mdx.Engine(N, L, 1.0, 0.002, integrator=mdx.INTEG_BEAD) as e:
            e.upload(x)
            e.set_bonds(b, np.full(len(b), c["k"]), c["r0"], c["kind"])
            F, _ = e.compute_forces()
        assert relmax(F, g[key]) < 1e-13, key
    for th in (0.3, 0.0):
        say("angles", th)
        with mdx.Engine(N, L, 1.0, 0.002, integrator=mdx.INTEG_BEAD) as e:
            e.upload(x)
            e.set_angles(a, np.full(len(a), 50.0), np.full(len(a), th), 0)
            F, _ = e.compute_forces()
        assert relmax(F, g[f"Fa_theta0_{th}"]) < 1e-12


def test_cosine_bending_and_active_vs_oracle():
    rng = np.random.default_rng(4)
    s = wl.active_polymer(nchains=6, nbeads=40, phi=0.05, seed=3)
    say("active polymer small fp64", s["N"])
    with engine_for(s) as e:
        F, T = e.compute_forces()
    Fo, To = po.compute_forces(s, list_cut=3.5)
    assert relmax(F, Fo) < 1e-12
    assert relmax(T, To) < 1e-12 or np.abs(To).max() == 0
    _ = rng


# --------------------------------------------------------------- trajectories
@pytest.mark.parametrize("N", [64, 512])
def test_rotor_trajectory_fp64_vs_reference(golden, N):
    g = golden(f"traj_rotor{N}.npz")
    s = wl.patchy_tetra(N)
    nsteps, snap = int(g["nsteps"]), int(g["snap_every"])
    noise = po.noise(int(g["noise_seed"]), 0, (nsteps + 1) * N * 6).reshape(nsteps + 1, N, 6)
    say("rotor trajectory fp64", N)
    with engine_for(s) as e:
        e.prime(noise[0])
        errs = []
        for k in range(nsteps // snap):
            e.step(snap, noise[1 + k * snap: 1 + (k + 1) * snap])
            st = e.download()
            errs.append(float(np.max(np.abs(st["x"] - g["snaps"][k]))))
    say("max |dx| per 100 steps", errs)
    assert errs[0] < 1e-10
    assert np.max(np.abs(st["x"] - g["x_out"])) < 1e-6
    assert np.max(np.abs(st["Q"] - g["Q_out"])) < 1e-6
    assert relmax(st["p"], g["p_out"]) < 1e-5
    assert relmax(st["L"], g["L_out"]) < 1e-5


def test_bead_trajectory_fp64_vs_reference(golden):
    g = golden("traj_bead128.npz")
    N = len(g["x"])
    nsteps = int(g["nsteps"])
    noise = po.noise(int(g["noise_seed"]), 0, nsteps * N * 3).reshape(nsteps, N, 3)
    say("bead trajectory fp64")
    with mdx.Engine(N, float(g["box"]), 1.0, 0.002, gamma=1.0, integrator=mdx.INTEG_BEAD) as e:
        e.upload(g["x"])
        e.set_wca(1.0, 1.0, 0.0)
        e.set_bonds(g["bonds"], np.full(len(g["bonds"]), 100.0), np.ones(len(g["bonds"])), 1)
        e.step(nsteps, noise)
        st = e.download()
    assert np.max(np.abs(st["x"] - g["x_out"])) < 1e-9
    assert relmax(st["p"], g["p_out"]) < 1e-9


def _first(s, k):
    """the first k particles of a rotor system (same box and parameters)"""
    t = dict(s)
    t["N"] = k
    for key in ("x", "p", "Q", "L"):
        t[key] = np.array(s[key][:k])
    t["patches"] = dict(s["patches"])
    t["patches"]["types"] = np.array(s["patches"]["types"][:k])
    return t


@pytest.mark.parametrize("k", [1, 2])
def test_one_and_two_particle_systems_vs_oracle(k):
    """Edge sizes: a lone particle (no pairs; friction, noise and rotation
    only) and an interacting pair, 30 FP64 steps with injected noise; a call
    of zero steps changes nothing"""
    s = _first(wl.patchy_tetra(64, phi=0.05), k)
    if k == 2:
        # in WCA range, with particle 1 on particle 0's first patch axis
        n0 = s["Q"][0].reshape(3, 3).T @ wl.TETRA_DIRS[0]
        s["x"][1] = (s["x"][0] + 1.05 * n0) % s["box"]
    nsteps, seed = 30, 11
    say("edge size", k)
    noise = po.noise(seed, 0, (nsteps + 1) * k * 6).reshape(nsteps + 1, k, 6)
    with engine_for(s) as e:
        e.prime(noise[0])
        before = e.download()
        e.step(0)
        same = e.download()
        e.step(nsteps, noise[1:])
        st = e.download()
    for key in ("x", "p", "Q", "L"):
        assert np.array_equal(before[key], same[key]), key
    ref = po.run_rotors(s, nsteps, seed, 0, list_cut=3.5, rebuild_every=20)
    for key in ("x", "p", "Q", "L"):
        err = relmax(st[key], ref[key])
        say(key, err)
        assert err < 1e-10, key


def test_ragged_batch_with_tiny_replicas_matches_independent_runs():
    """A batch whose replicas hold 1, 2, 700 and 3 particles (ragged, with
    replicas smaller than a cell stencil) equals the replicas run alone"""
    base = [wl.patchy_tetra(64, phi=0.05, seed_pos=40 + r) for r in range(4)]
    reps = [_first(base[0], 1), _first(base[1], 2), wl.patchy_tetra(700), _first(base[3], 3)]
    nsteps = 15
    Ntot = sum(r["N"] for r in reps)
    noise = po.noise(6, 0, (nsteps + 1) * Ntot * 6).reshape(nsteps + 1, Ntot, 6)
    offs = np.concatenate([[0], np.cumsum([r["N"] for r in reps])])
    say("ragged batch", [r["N"] for r in reps])
    with mdx.engine_from_system(None, replicas=reps) as e:
        e.prime(noise[0])
        e.step(nsteps, noise[1:])
        batch = e.download()
    for r, s in enumerate(reps):
        sl = slice(offs[r], offs[r + 1])
        with engine_for(s) as e:
            e.prime(noise[0, sl])
            e.step(nsteps, noise[1:, sl])
            one = e.download()
        for key in ("x", "p", "Q", "L"):
            assert relmax(batch[key][sl], one[key]) < 1e-12, (r, key)


# ------------------------------------------------------------ mixed precision
def f32_rounded(s):
    s = dict(s)
    s["Q"] = wl.mixed_orientations(s["Q"])
    return s


@pytest.mark.parametrize("builder", [lambda: wl.patchy_tetra(), lambda: wl.condensate_ab(32768)])
def test_mixed_forces_within_1e5_of_fp64(builder):
    """the FP64 engine gets exactly the orientations the mixed engine's
    kernels use (wl.mixed_orientations)"""
    s = builder()
    say("mixed vs fp64 forces", s["name"], s["N"])
    with engine_for(f32_rounded(s)) as e64:
        F64, T64 = e64.compute_forces()
    with engine_for(s, mdx.PREC_MIXED) as e32:
        F32, T32 = e32.compute_forces()
    ef, et = relmax(F32, F64), relmax(T32, T64)
    rms = float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64))
    say("max rel F", ef, "T", et, "rms F", rms)
    # the bound is the max-norm 1e-5; the RMS, the FP32 rounding of a few
    # terms per particle (about 1e-6), is held to twice that
    assert ef < 1e-5 and et < 1e-5 and rms < 2e-6


def test_determinism_bitwise():
    s = wl.patchy_tetra(2048)
    say("determinism")
    outs = []
    for _ in range(2):
        with engine_for(s, mdx.PREC_MIXED, seed=9) as e:
            e.prime()
            e.step(50)
            outs.append(e.download())
    for k in ("x", "p", "Q", "L"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


# ------------------------------------------------------------------ replicas
def test_replicas_match_independent_runs():
    reps = wl.replicas(6, seed=5, nmin=200, nmax=900)
    say("replicas", [r["N"] for r in reps])
    nsteps = 20
    Ntot = sum(r["N"] for r in reps)
    noise = po.noise(3, 0, (nsteps + 1) * Ntot * 6).reshape(nsteps + 1, Ntot, 6)
    offs = np.concatenate([[0], np.cumsum([r["N"] for r in reps])])
    with mdx.engine_from_system(None, replicas=reps) as e:
        e.prime(noise[0])
        e.step(nsteps, noise[1:])
        batch = e.download()
    for r, s in enumerate(reps):
        sl = slice(offs[r], offs[r + 1])
        with engine_for(s) as e:
            e.prime(noise[0, sl])
            e.step(nsteps, np.ascontiguousarray(noise[1:, sl]))
            one = e.download()
        for k in ("x", "p", "Q", "L"):
            assert np.array_equal(batch[k][sl], one[k]), (r, k)


# --------------------------------------------------------------- statistics
def test_equipartition_with_philox_noise():
    s = wl.patchy_tetra(4096)
    s["patches"]["params"][:, 0] = 0.0  # ideal rotors + WCA only
    say("equipartition")
    with engine_for(s, mdx.PREC_MIXED, seed=1) as e:
        e.prime()
        e.step(2000)
        acc = np.zeros(6)
        for _ in range(20):
            e.step(50)
            o = e.observables()
            acc += np.concatenate([o["ke_t"], o["ke_r"]])
    acc /= 20
    say("<KE> per component", acc)
    # the explicit friction scheme of the reference at dt=0.005 gives kT/2 to a few %
    assert np.all(np.abs(acc - 0.5) < 0.05)


# ----------------------------------------------------------- LDS tiles
@pytest.mark.parametrize("precision", [mdx.PREC_MIXED, mdx.PREC_FP64])
def test_lds_tiles_bitwise_equal_global_walk(monkeypatch, precision):
    """The LDS-staged tile walk and the global-memory walk visit neighbours in
    the same order, so the trajectories are bitwise identical."""
    s = wl.condensate_ab(1 << 18)
    say("tiles vs global", s["N"], precision)
    outs = []
    monkeypatch.setenv("MDX_SPLIT", "0")
    monkeypatch.setenv("MDX_SKIN", "0")
    for tiles in ("1", "0"):
        monkeypatch.setenv("MDX_TILES", tiles)
        with engine_for(s, precision, seed=4) as e:
            e.prime()
            e.step(20)
            outs.append(e.download())
            F, T = e.compute_forces()
            outs[-1]["F"], outs[-1]["T"] = F, T
    for k in ("x", "p", "Q", "L", "F", "T"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


# ------------------------------------------------- split walk / finish
@pytest.mark.parametrize("precision", [mdx.PREC_MIXED, mdx.PREC_FP64])
def test_split_walk_finish_bitwise_equal_fused(monkeypatch, precision):
    """The two-kernel form (neighbour walk, then patches + integrator) adds the
    same terms in the same order as the one-kernel form: bitwise identical
    trajectories, forces and torques."""
    s = wl.condensate_ab(1 << 18)
    say("split vs fused", s["N"], precision)
    outs = []
    monkeypatch.setenv("MDX_TILES", "0")
    monkeypatch.setenv("MDX_SKIN", "0")
    for split in ("1", "0"):
        monkeypatch.setenv("MDX_SPLIT", split)
        with engine_for(s, precision, seed=5) as e:
            e.prime()
            e.step(20)
            outs.append(e.download())
            F, T = e.compute_forces()
            outs[-1]["F"], outs[-1]["T"] = F, T
    for k in ("x", "p", "Q", "L", "F", "T"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_split_candidate_overflow_matches_oracle(monkeypatch):
    """A simple-cubic cluster at spacing 0.95 gives interior particles 18 patch
    candidates (6 at 0.95, 12 at 1.34, all inside max_distance_check), more
    than the 16 the walk hands over, so the finish kernel walks the
    neighbour cells again for the rest; forces and torques still match the
    oracle."""
    s = wl.patchy_tetra(512, phi=0.1)
    g = np.arange(8) * 0.95
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    s["x"] = lat + (s["box"] - 7 * 0.95) / 2
    say("candidate overflow cluster", s["N"])
    monkeypatch.setenv("MDX_SPLIT", "1")
    with engine_for(s) as e:
        F, T = e.compute_forces()
    Fo, To = po.compute_forces(s, list_cut=3.5)
    assert relmax(F, Fo) < 1e-12
    assert relmax(T, To) < 1e-12


@pytest.mark.parametrize("precision", [mdx.PREC_FP64, mdx.PREC_MIXED])
def test_overflow_steps_bitwise_equal_one_kernel(monkeypatch, precision):
    """The same simple-cubic cluster stepped 20 times: in the split form the
    walk raises the overflow gate, k_patch_ovf adds the candidates beyond 16
    and the integrator closes the gate each step; the one-kernel form handles
    them inline.  Same trajectory bit for bit."""
    s = wl.patchy_tetra(512, phi=0.1)
    g = np.arange(8) * 0.95
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    s["x"] = lat + (s["box"] - 7 * 0.95) / 2
    say("overflow steps", precision)
    outs = []
    for split in ("1", "0"):
        monkeypatch.setenv("MDX_SPLIT", split)
        with engine_for(s, precision, seed=12) as e:
            e.prime()
            for _ in range(4):
                e.step(5)
            outs.append(e.download())
    for k in ("x", "p", "Q", "L"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


# ------------------------------------------------------------- neighbour list
def _run(s, precision, nsteps, skin, kmax=64, resort=200, seed=7):
    with engine_for(s, precision, seed=seed) as e:
        e.set_neighbor_list(skin, kmax, resort)
        if s["integrator"] == "rotor":
            e.prime()
        e.step(nsteps)
        out = e.download()
        out["builds"] = e.info()["list_builds"]
    return out


@pytest.mark.parametrize("split", ["1", "0"])
@pytest.mark.parametrize("kmax", [64, 4])
def test_neighbor_list_rotor_matches_cell_walk(monkeypatch, kmax, split):
    """Steps that walk the neighbour list (rebuilt on the device from the
    displacement bound, re-sorted every 25 steps; kmax=4 sends most particles
    to the cell fallback) evaluate the same pairs as the exhaustive cell walk;
    only the summation order differs, so FP64 trajectories agree to rounding.
    Both the walk / patch / integrator kernels and the one-kernel form."""
    monkeypatch.setenv("MDX_SPLIT", split)
    s = wl.condensate_ab(1 << 15, phi=0.2)
    say("list vs cell walk (rotor)", s["N"], "kmax", kmax)
    ref = _run(s, mdx.PREC_FP64, 80, 0.0)
    lst = _run(s, mdx.PREC_FP64, 80, 0.3, kmax=kmax, resort=25)
    say("list builds", lst["builds"])
    assert lst["builds"] >= 4
    for k in ("x", "p", "Q", "L"):
        err = relmax(lst[k], ref[k])
        say(k, err)
        assert err < 1e-9, k


@pytest.mark.parametrize("split", ["1", "0"])
@pytest.mark.parametrize("builder", [
    lambda: wl.mt_bundle(2, 13, 60, spacing=6.0, box=(12.0, 12.0, 60.0)),
    lambda: wl.active_polymer(nchains=64, nbeads=64, phi=0.15, seed=3),
])
def test_neighbor_list_topology_matches_cell_walk(monkeypatch, builder, split):
    """Bead integrator (microtubule bundle: bonds, bending, exclusions) and a
    rotor system with topology (active polymer with patchy end beads)."""
    monkeypatch.setenv("MDX_SPLIT", split)
    s = builder()
    say("list vs cell walk", s["name"], s["integrator"], s["N"])
    ref = _run(s, mdx.PREC_FP64, 80, 0.0)
    lst = _run(s, mdx.PREC_FP64, 80, 0.3, kmax=8, resort=30)
    say("list builds", lst["builds"])
    assert lst["builds"] >= 3
    for k in ("x", "p"):
        err = relmax(lst[k], ref[k])
        say(k, err)
        assert err < 1e-9, k


def test_neighbor_list_mixed_close_to_cell_walk():
    s = wl.condensate_ab(1 << 15, phi=0.2)
    ref = _run(s, mdx.PREC_MIXED, 40, 0.0)
    lst = _run(s, mdx.PREC_MIXED, 40, 0.4)
    for k in ("x", "p", "Q", "L"):
        err = relmax(lst[k], ref[k])
        say("mixed", k, err)
        assert err < 1e-3, k


# ------------------------------------------------------------------ graphs
@pytest.mark.parametrize("split", ["1", "0"])
@pytest.mark.parametrize("builder", [
    lambda: wl.condensate_ab(1 << 16, phi=0.1),
    lambda: wl.mt_bundle(2, 13, 60, spacing=6.0, box=(12.0, 12.0, 60.0)),
])
def test_graph_replay_bitwise_equal_eager(monkeypatch, builder, split):
    """List steps replayed from captured graphs (two graphs, alternating
    state buffers, interleaved with eager re-sort steps, list rebuilds and
    host synchronisations between calls) give the same trajectory, bit for
    bit, as eager launches."""
    monkeypatch.setenv("MDX_SPLIT", split)
    s = builder()
    say("graph vs eager", s["name"], s["N"])
    outs = []
    for g in ("1", "0"):
        monkeypatch.setenv("MDX_GRAPH", g)
        with engine_for(s, mdx.PREC_MIXED, seed=3) as e:
            e.set_neighbor_list(0.3, 64, 40)
            if s["integrator"] == "rotor":
                e.prime()
            for _ in range(6):
                e.step(50)
                e.observables()
            out = e.download()
            out["builds"] = e.info()["list_builds"]
            outs.append(out)
    say("builds", outs[0]["builds"], outs[1]["builds"])
    assert outs[0]["builds"] == outs[1]["builds"] and outs[0]["builds"] > 6
    keys = ("x", "p", "Q", "L") if s["integrator"] == "rotor" else ("x", "p")
    for k in keys:
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("precision", [mdx.PREC_MIXED, mdx.PREC_FP64])
@pytest.mark.parametrize("builder", [
    lambda: wl.condensate_ab(1 << 16, phi=0.1),
    lambda: wl.mt_bundle(2, 13, 60, spacing=6.0, box=(12.0, 12.0, 60.0)),
])
def test_fused_integrator_bitwise_equal_separate(monkeypatch, builder, precision):
    """List steps that integrate inside the walk (particles without patch
    candidates) and the patch kernel (the others) give the trajectory of the
    separate integrator kernel bit for bit, with the same list rebuilds (the
    drift bound is reduced over both kernels' blocks)."""
    monkeypatch.setenv("MDX_SPLIT", "1")
    s = builder()
    say("fused vs separate", s["name"], s["N"], precision)
    outs = []
    for f in ("1", "0"):
        monkeypatch.setenv("MDX_FUSE", f)
        with engine_for(s, precision, seed=5) as e:
            e.set_neighbor_list(0.3, 64, 40)
            if s["integrator"] == "rotor":
                e.prime()
            for _ in range(4):
                e.step(60)
            out = e.download()
            out["builds"] = e.info()["list_builds"]
            outs.append(out)
    say("builds", outs[0]["builds"], outs[1]["builds"])
    assert outs[0]["builds"] == outs[1]["builds"] and outs[0]["builds"] > 4
    keys = ("x", "p", "Q", "L") if s["integrator"] == "rotor" else ("x", "p")
    for k in keys:
        assert np.array_equal(outs[0][k], outs[1][k]), k


# ---------------------------------------------------------------- analysis
@pytest.mark.parametrize("case", ["c4096", "c512", "c512n1"])
def test_pair_histogram_bit_exact(golden, case):
    """mdx_pair_histogram against the reference's calcg1 counts"""
    g = golden("calcg1.npz")
    x = g[f"{case}_x"]
    s = wl.patchy_tetra(len(x))
    s["x"] = x
    assert float(s["box"][0]) == float(g[f"{case}_L"])
    with engine_for(s) as e:
        c = e.pair_histogram(float(g[f"{case}_max"]), float(g[f"{case}_dx"]))
    assert np.array_equal(c, g[f"{case}_counts"])


@pytest.mark.parametrize("case", ["u4096", "u4096ex", "u4096c16", "n2", "chains"])
def test_clusters_bit_exact(golden, case):
    """mdx_clusters against the labels of the reference's distance_graph
    (calculatepairs_parallel + ConnectedComponentsParallel_Old)"""
    g = golden("dgraph.npz")
    x = g[f"{case}_x"]
    lo, hi = (int(v) for v in g[f"{case}_ex"])
    say("clusters", case, len(x))
    with mdx.Engine(len(x), float(g[f"{case}_L"]), 1.0, 0.005) as e:
        e.upload(x)
        lab, nc = e.clusters(int(g[f"{case}_n"]), float(g[f"{case}_cut"]), (lo, hi))
    assert np.array_equal(lab, g[f"{case}_labels"])
    assert nc == len(np.unique(g[f"{case}_labels"]))


def test_clusters_large_and_replica_vs_oracle():
    """262,144 condensate particles after 200 steps (binding distances 1.25
    and 1.6, the latter percolating), and one replica of a batch, against the
    oracle's union-find labels"""
    s = wl.condensate_ab(1 << 18, phi=0.1)
    L = float(s["box"][0])
    n = int(L // 3.5)
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=3) as e:
        e.prime()
        e.step(200)
        x = e.download()["x"]
        for cut in (1.25, 1.6):
            say("clusters large", s["N"], cut)
            lab, nc = e.clusters(n, cut)
            ref = po.dgraph_labels([L] * 3, x, n, cut)
            assert np.array_equal(lab, ref)
            say("components", nc, "largest", int(np.bincount(lab).max()))
    reps = wl.replicas(5, seed=8, nmin=300, nmax=2000)
    with mdx.engine_from_system(None, replicas=reps) as e:
        for r in (0, 3):
            Lr = float(reps[r]["box"][0])
            nr = max(int(Lr // 3.5), 1)
            lab, _ = e.clusters(nr, 1.3, (10, 40), replica=r)
            assert np.array_equal(lab, po.dgraph_labels([Lr] * 3, reps[r]["x"], nr, 1.3, 10, 40))


def test_autotune_tries_every_candidate_and_keeps_the_fastest(monkeypatch):
    monkeypatch.setenv("MDX_SPLIT", "1")
    s = wl.condensate_ab(1 << 16, phi=0.1)
    with engine_for(s, mdx.PREC_MIXED, seed=2) as e:
        e.prime()
        rep = e.autotune(10)
        say("autotune", rep)
        assert len(rep) == 8 and all(r["ms_per_step"] > 0 for r in rep)
        best = min(rep, key=lambda r: r["ms_per_step"])
        assert e.info()["skin"] == best["skin"]
        e.step(20)
        assert np.isfinite(e.download()["x"]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name,builder,fuse,steps,precision,tol", [
    ("condensate", lambda: wl.condensate_ab(), "2", 200, mdx.PREC_FP64, 1e-8),
    ("condensate", lambda: wl.condensate_ab(), "1", 200, mdx.PREC_FP64, 1e-8),
    ("mt_bundle", lambda: wl.mt_bundle(), "2", 500, mdx.PREC_FP64, 1e-8),
    ("polymer_fa0", lambda: wl.active_polymer(fa=0.0), "2", 200, mdx.PREC_FP64, 1e-8),
    ("condensate", lambda: wl.condensate_ab(), "2", 200, mdx.PREC_MIXED, 1e-2),
])
def test_full_size_momentum_conserved_without_thermostat(monkeypatch, name, builder, fuse, steps,
                                                         precision, tol):
    """Full-size systems stepped through the production list path (split
    kernels, compacted walk, gated list rebuilds; fuse "1" also integrates
    inside the walk) in FP64 with gamma = gammar = 0: the 1M two-species
    condensate (WCA + patches), the 416,000-bead microtubule bundle (WCA +
    harmonic bonds + cosine bending, bead integrator) and the 2.1M-bead
    polymer melt with its propulsion switched off (fa = 0: WCA + bonds +
    bending + patchy chain-end rotors).  The noise amplitude
    sqrt(2 gamma kT / dt) is then zero and there is no friction, so each half
    kick adds dt/2 * F and the total momentum moves only by the sum of the
    forces, which pair and three-body forces make zero (Newton's third law;
    MD.cpp:1684-1764, MD.cpp:1929-1983 and LangevinR.cpp:911-1047 add +F to i
    and -F to j, and the three angle terms sum to zero).  A pair seen from one
    side only, or twice from one side, moves it by about dt * |F_ij|, orders
    above the FP64 roundoff allowed here.  A particle moving more than half
    the list skin (0.4) forces a rebuild, so the run crosses at least one.
    The mixed-precision condensate keeps p in FP32: each step rounds every
    component by about 2^-24 |p|, a random walk of about
    6e-8 * sqrt(N * steps) = 9e-4 in the sum (7.6e-4 measured); its bound
    1e-2 stays below the ~0.05 that one contact pair seen from one side for
    one step adds (dt = 0.005, |F| ~ 10)."""
    monkeypatch.setenv("MDX_FUSE", fuse)
    s = wl.thermal_momenta(builder())
    s["gamma"] = 0.0
    s["gammar"] = 0.0
    say("momentum conservation", name, s["N"], "fuse", fuse, "steps", steps, "precision", precision)
    with mdx.engine_from_system(s, precision=precision, seed=5) as e:
        st0 = e.download()
        if s["integrator"] == "rotor":
            e.prime()
        e.step(steps)
        st = e.download()
    P0, P = st0["p"].sum(0), st["p"].sum(0)
    box = np.asarray(s["box"], dtype=np.float64)
    dx = st["x"] - st0["x"]
    dx -= box * np.round(dx / box)
    say("P0", P0, "P", P, "|dP|", np.abs(P - P0).max(), "max displacement", np.abs(dx).max())
    assert np.isfinite(st["p"]).all() and np.isfinite(st["x"]).all()
    assert np.abs(dx).max() > 0.2
    assert np.abs(P - P0).max() < tol
