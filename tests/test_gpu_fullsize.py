"""Full-size parity of the production step path on an MI355X, through the C ABI.

The step's own kernels (device-side binning, the neighbour list built by
k_nbuild, the compacted walk, the patch and integrator kernels, gated list
rebuilds and state re-sorts) are compared with the oracle at the full sizes
of the benchmark systems:
  - the list k_nbuild builds, as a pair set, equals the oracle's pair set at
    the list radius, bit for bit (1,048,576-particle condensate; the 256
    replicas of uneven size);
  - the forces and torques the step carries after production steps equal the
    oracle's at the downloaded state (FP64 <= 1e-12 of the largest), with the
    friction and the noise amplitude set to zero so that the carried values
    are the conservative ones (the sphere update then only rotates the
    torque into the body frame: LangevinR.cpp:475-578);
  - the same for the 416,000-bead bundle in its 200 x 200 x 600 box (bead
    integrator: with gamma = 0 one step adds dt F to p exactly,
    Langevin.cpp:254-269), which covers the orthorhombic box;
  - per-replica momentum conservation of the 256-replica batch without a
    thermostat across list rebuilds.
"""
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


@pytest.fixture(scope="module")
def c4():
    say("building the 256 replicas")
    return [wl.thermal_momenta(r, seed=99 + k) for k, r in enumerate(wl.replicas())]


@pytest.fixture(scope="module")
def c3():
    say("building the 1M condensate")
    return wl.thermal_momenta(wl.condensate_ab())


def mimage_r2(xa, xb, box):
    """r^2 with the engine's and the reference's minimum image
    (cube::distance_vector, geometry.h:277-294), in the same FP64 order"""
    d = xa - xb
    half2 = (box * 0.5) * (box * 0.5)
    d = np.where(d * d > half2, d - np.where(d >= 0, box, -box), d)
    return d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]


def keys(ij):
    ij = np.sort(np.asarray(ij, dtype=np.int64), axis=1)
    return np.sort(ij[:, 0] << 32 | ij[:, 1])


def oracle_list(x, box, radii, types=None, active=None):
    """the oracle's pairs within the list radius (MD::calculatepairs with
    cut = reach + skin), less the pairs of type pairs without patch rows
    beyond the WCA reach + skin"""
    rl, rw = radii
    n = np.maximum(np.floor(box / rl), 1).astype(int)
    pr = po.pairs_reference(box, x, n, rl)
    if types is not None and rw < rl:
        r2 = mimage_r2(x[pr[:, 0]], x[pr[:, 1]], box)
        act = active[types[pr[:, 0]], types[pr[:, 1]]]
        pr = pr[act | (r2 < rw * rw)]
    return pr


def test_list_pairs_full_condensate_bit_exact(c3):
    """k_nbuild's list at 1,048,576 particles after 30 mixed-precision
    production steps (live state in cell order) = the oracle's pair set:
    A-B pairs within reach + skin, A-A and B-B pairs (no patch rows) within
    the WCA reach + skin"""
    s = c3
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=3) as e:
        e.prime()
        e.step(30)
        st = e.download()
        say("list pairs 1M")
        got, radii = e.list_pairs()
    box = np.asarray(s["box"], dtype=np.float64)
    t = s["patches"]["types"]
    active = np.array([[False, True], [True, False]])
    ref = oracle_list(st["x"], box, radii, t, active)
    say("pairs", len(got), "oracle", len(ref), "radii", radii)
    assert len(got) > s["N"] // 2
    assert np.array_equal(keys(got), keys(ref))


def test_list_pairs_replica_batch_bit_exact(c4):
    """the list of the 256-replica batch (each replica its own box) equals,
    replica by replica, the oracle's pair set at the list radius"""
    reps = c4
    with mdx.engine_from_system(None, precision=mdx.PREC_MIXED, seed=4, replicas=reps) as e:
        e.prime()
        e.step(20)
        st = e.download()
        say("list pairs, replicas")
        got, radii = e.list_pairs()
    offs = np.concatenate([[0], np.cumsum([r["N"] for r in reps])])
    rep_of = np.searchsorted(offs, got[:, 0], side="right") - 1
    assert np.array_equal(rep_of, np.searchsorted(offs, got[:, 1], side="right") - 1)
    total = 0
    for r, s in enumerate(reps):
        sl = slice(offs[r], offs[r + 1])
        ref = oracle_list(st["x"][sl], np.asarray(s["box"], dtype=np.float64), radii)
        mine = got[rep_of == r] - offs[r]
        assert np.array_equal(keys(mine), keys(ref)), r
        total += len(ref)
    say("pairs", total)


def carried_vs_oracle(s, st, F, T):
    """carried F (lab), T (body) against the oracle's conservative force and
    the oracle's patch torque rotated into the body frame (orc_sphere with
    gamma = gamma_r = 0 and zero noise)"""
    t = dict(s)
    t["x"], t["Q"] = st["x"], st["Q"]
    Fo, To = po.compute_forces(t, list_cut=2.0)
    Fo, Tb = po.sphere(t, st["p"], st["L"], st["Q"], Fo, To, np.zeros((s["N"], 6)))
    return relmax(F, Fo), relmax(T, Tb)


@pytest.mark.parametrize("fuse", ["2", "1"])
def test_full_condensate_carried_forces_fp64(monkeypatch, c3, fuse):
    """1,048,576 particles, FP64, 25 production steps (a re-sort step, then
    list steps through the compacted walk, the patch kernel and the
    integrator; fuse 1: integration inside the walk and patch kernels), with
    gamma = gamma_r = 0: the carried F, T equal the oracle's to 1e-12"""
    monkeypatch.setenv("MDX_FUSE", fuse)
    s = dict(c3)
    s["gamma"] = s["gammar"] = 0.0
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=5) as e:
        e.prime()
        e.step(25)
        st = e.download()
        F, T = e.carried_forces()
        builds = e.info()["list_builds"]
    ef, et = carried_vs_oracle(s, st, F, T)
    say("fuse", fuse, "builds", builds, "F err", ef, "T err", et)
    assert builds >= 1
    assert ef < 1e-12 and et < 1e-12


def test_replica_batch_forces_fp64_and_mixed(c4):
    """The 256-replica batch: carried F, T after 20 FP64 production steps
    (gamma = 0) against the oracle for the smallest, the largest and 16
    more replicas (<= 1e-12); and mixed-precision frozen forces within 1e-5
    (max) / 1e-6 (RMS) of FP64 over the whole batch"""
    reps = [dict(r, gamma=0.0, gammar=0.0) for r in c4]
    Ns = np.array([r["N"] for r in reps])
    offs = np.concatenate([[0], np.cumsum(Ns)])
    pick = sorted(set([int(np.argmin(Ns)), int(np.argmax(Ns))] +
                      list(np.linspace(0, len(reps) - 1, 16).astype(int))))
    with mdx.engine_from_system(None, precision=mdx.PREC_FP64, seed=6, replicas=reps) as e:
        e.prime()
        e.step(20)
        st = e.download()
        F, T = e.carried_forces()
    worst = 0.0
    for r in pick:
        sl = slice(offs[r], offs[r + 1])
        sub = {k: st[k][sl] for k in st}
        ef, et = carried_vs_oracle(reps[r], sub, F[sl], T[sl])
        worst = max(worst, ef, et)
        assert ef < 1e-12 and et < 1e-12, (r, ef, et)
    say("replicas checked", len(pick), "N range", Ns[pick].min(), Ns[pick].max(), "worst", worst)
    rounded = [dict(r, Q=wl.mixed_orientations(r["Q"])) for r in c4]
    with mdx.engine_from_system(None, precision=mdx.PREC_FP64, replicas=rounded) as e:
        F64, T64 = e.compute_forces()
    with mdx.engine_from_system(None, precision=mdx.PREC_MIXED, replicas=c4) as e:
        F32, T32 = e.compute_forces()
    ef, et = relmax(F32, F64), relmax(T32, T64)
    rms = float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64))
    say("mixed vs fp64 batch: F", ef, "T", et, "rms", rms)
    # the bound is the max-norm 1e-5; the RMS, about 1e-6 (FP32 rounding of
    # a few terms per particle), is held to twice that
    assert ef < 1e-5 and et < 1e-5 and rms < 2e-6


def test_replica_batch_momentum_per_replica(c4):
    """No friction and no noise (gamma = gamma_r = 0): each replica's total
    momentum stays put to FP64 rounding over 200 production steps with list
    rebuilds (a pair missed from one side, or a pair across replicas, would
    move it by about dt |F|)"""
    reps = [dict(r, gamma=0.0, gammar=0.0) for r in c4]
    offs = np.concatenate([[0], np.cumsum([r["N"] for r in reps])])
    with mdx.engine_from_system(None, precision=mdx.PREC_FP64, seed=7, replicas=reps) as e:
        st0 = e.download()
        e.prime()
        e.step(200)
        st = e.download()
        builds = e.info()["list_builds"]
    dP = np.array([np.abs(st["p"][offs[r]:offs[r + 1]].sum(0) -
                          st0["p"][offs[r]:offs[r + 1]].sum(0)).max() for r in range(len(reps))])
    say("builds", builds, "max |dP| per replica", dP.max())
    assert builds >= 3
    assert dP.max() < 1e-9


@pytest.mark.parametrize("fuse", ["2", "0"])
def test_full_bundle_orthorhombic_forces_fp64(monkeypatch, fuse):
    """416,000-bead bundle in its 200 x 200 x 600 box after 40 FP64
    production steps: with gamma = 0 and kT = 0 one bead step adds exactly
    dt F to p (Langevin.cpp:254-269), so F = (p1 - p0) / dt equals the
    oracle's WCA (minus exclusions) + bond + bending force at the state
    before that step to 1e-12 of the largest"""
    monkeypatch.setenv("MDX_FUSE", fuse)
    s = wl.thermal_momenta(wl.mt_bundle())
    s["gamma"] = 0.0
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=8) as e:
        e.step(40)
        st0 = e.download()
        e.step(1)
        st1 = e.download()
    F = (st1["p"] - st0["p"]) / s["dt"]
    t = dict(s)
    t["x"] = st0["x"]
    Fo, _ = po.compute_forces(t, list_cut=2.0)
    err = relmax(F, Fo)
    say("bundle fuse", fuse, "F err", err)
    assert err < 1e-12
