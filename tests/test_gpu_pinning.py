"""Reductions, noise stream, resume, reflecting walls and the failure path of
the engine on an MI355X, through the C ABI, against the oracle and the
reference fixtures."""
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


# ---------------------------------------------------------------- reductions
@pytest.mark.parametrize("name,builder", [
    ("c1", lambda: wl.patchy_tetra()),
    ("c3", lambda: wl.condensate_ab(4096)),
    ("bundle", lambda: wl.mt_bundle(1, 13, 40, spacing=6.0, box=(10.0, 10.0, 60.0))),
    ("polymer", lambda: wl.active_polymer(nchains=24, nbeads=32, phi=0.1, seed=3, fa=0.0)),
])
def test_observables_fp64_vs_oracle(name, builder):
    """mdx_observables on a frozen FP64 state with thermal momenta: the
    kinetic temperatures per component against orc_measured_temperature
    (LangevinR.cpp:170-193, pinned by the reference fixture), the potential
    energy and the virial sum_i F_i x_i against orc_energy_virial (pinned by
    the reference's energies and by -dU/d strain), all within 1e-12"""
    s = wl.thermal_momenta(builder())
    say("observables", name, s["N"])
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64) as e:
        o = e.observables()
    U, W = po.energy_virial(s, list_cut=3.5)
    if s["integrator"] == "rotor":
        kt = po.measured_temperature(s, s["p"], s["L"])
    else:
        kt = np.concatenate([np.mean(s["p"] ** 2 / (2 * s["m"]), 0), np.zeros(3)])
    got = np.concatenate([o["ke_t"], o["ke_r"]])
    say("ke", got, "ref", kt, "U", o["epot"], U)
    assert relmax(got, kt) < 1e-12
    assert abs(o["epot"] - U) < 1e-12 * max(1.0, abs(U))
    assert relmax(o["virial"], W) < 1e-12


@pytest.mark.parametrize("tag,builder", [("c1", lambda: wl.patchy_tetra()),
                                         ("c3", lambda: wl.condensate_ab(4096))])
def test_stress_vs_reference(golden, tag, builder):
    """mdx_stress against the reference's MD::calculatestress (fixture) and
    the oracle on a state after steps"""
    g = golden("observables.npz")
    s = builder()
    assert np.array_equal(s["x"], g[f"stress_{tag}_x"])
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64) as e:
        st0 = e.stress()
        e.prime()
        e.step(30)
        x = e.download()["x"]
        st1 = e.stress()
    say("stress", tag, st0[0, 0], g[f"stress_{tag}_out"][0, 0])
    assert relmax(st0, g[f"stress_{tag}_out"]) < 1e-12
    L = float(s["box"][0])
    ref1 = po.stress(L, x, po.pairs_reference(L, x, int(L // 4), 3.5))
    assert relmax(st1, ref1) < 1e-12


# -------------------------------------------------------------------- noise
def recover_u(R):
    """the 32-bit Philox output behind a noise value of the reference's
    uniform transform 3.464101615 ((u + 0.5) 2^-32) - 1.732050808"""
    f = (R + 1.732050808) / 3.464101615
    return np.rint(f * 2.0 ** 32 - 0.5).astype(np.uint64)


def philox_rows(ids, counter, key, lanes):
    out = []
    for i in ids:
        row = []
        for half in range(2):
            c = np.array([i, counter & 0xffffffff, counter >> 32, half], dtype=np.uint32)
            row.extend(po.philox(c, key))
        out.append(row[:lanes] if lanes == 3 else row[:4] + row[4:6])
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("seed", [0, (7 << 32) | 12345])
def test_device_philox_stream_vs_oracle(seed):
    """The noise the step draws is Philox4x32-10 keyed by the engine seed with
    counter (particle, k lo, k hi, 0 / 1) for force evaluation k: recovered
    bit for bit from one force-free priming evaluation (p = L = 0, Q = 1:
    the carried F = R_t R[0:3] and T = R_r R[3:6], LangevinR.cpp:475-578)
    and compared with the oracle's Philox (pinned by the Random123
    known-answer vectors), at k = 0 and at a k above 2^32 set with
    mdx_set_noise_counter; and for the bead step (p1 = (c5 c4 + r) R,
    Langevin.cpp:254-269)"""
    N = 300
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 20, (N, 3))
    key = np.array([seed & 0xffffffff, seed >> 32], dtype=np.uint32)
    gamma, gammar, kT, dt = 1.0, 1.0 / 3.0, 1.0, 0.005
    Rt, Rr = np.sqrt(2 * gamma * kT / dt), np.sqrt(2 * gammar * kT / dt)
    with mdx.Engine(N, 20.0, kT, dt, gamma=gamma, gammar=gammar, seed=seed) as e:
        e.upload(x)
        for k in (0, (1 << 32) + 77):
            e.set_noise_counter(k)
            e.prime()
            assert e.noise_counter() == k + 1
            F, T = e.carried_forces()
            got = recover_u(np.concatenate([F / Rt, T / Rr], 1))
            ref = philox_rows(range(N), k, key, 6)
            say("rotor noise k", k, "first", got[0, :2], ref[0, :2])
            assert np.array_equal(got, ref)
    with mdx.Engine(N, 20.0, kT, 0.002, gamma=gamma, integrator=mdx.INTEG_BEAD, seed=seed) as e:
        e.upload(x)
        k = (3 << 32) + 5
        e.set_noise_counter(k)
        e.step(1)
        p = e.download()["p"]
        dtb = 0.002
        d = gamma * dtb / 2
        c2, c5 = 1.0 / (1.0 + d), 1 - d
        rr = np.sqrt(0.5 * kT * gamma * 1.0 * dtb)
        got = recover_u(p / (c5 * (c2 * rr) + rr))
        assert np.array_equal(got, philox_rows(range(N), k, key, 3))


# ------------------------------------------------------------------- resume
@pytest.mark.parametrize("precision", [mdx.PREC_FP64, mdx.PREC_MIXED])
def test_resume_bitwise(precision):
    """N steps, download (state, carried forces, noise counter), a new
    engine, upload, N more steps = 2N uninterrupted steps, bit for bit in
    FP64.  The re-sort period (every 25th step) puts a re-sort at step N + 1
    of the uninterrupted run, where the continuation re-sorts too; cells walk
    their members in ascending original index, so both then evaluate every
    sum in the same order.  The mixed engine holds orientations as FP32 unit
    quaternions, which the N x 9 matrices of the host interface carry only
    to FP32 rounding (the norm is not kept): its continuation starts within
    1e-6 of the uninterrupted run and stays within 1e-3 over the N steps."""
    s = wl.thermal_momenta(wl.condensate_ab(1 << 16, phi=0.1))
    n = 50
    with mdx.engine_from_system(s, precision=precision, seed=21) as a:
        a.set_neighbor_list(0.4, 64, 24)
        a.prime()
        a.step(n)
        a.step(n)
        full = a.download()
        ca = a.noise_counter()
    with mdx.engine_from_system(s, precision=precision, seed=21) as b0:
        b0.set_neighbor_list(0.4, 64, 24)
        b0.prime()
        b0.step(n)
        half = b0.download()
        F, T = b0.carried_forces()
        c = b0.noise_counter()
    t = dict(s)
    t.update({k: half[k] for k in ("x", "p", "Q", "L")})
    with mdx.engine_from_system(t, precision=precision, seed=21) as b:
        b.set_neighbor_list(0.4, 64, 24)
        b.upload_forces(F, T)
        b.set_noise_counter(c)
        b.step(n)
        resumed = b.download()
        assert b.noise_counter() == ca
    for k in ("x", "p", "Q", "L"):
        say(k, np.abs(resumed[k] - full[k]).max())
        if precision == mdx.PREC_FP64:
            assert np.array_equal(resumed[k], full[k]), k
        else:
            assert relmax(resumed[k], full[k]) < 1e-3, k


def test_resume_continues_the_noise_stream():
    """Without the counter a resumed run would repeat the noise of the first
    evaluations: with it, the first step of the continuation draws the
    noise of evaluation c (checked by the change it makes against a run
    that reuses counter 1)"""
    s = wl.thermal_momenta(wl.patchy_tetra(2048))
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=5) as a:
        a.prime()
        a.step(10)
        st = a.download()
        F, T = a.carried_forces()
        c = a.noise_counter()
    outs = []
    for counter in (c, 1):
        t = dict(s)
        t.update({k: st[k] for k in ("x", "p", "Q", "L")})
        with mdx.engine_from_system(t, precision=mdx.PREC_FP64, seed=5) as b:
            b.upload_forces(F, T)
            b.set_noise_counter(counter)
            b.step(1)
            outs.append(b.download()["p"])
    assert c == 11
    assert not np.array_equal(outs[0], outs[1])


# ------------------------------------------------------------ reflecting walls
def test_rotor_trajectory_reflecting_walls_vs_reference(golden):
    """1000 FP64 steps with reflecting walls (the reference's default cube,
    pb = false; geometry.h:433-444) and injected noise against the
    reference's Condensate::run loop"""
    g = golden("traj_rotor_walls256.npz")
    N = len(g["x"])
    s = wl.patchy_tetra(N)
    s["periodic"] = False
    nsteps, snap = int(g["nsteps"]), int(g["snap_every"])
    noise = po.noise(int(g["noise_seed"]), 0, (nsteps + 1) * N * 6).reshape(nsteps + 1, N, 6)
    with mdx.engine_from_system(s, precision=mdx.PREC_FP64) as e:
        e.prime(noise[0])
        errs = []
        for k in range(nsteps // snap):
            e.step(snap, noise[1 + k * snap: 1 + (k + 1) * snap])
            st = e.download()
            errs.append(float(np.max(np.abs(st["x"] - g["snaps"][k]))))
    say("walls max |dx| per 100 steps", errs)
    assert errs[0] < 1e-10
    assert np.max(np.abs(st["x"] - g["x_out"])) < 1e-6
    assert relmax(st["p"], g["p_out"]) < 1e-5
    assert st["x"].min() >= 0 and st["x"].max() <= float(g["box"])


# -------------------------------------------------------------- failure path
def _pair_system(r, precision_note=""):
    s = wl.patchy_tetra(64, phi=0.01)
    c = s["box"] / 2
    s["x"][0] = c - np.array([r / 2, 0, 0])
    s["x"][1] = c + np.array([r / 2, 0, 0])
    s["patches"]["params"][:, 0] = 0.0
    return s


def test_overlap_in_mixed_precision_stops_with_finite_state():
    """A pair uploaded at r = 1e-3 makes the FP32 WCA force overflow (the
    guard off): the priming evaluation reports MDX_ERR_NAN, the state stays
    finite, and steps after it move no particle and report the error again,
    naming the step; with the guard on (1e4, MD.cpp:1711-1749) a pair at
    r = 0.6 (|f| = 3.6e4) reports MDX_ERR_FORCE"""
    s = _pair_system(1e-3)
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=1) as e:
        e.set_force_guard(0.0)
        x0 = e.download()["x"]
        with pytest.raises(mdx.MdxError) as ex:
            e.prime()
        say("prime:", ex.value)
        assert ex.value.code == mdx.MDX_ERR_NAN
        with pytest.raises(mdx.MdxError) as ex:
            e.step(5)
        say("step:", ex.value)
        assert ex.value.code == mdx.MDX_ERR_NAN and "step 1" in str(ex.value)
        st = e.download()
        for k in ("x", "p", "Q", "L"):
            assert np.isfinite(st[k]).all(), k
        assert np.array_equal(st["x"][:2], x0[:2])
    s = _pair_system(0.6)
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=1) as e:
        with pytest.raises(mdx.MdxError) as ex:
            e.prime()
        say("guard:", ex.value)
        assert ex.value.code == mdx.MDX_ERR_FORCE
        assert np.isfinite(e.download()["x"]).all()


def test_guard_trips_mid_run_and_later_steps_hold():
    """Two particles driven into each other: the guard trips at some step k
    of the call (named in the error); the run stops there: positions stay
    finite and inside the box, and the step counter still advances"""
    s = _pair_system(1.05)
    s["p"] = np.zeros((s["N"], 3))
    s["p"][0, 0], s["p"][1, 0] = 60.0, -60.0
    with mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=2) as e:
        e.prime()
        with pytest.raises(mdx.MdxError) as ex:
            e.step(50)
        say("mid-run:", ex.value)
        assert ex.value.code in (mdx.MDX_ERR_FORCE, mdx.MDX_ERR_NAN)
        assert "at step" in str(ex.value)
        st = e.download()
        assert np.isfinite(st["x"]).all() and np.isfinite(st["p"]).all()
        assert st["x"].min() >= 0 and st["x"].max() <= s["box"][0]
