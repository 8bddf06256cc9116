"""The one-bond-per-patch binding on the GPU (mdx_bind_forces, through the C
ABI) against the oracle's restatement of
calculate_forces_and_torques3D_onlyone_nonlets (Rotational/LangevinRforce.cpp:
2874-4136), which tests/test_oracle_bind.py pins to the compiled reference.
Both draw Philox4x32-10 at {smallest patch of the component, draw, k} under
the engine's binding key.  Bars: bound state (isbound, boundto) bit-exact;
binding forces and torques within 1e-12 (FP64) / 1e-5 (mixed) of the oracle,
relative to the largest value."""
import os
import sys

import numpy as np
import pytest

import pyoracle as po
from mtcode_amd import mdx
from mtcode_amd import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_bind_golden as g  # noqa: E402

pytestmark = pytest.mark.gpu

SEED = 77
KEY = (SEED & 0xFFFFFFFF, (SEED >> 32) ^ 0xB1D0B1D0)   # the engine's binding key
Z = np.load(os.path.join(HERE, "golden", "bind.npz"))
AB, WIDE = g.systems()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def bind_engine(s, prec, on=g.ON, off=g.OFF):
    e = mdx.engine_from_system(s, precision=prec, seed=SEED)
    nh = e.set_binding(on, off)
    return e, nh


@pytest.mark.parametrize("case", ["ab0", "ab1", "wide0", "wide1"])
def test_bind_fp64_vs_oracle_two_calls(case):
    s = AB if case.startswith("ab") else WIDE
    say("bind fp64", case)
    e, nh = bind_engine(s, mdx.PREC_FP64)
    ib0, bt0 = Z[case + "_isbound_in"], Z[case + "_boundto_in"]
    e.set_bindings(ib0, bt0)
    F, T = e.bind_forces(draw=5)
    ib, bt = e.get_bindings()
    Fo, To, ibo, bto, st = po.bind_forces(s, ib0, bt0, g.ON, g.OFF, key=KEY, draw=5)
    assert np.array_equal(ib, ibo) and np.array_equal(bt, bto)
    assert rel(F, Fo) < 1e-12 and rel(T, To) < 1e-12
    # the next call starts from the state the device holds
    F2, T2 = e.bind_forces(draw=6)
    ib2, bt2 = e.get_bindings()
    Fo2, To2, ibo2, bto2, _ = po.bind_forces(s, ibo, bto, g.ON, g.OFF, key=KEY, draw=6)
    assert np.array_equal(ib2, ibo2) and np.array_equal(bt2, bto2)
    assert rel(F2, Fo2) < 1e-12 and rel(T2, To2) < 1e-12
    if case.startswith("wide"):
        assert st[2] > 0   # components of 4+ patches


@pytest.mark.parametrize("case", ["ab1", "wide1"])
def test_bind_mixed_vs_oracle(case):
    """mixed precision: the decisions are FP64 on the stored (FP32
    quaternion) orientations; on these configurations no decision sits
    within that rounding of its threshold, so the bound sets are equal"""
    s = AB if case.startswith("ab") else WIDE
    say("bind mixed", case)
    e, nh = bind_engine(s, mdx.PREC_MIXED)
    ib0, bt0 = Z[case + "_isbound_in"], Z[case + "_boundto_in"]
    e.set_bindings(ib0, bt0)
    F, T = e.bind_forces(draw=9)
    ib, bt = e.get_bindings()
    Fo, To, ibo, bto, _ = po.bind_forces(s, ib0, bt0, g.ON, g.OFF, key=KEY, draw=9)
    assert np.array_equal(ib, ibo) and np.array_equal(bt, bto)
    assert rel(F, Fo) < 1e-5 and rel(T, To) < 1e-5


W = np.load(os.path.join(HERE, "golden", "bind_wall.npz"))


@pytest.mark.parametrize("case", ["wall0", "wall1"])
def test_bind_walls_fp64_vs_oracle(case):
    """reflecting walls (the condensate's default box; the oracle is pinned
    to the reference on it): FP64 bound state bit-exact and forces within
    1e-12 over two calls"""
    s = g.walled(WIDE)
    say("bind walls fp64", case)
    e, nh = bind_engine(s, mdx.PREC_FP64)
    ib0, bt0 = W[case + "_isbound_in"], W[case + "_boundto_in"]
    e.set_bindings(ib0, bt0)
    for draw in (11, 12):
        F, T = e.bind_forces(draw=draw)
        ib, bt = e.get_bindings()
        Fo, To, ib0, bt0, _ = po.bind_forces(s, ib0, bt0, g.ON, g.OFF, key=KEY, draw=draw)
        assert np.array_equal(ib, ib0) and np.array_equal(bt, bt0)
        assert rel(F, Fo) < 1e-12 and rel(T, To) < 1e-12


PL = np.load(os.path.join(HERE, "golden", "plist.npz"))


@pytest.mark.parametrize("case", ["ab", "wide", "wall"])
def test_patch_list_vs_reference(case):
    """mdx_patch_list and mdx_bound_pairs on the reference's own pairs: FP64
    rows bit-exact against the compiled reference's calculate_patch_list and
    calculate_bound_pairs (order included);
    mixed precision (FP32 orientations) differs only in rows within rounding
    of the -0.4 threshold (at most 0.1 % of the rows)"""
    s = {"ab": AB, "wide": WIDE, "wall": g.walled(WIDE)}[case]
    ref = PL[case + "_plist"]
    say("patch list fp64", case)
    e = mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=SEED)
    rows = e.patch_list(PL[case + "_pairs"])
    assert np.array_equal(rows, ref)
    assert np.array_equal(po.adjacency(rows), PL[case + "_adjacency"])
    assert np.array_equal(e.patch_list(PL[case + "_pairs"], bound=True), PL[case + "_blist"])
    # a small cap reports the row count with MDX_ERR_CAPACITY
    with pytest.raises(mdx.MdxError) as ex:
        e.patch_list(PL[case + "_pairs"], cap=10)
    assert ex.value.code == mdx.MDX_ERR_CAPACITY
    say("patch list mixed", case)
    e = mdx.engine_from_system(s, precision=mdx.PREC_MIXED, seed=SEED)
    rows = e.patch_list(PL[case + "_pairs"])
    a = {tuple(r) for r in rows}
    b = {tuple(r) for r in ref}
    assert len(a ^ b) <= 1e-3 * len(b)


def test_patch_lists_full_size_properties():
    """the 1,048,576-particle condensate (mixed, the bench system) on the
    pairs of calculatepairs(2.5) (mdx_neighbor_pairs): every row has
    p1 < p2 and patches of its own particles; the bound-pair list lies inside
    the patch list (both cones above cos(theta_m) = 0.95 pass the -0.4
    allowance at max_ang 0.6, and max_check < max_check + 1); every bond the
    binding routine forms is a row of the bound-pair list"""
    s = wl.condensate_ab()
    say("patch lists full size")
    e, nh = bind_engine(s, mdx.PREC_MIXED, 0.5, 0.5)
    n = int(float(s["box"][0]) // 2.5)
    pairs = e.neighbor_pairs(n, 2.5)
    e.bind_forces(draw=0)
    ib, bt = e.get_bindings()
    rows_p = e.patch_list(pairs, cap=2 * len(pairs))
    rows_b = e.patch_list(pairs, cap=len(pairs), bound=True)
    say("  pairs", len(pairs), "patch list", len(rows_p), "bound pairs", len(rows_b))
    owner = np.repeat(np.arange(s["N"]), [s["patches"]["npatch"][t] for t in s["patches"]["types"]])
    for rows in (rows_p, rows_b):
        assert np.all(rows[:, 0] < rows[:, 1])
        assert np.all(owner[rows[:, 3]] == rows[:, 0]) and np.all(owner[rows[:, 4]] == rows[:, 1])
    key_p = rows_p[:, 3].astype(np.int64) * nh + rows_p[:, 4]
    key_b = rows_b[:, 3].astype(np.int64) * nh + rows_b[:, 4]
    assert np.all(np.isin(key_b, key_p))
    a = np.nonzero(ib)[0]
    b = bt[a]
    lo = np.where(owner[a] < owner[b], a, b)
    hi = np.where(owner[a] < owner[b], b, a)
    assert len(a) > 0
    assert np.all(np.isin(lo.astype(np.int64) * nh + hi, key_b))


def test_patch_list_rejects_bad_input():
    """host checks of mdx_patch_list: a pair index outside 0..N-1 is
    MDX_ERR_ARG; particles not numbered by type (GeneralPatch's type ranges)
    are MDX_ERR_STATE; both before any kernel runs"""
    say("patch list errors")
    e = mdx.engine_from_system(AB, precision=mdx.PREC_FP64, seed=SEED)
    with pytest.raises(mdx.MdxError) as ex:
        e.patch_list(np.array([[0, AB["N"]]]))
    assert ex.value.code == mdx.MDX_ERR_ARG
    assert len(e.patch_list(np.zeros((0, 2), np.int32))) == 0
    s = dict(AB)
    s["patches"] = dict(AB["patches"], types=AB["patches"]["types"][::-1].copy())
    e = mdx.engine_from_system(s, precision=mdx.PREC_FP64, seed=SEED)
    with pytest.raises(mdx.MdxError) as ex:
        e.patch_list(PL["ab_pairs"][:10])
    assert ex.value.code == mdx.MDX_ERR_STATE


def lattice_system(m, thetam, d, seed):
    s = g.lattice(wl.patchy_tetra(m ** 3, phi=0.05, seed_pos=seed, seed_q=seed + 1), m, 1.15,
                  seed + 2)
    s["patches"]["params"][:, 1] = d
    s["patches"]["params"][:, 2] = thetam
    return s


def test_bind_32k_vs_oracle_and_rates():
    """32,768 particles (131,072 patches), wide cones: the same bound state
    and forces as the oracle over three calls, with on/off rates 0.6/0.4"""
    s = lattice_system(32, 1.0, 1.1, 41)
    say("bind 32k")
    e, nh = bind_engine(s, mdx.PREC_FP64, 0.6, 0.4)
    ibo = np.zeros(nh, np.int32)
    bto = np.zeros(nh, np.int32)
    for draw in range(3):
        F, T = e.bind_forces(draw=draw)
        ib, bt = e.get_bindings()
        Fo, To, ibo, bto, st = po.bind_forces(s, ibo, bto, 0.6, 0.4, key=KEY, draw=draw)
        say("  call", draw, "edges", st[0], "large", st[2], "bound pairs", st[4])
        assert np.array_equal(ib, ibo) and np.array_equal(bt, bto)
        assert rel(F, Fo) < 1e-12 and rel(T, To) < 1e-12


def test_bind_full_size_properties():
    """the 1,048,576-particle condensate (the bench system) at its own
    packing: every bond mutual and between different particles, the bound
    pair forces sum to zero (third law), a doublet binds with probability
    on / (on + off) (Philox), and a second call keeps the invariants"""
    s = wl.condensate_ab()
    say("bind full size")
    e, nh = bind_engine(s, mdx.PREC_MIXED, 0.5, 0.5)
    owner = np.repeat(np.arange(s["N"]), [s["patches"]["npatch"][t] for t in s["patches"]["types"]])
    for draw in range(2):
        F, T = e.bind_forces(draw=draw)
        ib, bt = e.get_bindings()
        b = np.nonzero(ib)[0]
        assert np.all(ib[bt[b]] == 1) and np.all(bt[bt[b]] == b)
        assert np.all(owner[b] != owner[bt[b]])
        assert np.abs(F.sum(axis=0)).max() < 1e-6 * max(1.0, np.abs(F).max()) * np.sqrt(len(b) + 1)
        say("  call", draw, "bound patches", len(b))
    assert len(b) > 0


def test_bind_component_capacity_is_an_error():
    """cones of 86 degrees on a dense lattice join more than 64 patches into
    one component: the call returns MDX_ERR_CAPACITY and the engine stays
    usable"""
    s = lattice_system(8, 1.5, 1.1, 51)
    say("bind capacity")
    e, nh = bind_engine(s, mdx.PREC_FP64)
    with pytest.raises(mdx.MdxError) as ex:
        e.bind_forces(draw=0)
    assert ex.value.code == mdx.MDX_ERR_CAPACITY
    F = e.compute_forces()[0]
    assert np.all(np.isfinite(F))


def oracle_bind_run(s, nsteps, noise, on, off, draws, tables=None):
    """the loop of Condensate::run_singlebond (Condensate.cpp:464-478) on the
    oracle: half kick, drift, rotation, F = WCA + the binding routine's
    bound-pair terms, the sphere's friction and noise, half kick"""
    x, p, Q, L = (np.array(s[k], dtype=np.float64) for k in ("x", "p", "Q", "L"))
    ps = s["patches"]
    nh = int(sum(ps["npatch"][t] for t in ps["types"]))
    ib, bt = np.zeros(nh, np.int32), np.zeros(nh, np.int32)
    sw = dict(s, patches=None)

    def forces(x, Q, ib, bt, draw):
        F, T = po.compute_forces(sw, x=x, Q=Q)
        Fb, Tb, ib, bt, _ = po.bind_forces(s, ib, bt, on, off, key=KEY, draw=draw, x=x, Q=Q,
                                           tables=tables)
        return F + Fb, T + Tb, ib, bt

    F, T, ib, bt = forces(x, Q, ib, bt, draws[0])
    F, T = po.sphere(s, p, L, Q, F, T, noise[0])
    for k in range(1, nsteps + 1):
        p, L = po.halfkick(s, p, L, F, T)
        x, p = po.advance_pos(s, x, p)
        L, Q = po.rotate(s, L, Q)
        F, T, ib, bt = forces(x, Q, ib, bt, draws[k])
        F, T = po.sphere(s, p, L, Q, F, T, noise[k])
        p, L = po.halfkick(s, p, L, F, T)
    return {"x": x, "p": p, "Q": Q, "L": L}, ib, bt


@pytest.mark.parametrize("skin,walls", [(0.0, False), (0.4, False), (0.4, True)])
def test_bind_steps_fp64_vs_oracle(skin, walls):
    """20 steps of the binding loop with injected noise (cell walk, and the
    neighbour list with its rebuilds; periodic box and reflecting walls): the
    bound state equal to the oracle's and the state within 1e-10"""
    s = wl.thermal_momenta(lattice_system(11, 1.0, 1.1, 61), seed=5)
    if walls:
        s = g.walled(s)
    N, nsteps = s["N"], 20
    noise = po.noise(123, 0, (nsteps + 1) * N * 6).reshape(nsteps + 1, N, 6)
    say("bind steps fp64 skin", skin, "walls", walls)
    e, nh = bind_engine(s, mdx.PREC_FP64, 0.7, 0.3)
    e.set_neighbor_list(skin, 64, 200)
    c0 = e.noise_counter()
    e.prime(noise[0])
    e.step(nsteps, noise[1:])
    st = e.download()
    ib, bt = e.get_bindings()
    ref, ibo, bto = oracle_bind_run(s, nsteps, noise, 0.7, 0.3, [c0 + k for k in range(nsteps + 1)])
    assert ibo.sum() > 0
    assert np.array_equal(ib, ibo) and np.array_equal(bt, bto)
    for k in ("x", "Q"):
        assert np.max(np.abs(st[k] - ref[k])) < 1e-10, k
    for k in ("p", "L"):
        assert rel(st[k], ref[k]) < 1e-9, k


def test_bind_steps_mixed_long():
    """200 mixed-precision steps of the binding loop on 32,768 particles
    (Philox noise, neighbour list): bonds stay mutual, the state finite, and
    the bound fraction near the FP64 engine's"""
    s = wl.thermal_momenta(lattice_system(32, 0.8, 1.1, 71), seed=6)
    say("bind steps mixed 32k")
    frac = {}
    for prec in (mdx.PREC_MIXED, mdx.PREC_FP64):
        e, nh = bind_engine(s, prec, 0.7, 0.3)
        e.prime()
        e.step(200)
        st = e.download()
        ib, bt = e.get_bindings()
        b = np.nonzero(ib)[0]
        assert np.all(ib[bt[b]] == 1) and np.all(bt[bt[b]] == b)
        assert all(np.all(np.isfinite(v)) for v in st.values())
        frac[prec] = len(b) / nh
    say("  bound fraction", frac)
    assert abs(frac[mdx.PREC_MIXED] - frac[mdx.PREC_FP64]) < 0.02


@pytest.mark.parametrize("case", ["bin0", "bin1", "ter0", "ter1"])
def test_bind_table_models_vs_oracle(case):
    """BindingModelBinary / BindingModelTernary through
    mdx_set_binding_tables: FP64 bound state bit-exact and forces within
    1e-12 of the oracle's species-table form (pinned to the reference's
    models), over two calls"""
    tag = case[:3]
    tabs = {q: Z[f"{tag}_{q}"] for q in ("species", "dtab", "ttab", "tr4")}
    say("bind table model", case)
    e, nh = bind_engine(WIDE, mdx.PREC_FP64)
    e.set_binding_tables(tabs["species"], tabs["dtab"], tabs["ttab"], tabs["tr4"])
    ib0, bt0 = Z[case + "_isbound_in"], Z[case + "_boundto_in"]
    e.set_bindings(ib0, bt0)
    for draw in (3, 4):
        F, T = e.bind_forces(draw=draw)
        ib, bt = e.get_bindings()
        Fo, To, ib0, bt0, st = po.bind_forces(WIDE, ib0, bt0, 0.0, 0.0, key=KEY, draw=draw,
                                              tables=tabs)
        assert np.array_equal(ib, ib0) and np.array_equal(bt, bt0)
        assert rel(F, Fo) < 1e-12 and rel(T, To) < 1e-12
