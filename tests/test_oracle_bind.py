"""The oracle's one-bond-per-patch routine (orc_bind_forces) against the
compiled reference's calculate_forces_and_torques3D_onlyone_nonlets
(Rotational/LangevinRforce.cpp:2874-4136, BindingModelSingle) on the fixtures
of tests/golden/make_bind_golden.py: bound sets and partners bit-exact, forces
and torques to 1e-12, with the reference's own rand() draws."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_bind_golden as g  # noqa: E402
import pyoracle as po  # noqa: E402

Z = np.load(os.path.join(HERE, "golden", "bind.npz"))
AB, WIDE = g.systems()


@pytest.mark.parametrize("case", ["ab0", "ab1", "wide0", "wide1"])
def test_oracle_bind_matches_reference(case):
    s = AB if case.startswith("ab") else WIDE
    F, T, ib, bt, st = po.bind_forces(s, Z[case + "_isbound_in"], Z[case + "_boundto_in"], g.ON,
                                      g.OFF, seed=int(Z[case + "_seed"]))
    assert np.array_equal(ib, Z[case + "_isbound_out"])
    assert np.array_equal(bt, Z[case + "_boundto_out"])
    Fr, Tr = Z[case + "_F_out"], Z[case + "_T_out"]
    assert np.abs(F - Fr).max() <= 1e-12 * max(1.0, np.abs(Fr).max())
    assert np.abs(T - Tr).max() <= 1e-12 * max(1.0, np.abs(Tr).max())
    if case.startswith("wide"):
        assert st[2] > 0 and st[3] > 3   # components of 4+ patches were cut


W = np.load(os.path.join(HERE, "golden", "bind_wall.npz"))


@pytest.mark.parametrize("case", ["wall0", "wall1"])
def test_oracle_bind_walls_match_reference(case):
    """reflecting walls (the condensate's default box): the reference's
    routine on the wide-cone system without periodic images, reproduced bit
    for bit; the periodic box gives another bound set on the same inputs"""
    s = g.walled(WIDE)
    ib0, bt0, seed = W[case + "_isbound_in"], W[case + "_boundto_in"], int(W[case + "_seed"])
    F, T, ib, bt, st = po.bind_forces(s, ib0, bt0, g.ON, g.OFF, seed=seed)
    assert np.array_equal(ib, W[case + "_isbound_out"])
    assert np.array_equal(bt, W[case + "_boundto_out"])
    Fr, Tr = W[case + "_F_out"], W[case + "_T_out"]
    assert np.abs(F - Fr).max() <= 1e-12 * max(1.0, np.abs(Fr).max())
    assert np.abs(T - Tr).max() <= 1e-12 * max(1.0, np.abs(Tr).max())
    _, _, ibp, btp, _ = po.bind_forces(WIDE, ib0, bt0, g.ON, g.OFF, seed=seed)
    assert not (np.array_equal(ibp, ib) and np.array_equal(btp, bt))


PL = np.load(os.path.join(HERE, "golden", "plist.npz"))


@pytest.mark.parametrize("case", ["ab", "wide", "wall"])
def test_oracle_patch_list_matches_reference(case):
    """calculate_patch_list (LangevinR.cpp:1702-1853) on the reference's own
    pairs at cut 2.5: every row (p1, p2, potn, patches) and its order, the
    adjacency run starts, and the patch set's max_check / max_ang (0.6 for
    the condensate's cos(theta_m) = 0.95, cos(1.0) for the wide cones)"""
    s = {"ab": AB, "wide": WIDE, "wall": g.walled(WIDE)}[case]
    rows = po.patch_list(s, PL[case + "_pairs"])
    assert len(rows) > 1000
    assert np.array_equal(rows, PL[case + "_plist"])
    assert np.array_equal(po.adjacency(rows), PL[case + "_adjacency"])
    assert np.allclose(po.patch_limits(s["patches"]["params"]), PL[case + "_limits"],
                       rtol=0, atol=1e-15)
    # calculate_bound_pairs (LangevinR.cpp:1854-1998) on the same pairs
    b = po.bound_pairs(s, PL[case + "_pairs"])
    assert len(b) > 10
    assert np.array_equal(b, PL[case + "_blist"])


def test_oracle_patch_list_fixture_pins_the_constants():
    """plausible slips in the patch-list restatement change the two-species
    list, so the reference fixture rejects each: max_ang taken as the
    smallest cos(theta_m) (0.95) without its 0.6 start, the cut at max_check
    instead of max_check + 1, and the bound-pair cone test used in place of
    the -0.4 allowance"""
    ref = PL["ab_plist"]
    mk, ma = po.patch_limits(AB["patches"]["params"])
    assert ma == 0.6 and mk == pytest.approx(1.68)
    for kw in ({"maxang": 0.95}, {"maxdis": mk}):
        assert not np.array_equal(po.patch_list(AB, PL["ab_pairs"], **kw), ref)
    assert not np.array_equal(po.bound_pairs(AB, PL["ab_pairs"]), ref)


def test_oracle_bind_fixture_has_every_path():
    """the wide-cone fixture reaches the large-component cut (sizes to 27),
    breaks bonds and re-forms them"""
    s = WIDE
    _, _, ib, bt, st = po.bind_forces(s, Z["wide0_isbound_in"], Z["wide0_boundto_in"], g.ON,
                                      g.OFF, seed=101)
    assert st[2] > 100 and st[3] >= 10
    assert (Z["wide1_isbound_in"] != Z["wide1_isbound_out"]).any()


def bound_pairs(ib, bt):
    return {(i, int(bt[i])) for i in np.nonzero(ib)[0]}


@pytest.mark.parametrize("case", ["ab1", "wide1"])
def test_oracle_bind_philox_consistent(case):
    """Philox draws: every bond is mutual and is a candidate pair, each patch
    holds at most one bond, and the same key and draw give the same result"""
    s = AB if case.startswith("ab") else WIDE
    key = (0x1234, 0xBEEF)
    r1 = po.bind_forces(s, Z[case + "_isbound_in"], Z[case + "_boundto_in"], g.ON, g.OFF,
                        key=key, draw=7)
    r2 = po.bind_forces(s, Z[case + "_isbound_in"], Z[case + "_boundto_in"], g.ON, g.OFF,
                        key=key, draw=7)
    for a, b in zip(r1[:4], r2[:4]):
        assert np.array_equal(a, b)
    ib, bt = r1[2], r1[3]
    for i, j in bound_pairs(ib, bt):
        assert ib[j] and bt[j] == i and i != j
    F = r1[0]
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(F).max())


def test_oracle_bind_uniform_is_philox():
    """the per-component uniform is Philox4x32-10 (pinned by the Random123
    known-answer vectors elsewhere) of {root, draw lo, draw hi, k}"""
    key, draw = (3, 0xA5A5A5A5), (5 << 32) + 9
    o = po.philox([17, draw & 0xFFFFFFFF, draw >> 32, 2], key)
    want = ((int(o[0]) >> 5) * 67108864.0 + (int(o[1]) >> 6)) / 2.0 ** 53
    assert po.bind_uniform(key, draw, 17, 2) == want
    u = np.array([po.bind_uniform(key, 0, r, 0) for r in range(2000)])
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.03


def test_oracle_bind_on_off_limits():
    """from an unbound start, on_rate 0 and off_rate 1 bind nothing, and
    on_rate 1, off_rate 0 bind (BindingModelSingle's doublet and triplet
    rules)"""
    s = WIDE
    ib0, bt0 = Z["wide0_isbound_in"], Z["wide0_boundto_in"]
    _, _, ib, bt, st = po.bind_forces(s, ib0, bt0, 0.0, 1.0, key=(1, 2), draw=0)
    assert ib.sum() == 0
    _, _, ib, bt, st = po.bind_forces(s, ib0, bt0, 1.0, 0.0, key=(1, 2), draw=0)
    assert st[4] > 0 and ib.sum() == 2 * st[4]


@pytest.mark.parametrize("case", ["bin0", "bin1", "ter0", "ter1"])
def test_oracle_bind_table_models_match_reference(case):
    """BindingModelBinary (setup_equilibrium) and BindingModelTernary with
    SortingFunctionNonUniform (species not monotone in the patch index, so
    triplets are decided in a permuted frame): the oracle's species-table
    form with the tables and species the harness read out of the reference's
    models reproduces the reference's bound state bit for bit"""
    tag = case[:3]
    tabs = {q: Z[f"{tag}_{q}"] for q in ("species", "dtab", "ttab", "tr4")}
    F, T, ib, bt, st = po.bind_forces(WIDE, Z[case + "_isbound_in"], Z[case + "_boundto_in"], 0.0,
                                      0.0, seed=int(Z[case + "_seed"]), tables=tabs)
    assert np.array_equal(ib, Z[case + "_isbound_out"])
    assert np.array_equal(bt, Z[case + "_boundto_out"])
    Fr, Tr = Z[case + "_F_out"], Z[case + "_T_out"]
    assert np.abs(F - Fr).max() <= 1e-12 * max(1.0, np.abs(Fr).max())
    assert np.abs(T - Tr).max() <= 1e-12 * max(1.0, np.abs(Tr).max())
    if tag == "ter":
        sp = tabs["species"]
        assert set(np.unique(sp)) == {1, 2, 3} and np.any(np.diff(sp) < 0)


def test_oracle_bind_table_model_differs_from_single():
    """the table model changes the decisions (a mistaken table lookup would
    not go unnoticed): the ternary tables bind a different set than the
    single model from the same start and draws"""
    tabs = {q: Z[f"ter_{q}"] for q in ("species", "dtab", "ttab", "tr4")}
    a = po.bind_forces(WIDE, Z["ter0_isbound_in"], Z["ter0_boundto_in"], 0.0, 0.0, seed=303,
                       tables=tabs)
    b = po.bind_forces(WIDE, Z["ter0_isbound_in"], Z["ter0_boundto_in"], g.ON, g.OFF, seed=303)
    assert not np.array_equal(a[2], b[2])
