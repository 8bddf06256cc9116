"""Pins of the oracle that do not rely on running the reference: closed forms,
finite differences of the energies, brute force, invariants and the
published Philox known-answer vectors.  Each check is chosen so that a
dropped term, a flipped sign or a transposed operand fails it."""
import math

import numpy as np
import pytest

import pyoracle as po
from mtcode_amd import workloads as wl


def test_fasterfloor_quirks():
    # basic.h:58: truncation for x > 0, truncation - 1 otherwise
    assert po.fasterfloor(2.7) == 2
    assert po.fasterfloor(0.0) == -1
    assert po.fasterfloor(-0.5) == -1
    assert po.fasterfloor(-1.0) == -2
    assert po.fasterfloor(5.0) == 5


def test_cell_assignment_closed_form():
    rng = np.random.default_rng(0)
    L = 10.0
    x = rng.uniform(0, L, (500, 3))
    for n in (1, 2, 5, 7):
        c = po.assign_all(L, x, n)
        cx = np.clip(np.floor(n * x / L).astype(int), 0, n - 1)
        assert np.array_equal(c, cx[:, 0] + n * cx[:, 1] + n * n * cx[:, 2])


def test_stencil_brute_force():
    for n in (1, 2, 3, 4):
        rel = po.boxes_relationships(n)
        for c in range(n ** 3):
            cx, cy, cz = c % n, (c // n) % n, c // (n * n)
            s = sorted({(cx + a) % n + n * ((cy + b) % n) + n * n * ((cz + d) % n)
                        for a in (-1, 0, 1) for b in (-1, 0, 1) for d in (-1, 0, 1)})
            assert list(rel[c]) == s


def brute_pairs(x, L, cut, periodic=True):
    d = x[:, None, :] - x[None, :, :]
    if periodic:
        d = np.where(np.abs(d) > L * 0.5, d - np.sign(d) * L, d)
    r2 = np.sum(d * d, axis=2)
    i, j = np.nonzero(np.triu(r2 < cut * cut, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("periodic", [True, False])
def test_pair_set_brute_force(periodic):
    s = wl.patchy_tetra(700, phi=0.2, seed_pos=9)
    L = float(s["box"][0])
    for cut in (1.12246, 1.549, 3.5):
        n = max(int(L // cut), 1)
        pr = po.pairs_reference(L, s["x"], n, cut, periodic)
        got = {(min(a, b), max(a, b)) for a, b in pr.tolist()}
        assert len(got) == len(pr)
        assert got == brute_pairs(s["x"], L, cut, periodic)


def test_hs_force_is_minus_energy_derivative():
    for r in np.linspace(0.9, 1.12, 23):
        h = 1e-6
        fd = -(po.lib().orc_wca_energy(r + h, 1.0, 1.0, 0) -
               po.lib().orc_wca_energy(r - h, 1.0, 1.0, 0)) / (2 * h)
        f = po.lib().orc_hs_force(r, 1.0, 1.0, 0)
        assert abs(f - fd) < 1e-5 * max(1.0, abs(f))
        # closed form 24 eps (2 sigma^12 / r^13 - sigma^6 / r^7)
        assert abs(f - 24 * (2 / r ** 13 - 1 / r ** 7)) < 1e-12 * abs(f) + 1e-14
    assert po.lib().orc_hs_force(1.12246, 1.0, 1.0, 0) == 0.0
    assert po.lib().orc_hs_force(1.1224599, 1.0, 1.0, 0) != 0.0


def rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def random_patch_pair(rng, dirs, thm):
    """A at 0, B at r*d; patches ia of A and ib of B pointing at each other
    within the cone, away from the cone edge and the theta ~ 0 zone."""
    while True:
        RA = rot(rng.normal(size=3), rng.uniform(0, 6.3))
        RB = rot(rng.normal(size=3), rng.uniform(0, 6.3))
        ia, ib = rng.integers(len(dirs)), rng.integers(len(dirs))
        na = RA @ dirs[ia]
        # pick the pair direction close to na
        w = rng.normal(size=3)
        w -= w.dot(na) * na
        w /= np.linalg.norm(w)
        ti = rng.uniform(0.05, 0.8) * thm
        d = math.cos(ti) * na + math.sin(ti) * w
        # rotate B so that its patch ib points back along -d within the cone
        nb = RB @ dirs[ib]
        tgt = -d
        v = np.cross(nb, tgt)
        c = nb.dot(tgt)
        vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        RB = (np.eye(3) + vx + vx @ vx / (1 + c)) @ RB
        tj = rng.uniform(0.05, 0.8) * thm
        RB = rot(np.cross(tgt, rng.normal(size=3)), tj) @ RB
        r = rng.uniform(0.95, 1.5)
        return r * d, RA.T.ravel(), RB.T.ravel(), ia, ib


def pair_energy(xa, xb, QA, QB, ia, ib, dirs, eps, dd, thm):
    un = xa - xb
    r = np.linalg.norm(un)
    return po.kf2_energy(un / r, r, QA, QB, dirs[ia], dirs[ib], eps, dd, thm)


@pytest.mark.parametrize("dirs", [wl.TETRA_DIRS, wl.PLANAR3_DIRS, wl.BIVALENT_DIRS])
def test_patch_force_torque_finite_differences(dirs):
    """force = -dU/dx_A, torque_A = -dU/dphi_A, torque_B = -dU/dphi_B (lab
    frame, rotation of the body about a lab axis), Newton's third law."""
    rng = np.random.default_rng(5)
    eps, dd, thm = 6.0, 1.2, wl.THETAM_095
    checked = 0
    for _ in range(40):
        xb, QA, QB, ia, ib = random_patch_pair(rng, dirs, thm)
        xa = np.zeros(3)
        un = xa - xb
        r = np.linalg.norm(un)
        out = po.kf2_force_torque(un / r, r, QA, QB, dirs[ia], dirs[ib], eps, dd, thm)
        if np.all(out == 0):
            continue
        checked += 1
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = -(pair_energy(xa + e, xb, QA, QB, ia, ib, dirs, eps, dd, thm) -
                   pair_energy(xa - e, xb, QA, QB, ia, ib, dirs, eps, dd, thm)) / (2 * h)
            assert abs(out[k] - fd) < 2e-5 * max(1.0, np.abs(out[:3]).max())
            # torque on A: body->lab map R_A^T... rotate A's body by a lab rotation
            Rp, Rm = rot(np.eye(3)[k], h), rot(np.eye(3)[k], -h)
            QAp = (QA.reshape(3, 3) @ Rp.T).ravel()
            QAm = (QA.reshape(3, 3) @ Rm.T).ravel()
            fdt = -(pair_energy(xa, xb, QAp, QB, ia, ib, dirs, eps, dd, thm) -
                    pair_energy(xa, xb, QAm, QB, ia, ib, dirs, eps, dd, thm)) / (2 * h)
            assert abs(out[3 + k] - fdt) < 2e-5 * max(1.0, np.abs(out[3:]).max())
            QBp = (QB.reshape(3, 3) @ Rp.T).ravel()
            QBm = (QB.reshape(3, 3) @ Rm.T).ravel()
            fdt = -(pair_energy(xa, xb, QA, QBp, ia, ib, dirs, eps, dd, thm) -
                    pair_energy(xa, xb, QA, QBm, ia, ib, dirs, eps, dd, thm)) / (2 * h)
            assert abs(out[6 + k] - fdt) < 2e-5 * max(1.0, np.abs(out[3:]).max())
        # angular momentum balance: r x F + tau_A + tau_B = 0 (r = xa - xb)
        assert np.allclose(np.cross(un, out[:3]) + out[3:6] + out[6:9], 0, atol=1e-9)
    assert checked > 20


def test_patch_rotational_covariance():
    s = wl.patchy_tetra(400, phi=0.2, seed_pos=3)
    s["periodic"] = False
    s["box"] = np.repeat(1000.0, 3)
    s["x"] = s["x"] + 400.0
    F, T = po.compute_forces(s, list_cut=3.5, n_side=50)
    R = rot([0.3, -0.5, 0.8], 0.77)
    c = np.array([500.0, 500.0, 500.0])
    s2 = dict(s)
    s2["x"] = (s["x"] - c) @ R.T + c
    s2["Q"] = np.einsum("nij,kj->nik", s["Q"].reshape(-1, 3, 3), R).reshape(-1, 9)
    F2, T2 = po.compute_forces(s2, list_cut=3.5, n_side=50)
    assert np.abs(T).max() > 1.0
    assert np.allclose(F2, F @ R.T, atol=1e-9 * np.abs(F).max())
    assert np.allclose(T2, T @ R.T, atol=1e-9 * np.abs(T).max())
    # Newton's third law summed over the system
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()


def test_bond_and_angle_finite_differences():
    rng = np.random.default_rng(2)
    x = np.array([[1.0, 1.0, 1.0], [1.9, 1.3, 1.1], [2.5, 2.1, 1.4], [3.1, 2.4, 2.3]])
    x = x + rng.normal(scale=0.05, size=x.shape)
    L = 50.0
    bonds = np.array([[0, 1], [1, 2], [2, 3]])
    ang = np.array([[0, 1, 2], [1, 2, 3]])

    def e_bond(xx, kind):
        d = np.linalg.norm(xx[bonds[:, 0]] - xx[bonds[:, 1]], axis=1)
        if kind == 1:
            return np.sum(0.5 * 100.0 * (d - 1.0) ** 2)
        return np.sum(-(30.0 / 2) * 1.5 ** 2 * np.log(1 - d ** 2 / 1.5 ** 2))

    def e_angle(xx, form, th0):
        ab = xx[ang[:, 1]] - xx[ang[:, 0]]
        bc = xx[ang[:, 2]] - xx[ang[:, 1]]
        c = np.sum(ab * bc, 1) / np.linalg.norm(ab, axis=1) / np.linalg.norm(bc, axis=1)
        if form == 0:
            return np.sum(0.5 * 50.0 * (np.arccos(c) - th0) ** 2)
        return np.sum(20.0 * (1 - c))

    def fd(fun):
        g = np.zeros_like(x)
        h = 1e-6
        for i in range(len(x)):
            for k in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, k] += h
                xm[i, k] -= h
                g[i, k] = -(fun(xp) - fun(xm)) / (2 * h)
        return g

    for kind, p0, p1 in ((1, 100.0, 1.0), (2, 30.0, 1.5)):
        F = po.pair_forces(L, x, bonds, kind, p0, p1)
        assert np.allclose(F, fd(lambda xx: e_bond(xx, kind)), atol=1e-5)
    for form, th0 in ((0, 0.4), (0, 0.0), (1, 0.0)):
        F = po.angle_forces(L, x, ang, form, np.full(2, 50.0 if form == 0 else 20.0),
                            np.full(2, th0))
        assert np.allclose(F, fd(lambda xx: e_angle(xx, form, th0)), atol=1e-5)
        assert np.abs(F.sum(0)).max() < 1e-10


def test_active_force_closed_form():
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0], [1.0, 1.0, 1.0]]) + 5
    prev = np.array([-1, 0, 1, 2], np.int32)
    nxt = np.array([1, 2, 3, -1], np.int32)
    F = po.active_forces(20.0, x, prev, nxt, 2.0)
    t = np.array([[1, 0, 0], [1, 1, 0] / np.sqrt(2), [0, 1, 1] / np.sqrt(2), [0, 0, 1]])
    assert np.allclose(F, 2.0 * t)


def test_rotation_update_invariants():
    s = wl.patchy_tetra(64)
    rng = np.random.default_rng(1)
    L = rng.normal(scale=3.0, size=(64, 3))
    Q = s["Q"]
    for i in range(0, 64, 7):
        R = po.genfullmat(s, L[i]).reshape(3, 3)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-13)
        assert abs(np.linalg.det(R) - 1) < 1e-13
    L2, Q2 = po.rotate(s, L, Q)
    assert np.allclose(np.linalg.norm(L2, axis=1), np.linalg.norm(L, axis=1), rtol=1e-13)
    for q in Q2.reshape(-1, 3, 3):
        assert np.allclose(q @ q.T, np.eye(3), atol=1e-12)
    # small-angle limit: R = exp([phi]x) to O(phi^2) with phi = dt L / I; the
    # orientation then moves as Q <- R^T Q, i.e. dQ/dt = -[omega]x Q (body frame)
    phi = np.array([1e-4, -2e-4, 3e-4])
    R = po.genfullmat(s, phi * s["I"][0] / s["dt"]).reshape(3, 3)
    K = np.array([[0, -phi[2], phi[1]], [phi[2], 0, -phi[0]], [-phi[1], phi[0], 0]])
    assert np.allclose(R, np.eye(3) + K, atol=1e-7)


def test_langevin_updates_are_affine():
    s = wl.patchy_tetra(32)
    rng = np.random.default_rng(3)
    p = rng.normal(size=(32, 3))
    F1, F2 = rng.normal(size=(2, 32, 3))
    R1, R2 = rng.uniform(-1.7, 1.7, size=(2, 32, 3))
    a = po.advance_mom(s, p, F1, R1)
    b = po.advance_mom(s, p, F2, R2)
    c = po.advance_mom(s, p, 0.5 * (F1 + F2), 0.5 * (R1 + R2))
    assert np.allclose(c, 0.5 * (a + b), atol=1e-13)
    d = s["gamma"] * s["dt"] / 2
    c2 = 1 / (1 + d)
    assert np.allclose(po.advance_mom(s, p, 0 * F1, 0 * R1), (1 - d) * c2 * p, rtol=1e-14)
    # sphere: noise and friction enter linearly, torque is rotated to the body frame
    Q = s["Q"]
    L = rng.normal(size=(32, 3))
    T = rng.normal(size=(32, 3))
    Ff, Tf = po.sphere(s, 0 * p, 0 * L, Q, F1, T, np.zeros((32, 6)))
    assert np.allclose(Ff, F1, atol=1e-14)
    assert np.allclose(Tf, np.einsum("nij,nj->ni", Q.reshape(-1, 3, 3), T), atol=1e-13)
    RT = rng.uniform(-1.7, 1.7, size=(32, 6))
    Ff, Tf = po.sphere(s, p, L, Q, 0 * F1, 0 * T, RT)
    Rt = math.sqrt(2 * s["gamma"] * s["kT"] / s["dt"])
    Rr = math.sqrt(2 * s["gammar"] * s["kT"] / s["dt"])
    Qm = Q.reshape(-1, 3, 3)
    fb = -s["gamma"] / s["m"] * np.einsum("nij,nj->ni", Qm, p) + Rt * RT[:, :3]
    assert np.allclose(Ff, np.einsum("nji,nj->ni", Qm, fb), atol=1e-12)
    assert np.allclose(Tf, -s["gammar"] / 0.1 * L + Rr * RT[:, 3:], atol=1e-12)


def test_wrap_and_reflect():
    s = wl.patchy_tetra(4)
    x = np.array([[-0.5, 3.0, 27.0], [28.5, -60.0, 1.0], [1.0, 2.0, 3.0], [0.0, 0.0, 0.0]])
    p = np.zeros((4, 3))
    L = float(s["box"][0])
    s2 = dict(s, dt=0.0)
    x2, _ = po.advance_pos(s2, x, p)
    assert np.all((x2 >= 0) & (x2 <= L))
    assert np.allclose(np.mod(x2 - x, L) * (np.abs(np.mod(x2 - x, L) - L) > 1e-9), 0, atol=1e-9)
    s3 = dict(s2, periodic=False)
    p = np.ones((4, 3))
    x3, p3 = po.advance_pos(s3, x, p)
    assert x3[0, 0] == 0.5 and p3[0, 0] == -1.0
    assert abs(x3[1, 0] - (L - (28.5 - L))) < 1e-12 and p3[1, 0] == -1.0


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert po.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert po.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E,
                                                               0xA20BC7C6, 0x6D5451FD]
    assert po.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                     [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_injected_noise_moments():
    z = po.noise(123, 0, 400000)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    assert z.min() >= -1.732050808 and z.max() <= 1.732050808 + 1e-9
    assert np.array_equal(po.noise(123, 1000, 10), z[1000:1010])


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def brute_pairs_ortho(x, box, cut, periodic=True):
    box = np.asarray(box, dtype=np.float64)
    d = x[:, None, :] - x[None, :, :]
    if periodic:
        d = np.where(np.abs(d) > box * 0.5, d - np.sign(d) * box, d)
    r2 = np.sum(d * d, axis=2)
    i, j = np.nonzero(np.triu(r2 < cut * cut, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("periodic", [True, False])
def test_orthorhombic_pair_set_brute_force(periodic):
    """The oracle's box is an orthorhombic generalisation of the reference's
    cube (per-axis edges and cell counts); its pair set equals an O(N^2)
    search with the per-axis minimum image, on a 1:1.5:3 box"""
    rng = np.random.default_rng(12)
    box = np.array([9.0, 13.5, 27.0])
    x = rng.uniform(0, 1, (900, 3)) * box
    for cut in (1.12246, 1.949, 3.5):
        n = np.maximum(np.floor(box / cut), 1).astype(int)
        pr = po.pairs_reference(box, x, n, cut, periodic)
        got = {(min(a, b), max(a, b)) for a, b in pr.tolist()}
        assert len(got) == len(pr)
        assert got == brute_pairs_ortho(x, box, cut, periodic)


def test_orthorhombic_wca_forces_brute_force():
    """WCA forces on an orthorhombic periodic box against a vectorised
    all-pairs evaluation of 24 eps (2 sigma^12 / r^14 - sigma^6 / r^8) d with
    the per-axis minimum image (the closed form of HSPotential::force)"""
    rng = np.random.default_rng(13)
    box = np.array([8.0, 12.0, 24.0])
    x = rng.uniform(0, 1, (700, 3)) * box
    s = {"N": len(x), "box": box, "periodic": True, "x": x, "Q": None,
         "wca": (1.0, 1.0, 1.12246), "patches": None}
    F, _ = po.compute_forces(s, list_cut=1.5)
    d = x[:, None, :] - x[None, :, :]
    d = np.where(np.abs(d) > box * 0.5, d - np.sign(d) * box, d)
    r2 = np.sum(d * d, axis=2)
    np.fill_diagonal(r2, np.inf)
    m = r2 < 1.12246 ** 2
    f = np.where(m, 24 * (2 / r2 ** 7 - 1 / r2 ** 4), 0.0)
    Fb = np.sum(f[:, :, None] * d, axis=1)
    assert np.count_nonzero(m) > 20
    assert relmax(F, Fb) < 1e-12


def _open_system(builder):
    """a system moved into a box with reflecting walls and a margin, so an
    affine deformation of the positions deforms every pair vector exactly"""
    s = builder()
    s = dict(s)
    s["periodic"] = False
    x = np.array(s["x"], dtype=np.float64)
    box = np.asarray(s["box"], dtype=np.float64)
    # unwrap chains first: bead chains of the generator may cross the faces
    if s.get("bonds") is not None:
        for a, b in np.asarray(s["bonds"]):
            dd = x[b] - x[a]
            x[b] = x[a] + dd - box * np.round(dd / box)
    x = x - x.min(0) + 5.0
    s["box"] = np.full(3, float(x.max() + 5.0))
    s["x"] = x
    return s


@pytest.mark.parametrize("builder", [
    lambda: wl.patchy_tetra(400, phi=0.2),
    lambda: wl.condensate_ab(600, phi=0.2),
    lambda: wl.active_polymer(nchains=6, nbeads=30, phi=0.1, seed=3, fa=0.0),
    lambda: wl.mt_bundle(1, 13, 12, spacing=6.0, box=(8.0, 8.0, 14.0)),
])
def test_energy_virial_is_minus_strain_derivative(builder):
    """orc_energy_virial's W[k, l] = sum_i F_i,k x_i,l equals
    -dU/d eps of the energy under the affine deformation x -> (1 + eps e_k e_l^T) x
    (central differences), for every k, l: pins the virial of the pair, patch
    (non-central forces: W is not symmetric), bond and three-body terms on
    the energy, which the reference fixtures pin"""
    s = _open_system(builder)
    U, W = po.energy_virial(s)
    assert W.any()
    h = 1e-6
    for k in range(3):
        for l in range(3):
            E = np.zeros((3, 3))
            E[k, l] = 1.0
            xp = s["x"] + h * s["x"] @ E.T
            xm = s["x"] - h * s["x"] @ E.T
            Up, _ = po.energy_virial(s, x=xp)
            Um, _ = po.energy_virial(s, x=xm)
            fd = -(Up - Um) / (2 * h)
            assert abs(fd - W[k, l]) < 2e-5 * max(1.0, np.abs(W).max()), (k, l, fd, W[k, l])


def test_energy_parts_pinned():
    """U of orc_energy_virial = the WCA energies (fixture-pinned
    orc_wca_energy) + the patch energy (fixture-pinned orc_patch_energy) on
    the same list; for bonds and angles, -dU/dx equals the fixture-pinned
    forces and U vanishes at the rest length / rest angle"""
    s = wl.patchy_tetra(1000, phi=0.2)
    U, _ = po.energy_virial(s, list_cut=3.5)
    L = float(s["box"][0])
    pr = po.pairs_reference(L, s["x"], int(L // 4), 3.5)
    d = s["x"][pr[:, 0]] - s["x"][pr[:, 1]]
    d = np.where(np.abs(d) > L / 2, d - np.sign(d) * L, d)
    r = np.sqrt(np.sum(d * d, 1))
    ew = sum(po.lib().orc_wca_energy(float(v), 1.0, 1.0, 0) for v in r)
    assert abs(U - (ew + po.patch_energy(s))) < 1e-12 * max(1.0, abs(U))
    # bonds and angles: energy derivative = force
    c = _open_system(lambda: wl.active_polymer(nchains=4, nbeads=12, phi=0.1, seed=5, fa=0.0))
    c["wca"] = None
    c["patches"] = None
    for form in (0, 1):
        c["angle_form"] = form
        c["angle_theta0"] = np.full(len(c["angles"]), 0.4 if form == 0 else 0.0)
        U0, _ = po.energy_virial(c)
        F, _ = po.compute_forces(c)
        i = 7
        for k in range(3):
            xp, xm = c["x"].copy(), c["x"].copy()
            xp[i, k] += 1e-6
            xm[i, k] -= 1e-6
            fd = -(po.energy_virial(c, x=xp)[0] - po.energy_virial(c, x=xm)[0]) / 2e-6
            assert abs(fd - F[i, k]) < 1e-5 * max(1.0, np.abs(F).max())
    rest = {"N": 3, "box": np.full(3, 50.0), "periodic": False,
            "x": np.array([[10.0, 10, 10], [11.0, 10, 10], [11.0 + np.cos(0.4), 10 + np.sin(0.4), 10]]),
            "Q": None, "wca": None, "patches": None, "bonds": np.array([[0, 1], [1, 2]]),
            "bond_k": np.array([100.0, 100.0]), "bond_r0": np.array([1.0, 1.0]), "bond_kind": 1,
            "angles": np.array([[0, 1, 2]]), "angle_k": np.array([50.0]),
            "angle_theta0": np.array([0.4]), "angle_form": 0}
    assert abs(po.energy_virial(rest)[0]) < 1e-20
