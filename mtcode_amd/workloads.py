"""Seeded synthetic inputs for the five Langevin workloads (host-side, numpy).

Every builder returns a plain ``dict`` describing one system (or a list of
them for the batched-replica workload).  Arrays follow the reference's host
layout: positions, momenta, torques N x 3 row-major FP64, orientations N x 9
row-major (lab -> body, ``DataStructures/matrix2.h:33-36``,
``MDBase/Rotational/LangevinR.cpp:1-2``).

Workloads (sizes as named by the project's benchmark list):

* ``patchy_tetra``   N=4096 patchy fluid, packing fraction 0.1, 4 tetrahedral
                     patches, WCA + patches eps_p=6, d=1.2, cos(theta_m)=0.95.
* ``mt_bundle``      64 tubes x 13 protofilaments x 500 beads on a 13_3
                     helical lattice, harmonic bonds k=100 r0=1, cosine
                     bending kappa=50, WCA between non-bonded beads.
* ``condensate_ab``  N=1,048,576 two-species patchy condensate, packing
                     fraction 0.05, A: 3 planar patches, B: 2 patches, A-B
                     attraction only (eps_p=8) plus WCA.
* ``replicas``       256 independent tetrahedral systems, N log-uniform in
                     [512, 16384], kT in [0.8, 1.2], eps_p in [4, 10].
* ``active_polymer`` 8192 chains x 256 beads, harmonic k=200, cosine
                     bending kappa=20, tangential active force f_a=1, chain
                     ends carry 2 patches (eps_p=5).
"""
from __future__ import annotations

import math

import numpy as np

# Patch geometries (body frame unit vectors).
# TetrahedralPatch (MDBase/Potentials/ComboPatch/TetrahedralPatch.h:7-21)
TETRA_DIRS = np.array(
    [
        [math.sqrt(8.0 / 9.0), 0.0, -1.0 / 3.0],
        [-math.sqrt(2.0 / 9.0), math.sqrt(2.0 / 3.0), -1.0 / 3.0],
        [-math.sqrt(2.0 / 9.0), -math.sqrt(2.0 / 3.0), -1.0 / 3.0],
        [0.0, 0.0, 1.0],
    ]
)
# three planar arms (mainPatchy.cpp:152-174)
PLANAR3_DIRS = np.array(
    [[1.0, 0.0, 0.0], [-0.5, math.sqrt(3.0 / 4.0), 0.0], [-0.5, -math.sqrt(3.0 / 4.0), 0.0]]
)
# BivalentPatch (MDBase/Potentials/ComboPatch/BivalentPatch.h:6-12)
BIVALENT_DIRS = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])

WCA_CUT_LITERAL = 1.12246  # HSPotential paramlimit (Potentials/potential.h:84)
THETAM_095 = math.acos(0.95)


def box_for_packing(n: int, phi: float, sigma: float = 1.0) -> float:
    """Cubic box edge for n spheres of diameter sigma at packing fraction phi."""
    return (n * math.pi * sigma**3 / 6.0 / phi) ** (1.0 / 3.0)


def nonoverlapping_uniform(n: int, box, rng: np.random.Generator, rmin: float = 1.0,
                           periodic: bool = True) -> np.ndarray:
    """Uniform positions in [0, L) with every pair farther apart than rmin.

    Overlapping particles are redrawn until none remain (random sequential
    rejection; deterministic for a given generator state)."""
    from scipy.spatial import cKDTree

    box = np.asarray(box, dtype=np.float64).reshape(-1)
    if box.size == 1:
        box = np.repeat(box, 3)
    x = rng.uniform(0.0, 1.0, size=(n, 3)) * box
    x = np.minimum(x, np.nextafter(box, 0))
    if n > 200000:
        # large systems: after the first full pass only redrawn particles are
        # re-checked (a different, equally valid, deterministic sequence)
        kw = dict(boxsize=box if periodic else None, balanced_tree=False, compact_nodes=False)
        tree = cKDTree(x, **kw)
        pairs = tree.query_pairs(rmin, output_type="ndarray")
        bad = np.unique(pairs[:, 1]) if len(pairs) else np.zeros(0, dtype=int)
        for _ in range(1000):
            if len(bad) == 0:
                return x
            x[bad] = rng.uniform(0.0, 1.0, size=(len(bad), 3)) * box
            x[bad] = np.minimum(x[bad], np.nextafter(box, 0))
            tree = cKDTree(x, **kw)
            cnt = tree.query_ball_point(x[bad], rmin, return_length=True, workers=8)
            bad = bad[cnt > 1]
        raise RuntimeError("could not place particles without overlap")
    for _ in range(1000):
        tree = cKDTree(x, boxsize=box if periodic else None)
        pairs = tree.query_pairs(rmin, output_type="ndarray")
        if len(pairs) == 0:
            return x
        bad = np.unique(pairs[:, 1])
        x[bad] = rng.uniform(0.0, 1.0, size=(len(bad), 3)) * box
        x[bad] = np.minimum(x[bad], np.nextafter(box, 0))
    raise RuntimeError("could not place particles without overlap")


def arvo_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation matrices, N x 9 row-major, by Arvo's method in
    the element order of LangevinNVTR::initialize (LangevinR.cpp:237-267)."""
    x1, x2, x3 = rng.uniform(0.0, 1.0, size=(3, n))
    v1 = np.cos(2 * math.pi * x2) * np.sqrt(x3)
    v2 = np.sin(2 * math.pi * x2) * np.sqrt(x3)
    v3 = np.sqrt(1.0 - x3)
    r1 = np.cos(2 * math.pi * x1)
    r2 = np.sin(2 * math.pi * x1)
    q = np.empty((n, 9))
    q[:, 0] = -(r1 * (1 - 2 * v1**2)) - 2 * r2 * v1 * v2
    q[:, 1] = -(r2 * (1 - 2 * v1**2)) + 2 * r1 * v1 * v2
    q[:, 2] = 2 * v1 * v3
    q[:, 3] = 2 * r1 * v1 * v2 + r2 * (1 - 2 * v2**2)
    q[:, 4] = 2 * r2 * v1 * v2 - r1 * (1 - 2 * v2**2)
    q[:, 5] = 2 * v2 * v3
    q[:, 6] = 2 * r1 * v1 * v3 - 2 * r2 * v2 * v3
    q[:, 7] = 2 * r2 * v1 * v3 + 2 * r1 * v2 * v3
    q[:, 8] = -1 + 2 * v3**2
    return q


def tetra_patchset(eps_p: float, d: float = 1.2, thetam: float = THETAM_095) -> dict:
    """TetrahedralPatch(eps_p, d, thetam): one type, 4 patches, 16 equal rows
    (TetrahedralPatch.cpp:35-47)."""
    return {
        "npatch": [4],
        "dirs": TETRA_DIRS.copy(),
        "params": np.tile([eps_p, d, thetam], (16, 1)).astype(np.float64),
    }


def two_species_patchset(eps_ab: float, d: float = 1.2, thetam: float = THETAM_095) -> dict:
    """GeneralPatch with type 0 = 3 planar patches, type 1 = 2 bivalent
    patches, rows in GeneralPatch order (GeneralPatch.cpp:75-106): (0,0) 9 rows
    eps=0, (0,1) 6 rows eps=eps_ab, (1,1) 4 rows eps=0."""
    rows = []
    for (ti, tj) in [(0, 0), (0, 1), (1, 1)]:
        npi = 3 if ti == 0 else 2
        npj = 3 if tj == 0 else 2
        e = eps_ab if ti != tj else 0.0
        rows += [[e, d, thetam]] * (npi * npj)
    return {
        "npatch": [3, 2],
        "dirs": np.vstack([PLANAR3_DIRS, BIVALENT_DIRS]),
        "params": np.array(rows, dtype=np.float64),
    }


def _base(name: str, n: int, box, dt: float, kT: float, gamma: float, gammar: float,
          integrator: str) -> dict:
    box = np.asarray(box, dtype=np.float64).reshape(-1)
    if box.size == 1:
        box = np.repeat(box, 3)
    return {
        "name": name,
        "N": int(n),
        "box": box,
        "periodic": True,
        "dt": float(dt),
        "m": 1.0,
        "gamma": float(gamma),
        "gammar": float(gammar),
        "kT": float(kT),
        "I": np.array([0.1, 0.1, 0.1]),
        "integrator": integrator,
        "wca": (1.0, 1.0, WCA_CUT_LITERAL),
        "patches": None,
        "bonds": None,
        "angles": None,
        "exclusions": None,
        "active": None,
        "p": np.zeros((n, 3)),
        "L": np.zeros((n, 3)),
    }


def mixed_quaternions(Q: np.ndarray) -> np.ndarray:
    """The unit quaternions (N x 4 float32, w x y z) the mixed-precision
    engine stores for the orientations Q (N x 9 lab -> body matrices): its
    upload's Shepperd's method in FP64, normalised, rounded once to FP32,
    with the same operations."""
    m = np.asarray(Q, dtype=np.float64).reshape(-1, 9).T
    tr = (m[0] + m[4]) + m[8]
    c0 = tr > 0
    c1 = ~c0 & (m[0] > m[4]) & (m[0] > m[8])
    c2 = ~c0 & ~c1 & (m[4] > m[8])
    q = np.zeros((4, m.shape[1]))
    with np.errstate(invalid="ignore", divide="ignore"):
        S0 = np.sqrt(tr + 1) * 2
        S1 = np.sqrt(((1 + m[0]) - m[4]) - m[8]) * 2
        S2 = np.sqrt(((1 + m[4]) - m[0]) - m[8]) * 2
        S3 = np.sqrt(((1 + m[8]) - m[0]) - m[4]) * 2
        br = [(c0, [0.25 * S0, (m[7] - m[5]) / S0, (m[2] - m[6]) / S0, (m[3] - m[1]) / S0]),
              (c1, [(m[7] - m[5]) / S1, 0.25 * S1, (m[1] + m[3]) / S1, (m[2] + m[6]) / S1]),
              (c2, [(m[2] - m[6]) / S2, (m[1] + m[3]) / S2, 0.25 * S2, (m[5] + m[7]) / S2]),
              (~(c0 | c1 | c2), [(m[3] - m[1]) / S3, (m[2] + m[6]) / S3, (m[5] + m[7]) / S3,
                                 0.25 * S3])]
    for c, comp in br:
        for k in range(4):
            q[k][c] = comp[k][c]
    inv = 1 / np.sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3])
    return (q * inv).astype(np.float32).T


def mixed_orientations(Q: np.ndarray) -> np.ndarray:
    """The matrices (N x 9, FP64) that the mixed-precision engine's kernels
    use for the orientations Q it is given: the quaternions of
    mixed_quaternions, turned back into matrices in FP32 arithmetic with the
    kernels' operations.  An FP64 engine handed them sees exactly the
    orientations a mixed engine handed Q works with."""
    w, x, y, z = mixed_quaternions(Q).T
    one, two = np.float32(1), np.float32(2)
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz, wx, wy, wz = x * y, x * z, y * z, w * x, w * y, w * z
    out = np.stack([one - two * (yy + zz), two * (xy - wz), two * (xz + wy),
                    two * (xy + wz), one - two * (xx + zz), two * (yz - wx),
                    two * (xz - wy), two * (yz + wx), one - two * (xx + yy)], 1)
    return out.astype(np.float64)


def thermal_momenta(s: dict, seed: int = 99) -> dict:
    """A copy of system s with momenta drawn from the Maxwell-Boltzmann
    distribution at its kT: p ~ N(0, m kT) per component, and for rotors the
    body-frame angular momenta L ~ N(0, I_k kT).  The generators start from
    rest; a run from these momenta is at its steady temperature from the first
    step (the friction would otherwise take about 1 / gamma to heat it)."""
    t = dict(s)
    rng = np.random.default_rng(seed)
    n = s["N"]
    t["p"] = rng.normal(0.0, np.sqrt(s["m"] * s["kT"]), (n, 3))
    if s["integrator"] == "rotor":
        t["L"] = rng.normal(0.0, 1.0, (n, 3)) * np.sqrt(np.asarray(s["I"]) * s["kT"])
    return t


def patchy_tetra(n: int = 4096, phi: float = 0.1, seed_pos: int = 1, seed_q: int = 2,
                 eps_p: float = 6.0, kT: float = 1.0) -> dict:
    """N patchy particles, 4 tetrahedral patches, cubic periodic box at packing
    fraction phi; kT=1, gamma=1, gamma_r=1/3, I=0.1, m=1, dt=0.005."""
    L = box_for_packing(n, phi)
    s = _base("patchy_tetra", n, L, 0.005, kT, 1.0, 1.0 / 3.0, "rotor")
    s["x"] = nonoverlapping_uniform(n, L, np.random.default_rng(seed_pos))
    s["Q"] = arvo_rotations(n, np.random.default_rng(seed_q))
    ps = tetra_patchset(eps_p)
    ps["types"] = np.zeros(n, dtype=np.int32)
    ps["maxd2"] = 2.4
    s["patches"] = ps
    return s


def condensate_ab(n: int = 1 << 20, phi: float = 0.05, seed_pos: int = 3, seed_q: int = 4,
                  eps_ab: float = 8.0) -> dict:
    """Two-species condensate: indices [0, n/2) type A (3 planar patches),
    [n/2, n) type B (2 patches), A-B attraction only, WCA for all."""
    L = box_for_packing(n, phi)
    s = _base("condensate_ab", n, L, 0.005, 1.0, 1.0, 1.0 / 3.0, "rotor")
    s["x"] = nonoverlapping_uniform(n, L, np.random.default_rng(seed_pos))
    s["Q"] = arvo_rotations(n, np.random.default_rng(seed_q))
    ps = two_species_patchset(eps_ab)
    t = np.zeros(n, dtype=np.int32)
    t[n // 2:] = 1
    ps["types"] = t
    ps["maxd2"] = 2.4
    s["patches"] = ps
    return s


def condensate_bind(n: int = 1 << 20, phi: float = 0.05) -> dict:
    """The two-species condensate under the one-bond-per-patch binding model
    of the reference's patchy driver: BindingModelSingle(0.99999, 1 - 0.99999)
    (mainPatchy.cpp:97-101, run_singlebond); all patches start unbound."""
    s = condensate_ab(n, phi)
    s["name"] = "condensate_bind"
    s["binding"] = (0.99999, 1.0 - 0.99999)
    return s


def replicas(count: int = 256, seed: int = 5, nmin: int = 512, nmax: int = 16384,
             phi: float = 0.1, scale: float = 1.0, progress=None) -> list:
    """Independent tetrahedral systems with N log-uniform in [nmin, nmax],
    kT ~ U[0.8, 1.2], eps_p ~ U[4, 10], each in its own box at packing
    fraction phi.  ``scale`` shrinks every N (for tests)."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=count))).astype(int)
    ns = np.maximum((ns * scale).astype(int), 8)
    kts = rng.uniform(0.8, 1.2, size=count)
    eps = rng.uniform(4.0, 10.0, size=count)
    out = []
    for r in range(count):
        s = patchy_tetra(int(ns[r]), phi, seed_pos=1000 + r, seed_q=5000 + r, eps_p=float(eps[r]),
                         kT=float(kts[r]))
        s["name"] = f"replica{r}"
        out.append(s)
        if progress and (r + 1) % 32 == 0:
            progress(f"replica {r + 1}/{count}")
    return out


def mt_bundle(ntubes_side: int = 8, nproto: int = 13, nbeads: int = 500, spacing: float = 25.0,
              box=(200.0, 200.0, 600.0), kbond: float = 100.0, kappa: float = 50.0,
              seed: int = 7) -> dict:
    """Microtubule-like bundle on an 8 x 8 xy grid of tubes.

    Bead (p, k) of a tube sits at radius R = sqrt(1 - (3/13)^2) / (2 sin(pi/13))
    and angle 2 pi p / 13 with z = k + 3 p / 13 (a 13_3 helical lattice: lateral
    and longitudinal neighbours one unit apart).  Bonds: (p,k)-(p,k+1) along
    protofilaments, (p,k)-(p+1,k) laterally, seam (12,k)-(0,k+3).  Angles:
    consecutive protofilament triplets.  Non-bonded pairs interact via WCA."""
    R = math.sqrt(1.0 - (3.0 / 13.0) ** 2) / (2.0 * math.sin(math.pi / nproto))
    nt = ntubes_side * ntubes_side
    per_tube = nproto * nbeads
    n = nt * per_tube
    box = np.asarray(box, dtype=np.float64)
    s = _base("mt_bundle", n, box, 0.002, 1.0, 1.0, 1.0 / 3.0, "bead")
    x = np.empty((n, 3))
    p_idx, k_idx = np.meshgrid(np.arange(nproto), np.arange(nbeads), indexing="ij")
    p_idx = p_idx.ravel()
    k_idx = k_idx.ravel()
    local = np.stack(
        [
            R * np.cos(2 * math.pi * p_idx / nproto),
            R * np.sin(2 * math.pi * p_idx / nproto),
            k_idx + 3.0 * p_idx / nproto,
        ],
        axis=1,
    )
    off = (box[:2] - spacing * (ntubes_side - 1)) / 2.0
    z0 = (box[2] - (nbeads + 3)) / 2.0
    bonds, angles = [], []
    lid = lambda pp, kk: pp * nbeads + kk  # noqa: E731
    for t in range(nt):
        cx = off[0] + spacing * (t % ntubes_side)
        cy = off[1] + spacing * (t // ntubes_side)
        base = t * per_tube
        x[base:base + per_tube] = local + np.array([cx, cy, z0])
        for pp in range(nproto):
            ks = np.arange(nbeads - 1)
            bonds.append(np.stack([base + lid(pp, ks), base + lid(pp, ks + 1)], 1))
            ks3 = np.arange(nbeads - 2)
            angles.append(np.stack([base + lid(pp, ks3), base + lid(pp, ks3 + 1),
                                    base + lid(pp, ks3 + 2)], 1))
            ks = np.arange(nbeads)
            if pp < nproto - 1:
                bonds.append(np.stack([base + lid(pp, ks), base + lid(pp + 1, ks)], 1))
            else:
                ks = np.arange(nbeads - 3)
                bonds.append(np.stack([base + lid(pp, ks), base + lid(0, ks + 3)], 1))
    s["x"] = np.mod(x, box)
    b = np.concatenate(bonds).astype(np.int32)
    a = np.concatenate(angles).astype(np.int32)
    s["bonds"] = b
    s["bond_k"] = np.full(len(b), kbond)
    s["bond_r0"] = np.full(len(b), 1.0)
    s["bond_kind"] = 1
    s["angles"] = a
    s["angle_k"] = np.full(len(a), kappa)
    s["angle_theta0"] = np.zeros(len(a))
    s["angle_form"] = 1
    s["exclusions"] = b.copy()
    s["Q"] = np.tile(np.eye(3).ravel(), (n, 1))
    _ = seed
    return s


def active_polymer(nchains: int = 8192, nbeads: int = 256, phi: float = 0.15, seed: int = 6,
                   kbond: float = 200.0, kappa: float = 20.0, fa: float = 1.0,
                   eps_p: float = 5.0, progress=None) -> dict:
    """Chains grown as off-lattice self-avoiding random walks (unit steps, no
    bead closer than 0.9 to an earlier bead of any chain), periodic box at
    volume fraction phi.  Harmonic bonds, cosine bending, tangential active
    force along the local bond; the two end beads of each chain are rotors
    with 2 patches (type 1), all other beads are type 0 without patches."""
    n = nchains * nbeads
    L = box_for_packing(n, phi)
    rng = np.random.default_rng(seed)
    s = _base("active_polymer", n, L, 0.005, 1.0, 1.0, 1.0 / 3.0, "rotor")
    x = _grow_chains(nchains, nbeads, L, rng, progress=progress)
    s["x"] = x
    idx = np.arange(n).reshape(nchains, nbeads)
    b = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1).astype(np.int32)
    a = np.stack([idx[:, :-2].ravel(), idx[:, 1:-1].ravel(), idx[:, 2:].ravel()], 1).astype(np.int32)
    s["bonds"] = b
    s["bond_k"] = np.full(len(b), kbond)
    s["bond_r0"] = np.full(len(b), 1.0)
    s["bond_kind"] = 1
    s["angles"] = a
    s["angle_k"] = np.full(len(a), kappa)
    s["angle_theta0"] = np.zeros(len(a))
    s["angle_form"] = 1
    s["exclusions"] = b.copy()
    prev = np.full(n, -1, dtype=np.int32)
    nxt = np.full(n, -1, dtype=np.int32)
    prev[idx[:, 1:].ravel()] = idx[:, :-1].ravel()
    nxt[idx[:, :-1].ravel()] = idx[:, 1:].ravel()
    s["active"] = {"fa": fa, "prev": prev, "next": nxt}
    s["Q"] = arvo_rotations(n, np.random.default_rng(seed + 100))
    types = np.zeros(n, dtype=np.int32)
    types[idx[:, 0]] = 1
    types[idx[:, -1]] = 1
    rows = [[0.0, 1.2, THETAM_095]] * 0  # type 0 has no patches
    rows = []
    # type pairs (0,0): 0 rows, (0,1): 0 rows, (1,1): 2 x 2 rows
    rows += [[eps_p, 1.2, THETAM_095]] * 4
    s["patches"] = {
        "npatch": [0, 2],
        "dirs": BIVALENT_DIRS.copy(),
        "params": np.array(rows, dtype=np.float64),
        "types": types,
        "maxd2": 2.4,
    }
    return s


class _BeadGrid:
    """Spatial hash of placed beads (cubic cells of edge >= rmin, K slots per
    cell) for the distance test of chain growth in a periodic cube."""

    K = 8  # points >= rmin apart in a cell of edge < 2 rmin: at most 8

    def __init__(self, L: float, rmin: float, cap: int):
        self.L, self.rmin = float(L), float(rmin)
        self.nc = max(3, min(int(L // rmin), 512))
        if L / self.nc >= 2 * rmin:
            self.nc = int(np.ceil(L / (1.999 * rmin)))
        self.slots = np.full((self.nc ** 3, self.K), -1, dtype=np.int32)
        self.fill = np.zeros(self.nc ** 3, dtype=np.int8)
        self.pts = np.zeros((cap, 3))
        self.n = 0

    def _cell3(self, p):
        return np.minimum((p / (self.L / self.nc)).astype(np.int64), self.nc - 1)

    def clear_of(self, p):
        """True where p (m x 3, in [0, L)) is at least rmin from every placed bead"""
        nc, L = self.nc, self.L
        c3 = self._cell3(p)
        off = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)])
        nb = (c3[:, None, :] + off[None]) % nc
        cid = (nb[..., 0] * nc + nb[..., 1]) * nc + nb[..., 2]
        ids = self.slots[cid].reshape(len(p), -1)
        valid = ids >= 0
        q = self.pts[np.where(valid, ids, 0)]
        d = p[:, None, :] - q
        d -= L * np.round(d / L)
        r2 = np.einsum("mkj,mkj->mk", d, d)
        return ~np.any(valid & (r2 < self.rmin ** 2), axis=1)

    def add(self, p):
        """insert points; returns their ids"""
        nc = self.nc
        c3 = self._cell3(p)
        cid = (c3[:, 0] * nc + c3[:, 1]) * nc + c3[:, 2]
        ids = np.arange(self.n, self.n + len(p), dtype=np.int32)
        self.pts[self.n:self.n + len(p)] = p
        self.n += len(p)
        order = np.argsort(cid, kind="stable")
        cs = cid[order]
        first = np.r_[0, np.flatnonzero(cs[1:] != cs[:-1]) + 1]
        rank = np.arange(len(cs)) - np.repeat(first, np.diff(np.r_[first, len(cs)]))
        slot = self.fill[cs].astype(np.int64) + rank
        if np.any(slot >= self.K):
            raise RuntimeError("bead grid cell overflow")
        self.slots[cs, slot] = ids[order]
        np.add.at(self.fill, cs, 1)
        return ids

    def remove(self, ids):
        nc = self.nc
        for i in np.asarray(ids).ravel():
            c3 = self._cell3(self.pts[i:i + 1])[0]
            c = (c3[0] * nc + c3[1]) * nc + c3[2]
            f = int(self.fill[c])
            k = int(np.flatnonzero(self.slots[c, :f] == i)[0])
            self.slots[c, k] = self.slots[c, f - 1]
            self.slots[c, f - 1] = -1
            self.fill[c] = f - 1


def _grow_chains(nchains: int, nbeads: int, L: float, rng: np.random.Generator,
                 rmin: float = 0.9, tries: int = 50, back: int = 4, progress=None) -> np.ndarray:
    """Off-lattice self-avoiding walks grown in parallel, one bead per chain
    per round: unit steps in random directions, a trial rejected when closer
    than rmin (minimum image) to any placed bead or to another accepted trial
    of the same round.  A chain that fails `tries` times in a row drops its
    last `back` beads and grows again from there."""
    from scipy.spatial import cKDTree

    n = nchains * nbeads
    x = np.zeros((nchains, nbeads, 3))
    gid = np.full((nchains, nbeads), -1, dtype=np.int64)
    grid = _BeadGrid(L, rmin, n + 64 * nchains * back)
    x[:, 0] = nonoverlapping_uniform(nchains, L, rng, rmin=2.0)
    gid[:, 0] = grid.add(x[:, 0])
    length = np.ones(nchains, dtype=np.int64)
    fails = np.zeros(nchains, dtype=np.int64)
    rounds = 0
    while True:
        act = np.flatnonzero(length < nbeads)
        if len(act) == 0:
            break
        rounds += 1
        if progress and rounds % 64 == 0:
            progress(f"chain growth round {rounds}, {int(length.sum())}/{n} beads")
        if rounds > 50 * nbeads + 1000:
            raise RuntimeError("chain growth did not finish")
        v = rng.normal(size=(len(act), 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        cand = np.mod(x[act, length[act] - 1] + v, L)
        ok = grid.clear_of(cand)
        idx = np.flatnonzero(ok)
        if len(idx) > 1:
            pairs = cKDTree(cand[idx], boxsize=L).query_pairs(rmin, output_type="ndarray")
            if len(pairs):
                ok[idx[np.unique(pairs[:, 1])]] = False
        acc = act[ok]
        x[acc, length[acc]] = cand[ok]
        gid[acc, length[acc]] = grid.add(cand[ok])
        length[acc] += 1
        fails[acc] = 0
        rej = act[~ok]
        fails[rej] += 1
        for c in rej[fails[rej] >= tries]:
            nb = int(min(back, length[c] - 1))
            grid.remove(gid[c, length[c] - nb:length[c]])
            gid[c, length[c] - nb:length[c]] = -1
            length[c] -= nb
            fails[c] = 0
    if grid.n > len(grid.pts):
        raise RuntimeError("bead grid capacity exceeded")
    return np.mod(x.reshape(n, 3), L)
