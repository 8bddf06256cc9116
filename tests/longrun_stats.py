"""Statistics of long patchy runs (test infrastructure): radial distribution
function, patch-bond counts and kinetic temperatures from snapshots, with
block-average error bars.  A patch bond is a particle pair with
r < 1.4 d and one patch pair passing both cone tests of
KernFrenkelOnePatch2 (potentialtheta.h:517-750): -n_i.u > cos(theta_m) and
n_j.u > cos(theta_m), u = (x_i - x_j)/r, n = Q^T n_body."""
import numpy as np
from scipy.spatial import cKDTree

RMAX, DR = 4.0, 0.05
EDGES = np.arange(0.0, RMAX + DR / 2, DR)


def gr(x, L):
    N = len(x)
    t = cKDTree(np.mod(x, L), boxsize=L)
    pr = t.query_pairs(RMAX, output_type="ndarray")
    d = x[pr[:, 0]] - x[pr[:, 1]]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(1))
    h, _ = np.histogram(r, EDGES)
    rho = N / L ** 3
    shell = 4.0 / 3.0 * np.pi * (EDGES[1:] ** 3 - EDGES[:-1] ** 3)
    return 2.0 * h / (N * rho * shell)


def bonds(x, Q, L, dirs, d=1.2, costm=0.95, rmax=None):
    """bonds per particle; rmax: also require r < rmax (the force loop's
    max_distance_check, LangevinR.cpp:938)"""
    N = len(x)
    t = cKDTree(np.mod(x, L), boxsize=L)
    pr = t.query_pairs(1.4 * d if rmax is None else min(1.4 * d, rmax), output_type="ndarray")
    i, j = pr[:, 0], pr[:, 1]
    dv = x[i] - x[j]
    dv -= L * np.round(dv / L)
    u = dv / np.linalg.norm(dv, axis=1, keepdims=True)
    Qm = Q.reshape(N, 3, 3)
    # lab normals n = Q^T n_body for every patch: (N, P, 3)
    n = np.einsum("nkl,pk->npl", Qm, dirs)
    ai = -np.einsum("mpl,ml->mp", n[i], u)
    aj = np.einsum("mpl,ml->mp", n[j], u)
    ok = ((ai > costm)[:, :, None] & (aj > costm)[:, None, :]).any((1, 2))
    cnt = np.zeros(N, dtype=np.int64)
    np.add.at(cnt, i[ok], 1)
    np.add.at(cnt, j[ok], 1)
    return cnt


def frame_stats(x, p, Q, Lm, L, dirs, m=1.0, I=0.1):
    b = bonds(x, Q, L, dirs)
    hist = np.bincount(np.minimum(b, 4), minlength=5) / len(b)
    return {
        "gr": gr(x, L),
        "bond_mean": np.array([b.mean()]),
        "bond_hist": hist,
        "ke_t": (p * p).mean(0) / (2 * m),
        "ke_r": (Lm * Lm).mean(0) / (2 * I),
    }


def series(frames, L, dirs):
    """frames: iterable of (x, p, Q, Lm); returns {key: (nframes, ...)}"""
    out = {}
    for (x, p, Q, Lm) in frames:
        for k, v in frame_stats(x, p, Q, Lm, L, dirs).items():
            out.setdefault(k, []).append(v)
    return {k: np.array(v) for k, v in out.items()}


def mean_err(s, nblocks=8):
    """block means and the standard error of their mean"""
    nb = nblocks
    n = (len(s) // nb) * nb
    blocks = s[:n].reshape(nb, n // nb, *s.shape[1:]).mean(1)
    return blocks.mean(0), blocks.std(0, ddof=1) / np.sqrt(nb)
