"""ctypes front end of liboracle.so (the FP64 C restatement in oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "mtref")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
MAXT = 8


class Box(C.Structure):
    _fields_ = [("l", C.c_double * 3), ("periodic", C.c_int)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("m", C.c_double), ("gamma", C.c_double),
                ("gammar", C.c_double), ("kT", C.c_double), ("I", C.c_double * 3)]


class PatchSet(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("type", _ip), ("npatch", C.c_int * MAXT),
                ("dirs", _dp), ("params", _dp)]


class BindModel(C.Structure):
    _fields_ = [("on_rate", C.c_double), ("off_rate", C.c_double), ("rng", C.c_int),
                ("seed", C.c_uint), ("key", C.c_uint32 * 2), ("draw", C.c_uint64),
                ("table", C.c_int), ("species", _ip), ("dtab", _dp), ("ttab", _dp),
                ("tr4", _dp)]


class Interactions(C.Structure):
    _fields_ = [
        ("wca_on", C.c_int), ("wca_eps", C.c_double), ("wca_sigma", C.c_double),
        ("wca_rcut", C.c_double),
        ("patches", C.POINTER(PatchSet)), ("maxd2", C.c_double),
        ("bonds", _ip), ("nbonds", C.c_longlong), ("bond_kind", C.c_int),
        ("bond_k", _dp), ("bond_r0", _dp),
        ("angles", _ip), ("nangles", C.c_longlong), ("angle_form", C.c_int),
        ("angle_k", _dp), ("angle_theta0", _dp),
        ("excl", _ip), ("nexcl", C.c_longlong),
        ("fa", C.c_double), ("prev", _ip), ("next", _ip),
        ("list_cut", C.c_double), ("rebuild_every", C.c_int), ("cells_per_side", C.c_int * 3),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_fasterfloor.restype = C.c_int
        L.orc_fasterfloor.argtypes = [C.c_double]
        L.orc_calcg1.argtypes = [C.POINTER(Box), C.c_int, _dp, C.c_double, C.c_double,
                                 C.POINTER(C.c_longlong), C.c_int]
        L.orc_assign_all.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, _ip]
        L.orc_dgraph_labels.restype = C.c_longlong
        L.orc_dgraph_labels.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, C.c_double, C.c_int,
                                        C.c_int, _ip]
        L.orc_boxes_relationships.restype = C.c_int
        L.orc_boxes_relationships.argtypes = [_ip, _ip]
        L.orc_pairs_reference.restype = C.c_longlong
        L.orc_pairs_reference.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, C.c_double, _ip,
                                          C.c_longlong]
        L.orc_hs_force.restype = C.c_double
        L.orc_hs_force.argtypes = [C.c_double] * 4
        L.orc_wca_energy.restype = C.c_double
        L.orc_wca_energy.argtypes = [C.c_double] * 4
        L.orc_kf2_force_torque.argtypes = [_dp, C.c_double, _dp, _dp, _dp, _dp, C.c_double,
                                           C.c_double, C.c_double, _dp]
        L.orc_kf2_energy.restype = C.c_double
        L.orc_kf2_energy.argtypes = [_dp, C.c_double, _dp, _dp, _dp, _dp, C.c_double,
                                     C.c_double, C.c_double]
        L.orc_pair_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, C.c_longlong, C.c_int,
                                      C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_angle_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, C.c_longlong, C.c_int,
                                       _dp, _dp, _dp]
        L.orc_active_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, _ip, C.c_double, _dp]
        L.orc_patch_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp, _ip, C.c_longlong,
                                       C.POINTER(PatchSet), C.c_double, _dp, _dp]
        L.orc_patch_energy.restype = C.c_double
        L.orc_patch_energy.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp, _ip, C.c_longlong,
                                       C.POINTER(PatchSet), C.c_double]
        L.orc_sphere.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_halfkick.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _dp, _dp]
        L.orc_advance_pos.argtypes = [C.POINTER(Params), C.POINTER(Box), C.c_int, _dp, _dp]
        L.orc_genfullmat.argtypes = [C.POINTER(Params), _dp, _dp]
        L.orc_rotate.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp]
        L.orc_advance_mom.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_noise_uniform.restype = C.c_double
        L.orc_noise_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_noise_fill.argtypes = [C.c_uint64, C.c_uint64, C.c_longlong, _dp]
        L.orc_compute_forces.restype = C.c_int
        L.orc_compute_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp,
                                         C.POINTER(Interactions), _dp, _dp]
        L.orc_run_rotors.restype = C.c_int
        L.orc_run_rotors.argtypes = [C.POINTER(Params), C.POINTER(Box), C.c_int, _dp, _dp, _dp,
                                     _dp, C.POINTER(Interactions), C.c_int, _dp, C.c_int, _dp]
        L.orc_run_beads.restype = C.c_int
        L.orc_run_beads.argtypes = [C.POINTER(Params), C.POINTER(Box), C.c_int, _dp, _dp,
                                    C.POINTER(Interactions), C.c_int, _dp, C.c_int, _dp]
        L.orc_measured_temperature.argtypes = [C.POINTER(Params), C.c_int, _dp, _dp, _dp]
        L.orc_stress.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip, C.c_longlong, C.c_double,
                                 C.c_double, C.c_double, _dp]
        L.orc_energy_virial.restype = C.c_int
        L.orc_energy_virial.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp,
                                        C.POINTER(Interactions), _dp, _dp]
        L.orc_bind_uniform.restype = C.c_double
        L.orc_bind_uniform.argtypes = [C.POINTER(C.c_uint32), C.c_uint64, C.c_int, C.c_int]
        L.orc_bind_forces.restype = C.c_int
        L.orc_bind_forces.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp, C.POINTER(PatchSet),
                                      C.POINTER(BindModel), _ip, _ip, _dp, _dp,
                                      C.POINTER(C.c_longlong)]
        L.orc_patch_list.restype = C.c_longlong
        L.orc_patch_list.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp, _ip, C.c_longlong,
                                     C.POINTER(PatchSet), C.c_double, C.c_double, _ip,
                                     C.c_longlong]
        L.orc_bound_pairs.restype = C.c_longlong
        L.orc_bound_pairs.argtypes = [C.POINTER(Box), C.c_int, _dp, _dp, _ip, C.c_longlong,
                                      C.POINTER(PatchSet), C.c_double, _ip, C.c_longlong]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def make_box(box, periodic=True):
    b = Box()
    box = np.asarray(box, dtype=np.float64).reshape(-1)
    if box.size == 1:
        box = np.repeat(box, 3)
    for k in range(3):
        b.l[k] = float(box[k])
    b.periodic = 1 if periodic else 0
    return b


def make_params(s):
    P = Params()
    P.dt, P.m, P.gamma, P.gammar, P.kT = s["dt"], s["m"], s["gamma"], s["gammar"], s["kT"]
    for k in range(3):
        P.I[k] = float(s["I"][k])
    return P


def fasterfloor(x):
    return lib().orc_fasterfloor(float(x))


def assign_all(box, x, n, periodic=True):
    b = make_box(box, periodic)
    x, xp = _d(x)
    nn, np_ = _i(np.broadcast_to(np.asarray(n), (3,)).copy())
    out = np.empty(len(x), dtype=np.int32)
    lib().orc_assign_all(C.byref(b), len(x), xp, np_, out.ctypes.data_as(_ip))
    return out


def boxes_relationships(n):
    n3 = np.broadcast_to(np.asarray(n), (3,)).astype(np.int32).copy()
    nc = int(np.prod(n3))
    out = np.zeros(nc * 27, dtype=np.int32)
    ncols = lib().orc_boxes_relationships(n3.ctypes.data_as(_ip), out.ctypes.data_as(_ip))
    return out[: nc * ncols].reshape(nc, ncols)


def pairs_reference(box, x, n, cut, periodic=True):
    b = make_box(box, periodic)
    x, xp = _d(x)
    n3, np_ = _i(np.broadcast_to(np.asarray(n), (3,)).copy())
    cap = len(x) * 32 + 64
    while True:
        out = np.empty(2 * cap, dtype=np.int32)
        cnt = lib().orc_pairs_reference(C.byref(b), len(x), xp, np_, float(cut),
                                        out.ctypes.data_as(_ip), cap)
        if cnt <= cap:
            return out[: 2 * cnt].reshape(cnt, 2)
        cap = cnt


class _Keep:
    """holds numpy arrays alive while ctypes structs point into them"""

    def __init__(self):
        self.refs = []

    def d(self, a):
        a, p = _d(a)
        self.refs.append(a)
        return p

    def i(self, a):
        a, p = _i(a)
        self.refs.append(a)
        return p


def make_patchset(ps, keep):
    P = PatchSet()
    P.ntypes = len(ps["npatch"])
    P.type = keep.i(ps["types"])
    for t, v in enumerate(ps["npatch"]):
        P.npatch[t] = int(v)
    P.dirs = keep.d(np.asarray(ps["dirs"]).reshape(-1, 3) if len(ps["dirs"]) else np.zeros(3))
    P.params = keep.d(ps["params"])
    keep.refs.append(P)
    return P


def make_interactions(s, keep, list_cut=None, rebuild_every=20, n_side=None):
    it = Interactions()
    eps, sig, rc = s["wca"] if s.get("wca") else (0.0, 1.0, 0.0)
    it.wca_on = 1 if s.get("wca") else 0
    it.wca_eps, it.wca_sigma, it.wca_rcut = eps, sig, rc
    ps = s.get("patches")
    if ps is not None:
        it.patches = C.pointer(make_patchset(ps, keep))
        it.maxd2 = float(ps.get("maxd2", 2.4))
    if s.get("bonds") is not None:
        it.bonds = keep.i(s["bonds"])
        it.nbonds = len(s["bonds"])
        it.bond_kind = int(s.get("bond_kind", 1))
        it.bond_k = keep.d(s["bond_k"])
        it.bond_r0 = keep.d(s["bond_r0"])
    if s.get("angles") is not None:
        it.angles = keep.i(s["angles"])
        it.nangles = len(s["angles"])
        it.angle_form = int(s.get("angle_form", 0))
        it.angle_k = keep.d(s["angle_k"])
        it.angle_theta0 = keep.d(s["angle_theta0"])
    if s.get("exclusions") is not None:
        it.excl = keep.i(s["exclusions"])
        it.nexcl = len(s["exclusions"])
    if s.get("active") is not None:
        it.fa = float(s["active"]["fa"])
        it.prev = keep.i(s["active"]["prev"])
        it.next = keep.i(s["active"]["next"])
    box = np.asarray(s["box"], dtype=np.float64)
    if list_cut is None:
        list_cut = 3.5
    it.list_cut = float(list_cut)
    it.rebuild_every = int(rebuild_every)
    if n_side is None:
        n_side = np.maximum(np.floor(box / 4.0), 1).astype(int)  # Condensate.cpp:4
    n_side = np.broadcast_to(np.asarray(n_side), (3,))
    for k in range(3):
        it.cells_per_side[k] = int(n_side[k])
    return it


def compute_forces(s, list_cut=None, n_side=None, x=None, Q=None):
    """Conservative forces (lab) and patch torques (lab) on a frozen
    configuration, summed over the reference's cell-list pair list."""
    keep = _Keep()
    it = make_interactions(s, keep, list_cut, 20, n_side)
    b = make_box(s["box"], s.get("periodic", True))
    x = s["x"] if x is None else x
    Q = s.get("Q") if Q is None else Q
    N = len(x)
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    xp = keep.d(x)
    Qp = keep.d(Q if Q is not None else np.tile(np.eye(3).ravel(), (N, 1)))
    rc = lib().orc_compute_forces(C.byref(b), N, xp, Qp, C.byref(it), F.ctypes.data_as(_dp),
                                  T.ctypes.data_as(_dp))
    if rc:
        raise RuntimeError("orc_compute_forces failed")
    return F, T


def energy_virial(s, list_cut=None, n_side=None, x=None):
    """total conservative energy U and virial W (3 x 3, W[k, l] = sum_i
    F_i,k x_i,l over the terms) on a frozen configuration"""
    keep = _Keep()
    it = make_interactions(s, keep, list_cut, 20, n_side)
    b = make_box(s["box"], s.get("periodic", True))
    x = s["x"] if x is None else x
    N = len(x)
    Q = s.get("Q")
    U = np.zeros(1)
    W = np.zeros(9)
    rc = lib().orc_energy_virial(C.byref(b), N, keep.d(x),
                                 keep.d(Q if Q is not None else np.tile(np.eye(3).ravel(), (N, 1))),
                                 C.byref(it), U.ctypes.data_as(_dp), W.ctypes.data_as(_dp))
    if rc:
        raise RuntimeError("orc_energy_virial failed")
    return float(U[0]), W.reshape(3, 3)


def stress(box, x, ij, eps=1.0, sigma=1.0, rcut=0.0, periodic=True):
    """MD::calculatestress over the pairs ij with an HSPotential interaction"""
    b = make_box(box, periodic)
    keep = _Keep()
    out = np.zeros(9)
    lib().orc_stress(C.byref(b), len(x), keep.d(x), keep.i(ij), len(ij), float(eps),
                     float(sigma), float(rcut), out.ctypes.data_as(_dp))
    return out.reshape(3, 3)


def patch_energy(s, list_cut=3.5, n_side=None):
    keep = _Keep()
    b = make_box(s["box"], s.get("periodic", True))
    if n_side is None:
        n_side = np.maximum(np.floor(np.asarray(s["box"]) / 4.0), 1).astype(int)
    pr = pairs_reference(s["box"], s["x"], n_side, list_cut, s.get("periodic", True))
    ps = make_patchset(s["patches"], keep)
    return lib().orc_patch_energy(C.byref(b), len(s["x"]), keep.d(s["x"]), keep.d(s["Q"]),
                                  keep.i(pr), len(pr), C.byref(ps),
                                  float(s["patches"].get("maxd2", 2.4)))


def kf2_force_torque(un, rij, Q1, Q2, nb1, nb2, att, dis, thetam):
    out = np.zeros(9)
    keep = _Keep()
    lib().orc_kf2_force_torque(keep.d(un), float(rij), keep.d(Q1), keep.d(Q2), keep.d(nb1),
                               keep.d(nb2), float(att), float(dis), float(thetam),
                               out.ctypes.data_as(_dp))
    return out


def kf2_energy(un, rij, Q1, Q2, nb1, nb2, att, dis, thetam):
    keep = _Keep()
    return lib().orc_kf2_energy(keep.d(un), float(rij), keep.d(Q1), keep.d(Q2), keep.d(nb1),
                                keep.d(nb2), float(att), float(dis), float(thetam))


def calcg1(box, x, max_, dx, periodic=True):
    b = make_box(box, periodic)
    x = np.ascontiguousarray(x, dtype=np.float64)
    cap = int(np.floor(max_ / dx)) + 1
    out = np.zeros(cap, dtype=np.int64)
    gs = lib().orc_calcg1(C.byref(b), len(x), x.ctypes.data_as(_dp), C.c_double(max_),
                          C.c_double(dx), out.ctypes.data_as(C.POINTER(C.c_longlong)), cap)
    return out[:gs]


def dgraph_labels(box, x, n, cut, lo=0, hi=-1, periodic=True):
    """distance_graph labels: the smallest index of each particle's connected
    component among the pairs within cut (orc_dgraph_labels)"""
    b = make_box(box, periodic)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n3 = np.ascontiguousarray(np.broadcast_to(np.asarray(n), (3,)), dtype=np.int32)
    out = np.empty(len(x), dtype=np.int32)
    np_ = lib().orc_dgraph_labels(C.byref(b), len(x), x.ctypes.data_as(_dp),
                                  n3.ctypes.data_as(_ip), C.c_double(cut), int(lo), int(hi),
                                  out.ctypes.data_as(_ip))
    if np_ < 0:
        raise MemoryError("orc_dgraph_labels")
    return out


def bind_forces(s, isbound, boundto, on_rate, off_rate, seed=None, key=None, draw=0, x=None,
                Q=None, tables=None):
    """One call of the one-bond-per-patch routine (orc_bind_forces): the
    reference's rand() draws after srand(seed) when seed is given, else Philox
    with key (2 words) and draw; tables (species, dtab, ttab, tr4) selects the
    species-table model.  Returns (F, T, isbound, boundto, stats)."""
    keep = _Keep()
    x = s["x"] if x is None else x
    Q = s["Q"] if Q is None else Q
    N = len(x)
    ps = make_patchset(s["patches"], keep)
    m = BindModel()
    m.on_rate, m.off_rate = float(on_rate), float(off_rate)
    if seed is not None:
        m.rng, m.seed = 0, int(seed)
    else:
        m.rng = 1
        m.key[0], m.key[1] = int(key[0]), int(key[1])
        m.draw = int(draw)
    if tables is not None:
        m.table = 1
        m.species = keep.i(tables["species"])
        m.dtab = keep.d(tables["dtab"])
        m.ttab = keep.d(tables["ttab"])
        m.tr4 = keep.d(tables["tr4"])
    ib, ibp = _i(np.array(isbound, dtype=np.int32))
    bt, btp = _i(np.array(boundto, dtype=np.int32))
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    st = np.zeros(5, dtype=np.int64)
    b = make_box(s["box"], s.get("periodic", True))
    xa, xp = _d(x)
    qa, qp = _d(Q)
    rc = lib().orc_bind_forces(C.byref(b), N, xp, qp, C.byref(ps), C.byref(m), ibp, btp,
                               F.ctypes.data_as(_dp), T.ctypes.data_as(_dp),
                               st.ctypes.data_as(C.POINTER(C.c_longlong)))
    if rc != 0:
        raise RuntimeError(f"orc_bind_forces: {rc}")
    return F, T, ib, bt, st


def patch_limits(params):
    """ComboPatch::max_check and max_ang of a GeneralPatch with these rows
    (eps, d, theta_m): 1.0 raised to the largest 1.4 d, 0.6 lowered to the
    smallest cos theta_m (combopatch.h:15, GeneralPatch.cpp:93-99)"""
    params = np.asarray(params, dtype=np.float64).reshape(-1, 3)
    mk = max(1.0, float(np.max(1.4 * params[:, 1])))
    ma = min(0.6, float(np.min(np.cos(params[:, 2]))))
    return mk, ma


def patch_list(s, pairs, maxdis=None, maxang=None):
    """calculate_patch_list on the given pairs (orc_patch_list): an M x 5
    int32 array (p1, p2, potn, patch of p1, patch of p2).  maxdis / maxang
    replace max_check + 1 / max_ang (for tests that plant a mistake)."""
    keep = _Keep()
    N = len(s["x"])
    ps = make_patchset(s["patches"], keep)
    mk, ma = patch_limits(s["patches"]["params"])
    if maxdis is not None:
        mk = maxdis - 1.0
    if maxang is not None:
        ma = maxang
    b = make_box(s["box"], s.get("periodic", True))
    xa, xp = _d(s["x"])
    qa, qp = _d(s["Q"])
    ij, ijp = _i(np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)))
    n = lib().orc_patch_list(C.byref(b), N, xp, qp, ijp, len(ij), C.byref(ps), mk + 1.0, ma,
                             None, 0)
    if n < 0:
        raise RuntimeError("orc_patch_list: particles not numbered by type")
    out = np.zeros((max(n, 1), 5), dtype=np.int32)
    lib().orc_patch_list(C.byref(b), N, xp, qp, ijp, len(ij), C.byref(ps), mk + 1.0, ma,
                         out.ctypes.data_as(_ip), n)
    return out[:n]


def bound_pairs(s, pairs):
    """calculate_bound_pairs on the given pairs (orc_bound_pairs): an M x 5
    int32 array as patch_list's"""
    keep = _Keep()
    N = len(s["x"])
    ps = make_patchset(s["patches"], keep)
    mk, _ = patch_limits(s["patches"]["params"])
    b = make_box(s["box"], s.get("periodic", True))
    xa, xp = _d(s["x"])
    qa, qp = _d(s["Q"])
    ij, ijp = _i(np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)))
    n = lib().orc_bound_pairs(C.byref(b), N, xp, qp, ijp, len(ij), C.byref(ps), mk, None, 0)
    if n < 0:
        raise RuntimeError("orc_bound_pairs: particles not numbered by type")
    out = np.zeros((max(n, 1), 5), dtype=np.int32)
    lib().orc_bound_pairs(C.byref(b), N, xp, qp, ijp, len(ij), C.byref(ps), mk,
                          out.ctypes.data_as(_ip), n)
    return out[:n]


def adjacency(rows):
    """adjacency (LangevinR.cpp:2000-2036): the rows whose particle pair
    differs from the row before"""
    rows = np.asarray(rows).reshape(-1, 5)
    if len(rows) < 2:
        return np.zeros(0, dtype=np.int32)
    d = (rows[1:, 0] != rows[:-1, 0]) | (rows[1:, 1] != rows[:-1, 1])
    return (np.nonzero(d)[0] + 1).astype(np.int32)


def bind_uniform(key, draw, root, k):
    kk = (C.c_uint32 * 2)(int(key[0]), int(key[1]))
    return lib().orc_bind_uniform(kk, int(draw), int(root), int(k))


def noise(seed, first, count):
    out = np.empty(int(count))
    lib().orc_noise_fill(C.c_uint64(seed), C.c_uint64(first), int(count), out.ctypes.data_as(_dp))
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def run_rotors(s, nsteps, noise_seed, snap_every=0, list_cut=3.5, rebuild_every=20, n_side=None):
    """Condensate::run order with injected noise orc_noise_uniform(seed, .)."""
    keep = _Keep()
    it = make_interactions(s, keep, list_cut, rebuild_every, n_side)
    b = make_box(s["box"], s.get("periodic", True))
    P = make_params(s)
    N = s["N"]
    x = np.array(s["x"], dtype=np.float64)
    p = np.array(s["p"], dtype=np.float64)
    Q = np.array(s["Q"], dtype=np.float64)
    L = np.array(s["L"], dtype=np.float64)
    nz = noise(noise_seed, 0, (nsteps + 1) * N * 6)
    nsnap = nsteps // snap_every if snap_every > 0 else 0
    snaps = np.zeros((max(nsnap, 1), N, 3))
    rc = lib().orc_run_rotors(C.byref(P), C.byref(b), N, x.ctypes.data_as(_dp),
                              p.ctypes.data_as(_dp), Q.ctypes.data_as(_dp), L.ctypes.data_as(_dp),
                              C.byref(it), int(nsteps), nz.ctypes.data_as(_dp), int(snap_every),
                              snaps.ctypes.data_as(_dp))
    if rc:
        raise RuntimeError("orc_run_rotors failed")
    return {"x": x, "p": p, "Q": Q, "L": L, "snaps": snaps[:nsnap]}


def run_beads(s, nsteps, noise_seed, snap_every=0, list_cut=3.5, rebuild_every=20, n_side=None):
    keep = _Keep()
    it = make_interactions(s, keep, list_cut, rebuild_every, n_side)
    b = make_box(s["box"], s.get("periodic", True))
    P = make_params(s)
    N = s["N"]
    x = np.array(s["x"], dtype=np.float64)
    p = np.array(s["p"], dtype=np.float64)
    nz = noise(noise_seed, 0, nsteps * N * 3)
    nsnap = nsteps // snap_every if snap_every > 0 else 0
    snaps = np.zeros((max(nsnap, 1), N, 3))
    rc = lib().orc_run_beads(C.byref(P), C.byref(b), N, x.ctypes.data_as(_dp),
                             p.ctypes.data_as(_dp), C.byref(it), int(nsteps),
                             nz.ctypes.data_as(_dp), int(snap_every), snaps.ctypes.data_as(_dp))
    if rc:
        raise RuntimeError("orc_run_beads failed")
    return {"x": x, "p": p, "snaps": snaps[:nsnap]}


def sphere(s, p, L, Q, F, T, RT):
    P = make_params(s)
    F = np.array(F, dtype=np.float64)
    T = np.array(T, dtype=np.float64)
    keep = _Keep()
    lib().orc_sphere(C.byref(P), len(p), keep.d(p), keep.d(L), keep.d(Q), F.ctypes.data_as(_dp),
                     T.ctypes.data_as(_dp), keep.d(RT))
    return F, T


def halfkick(s, p, L, F, T):
    """LangevinNVTR::advancemom_halfstep (orc_halfkick); returns (p, L)"""
    P = make_params(s)
    p = np.array(p, dtype=np.float64)
    L = np.array(L, dtype=np.float64)
    keep = _Keep()
    lib().orc_halfkick(C.byref(P), len(p), p.ctypes.data_as(_dp), L.ctypes.data_as(_dp),
                       keep.d(F), keep.d(T))
    return p, L


def genfullmat(s, L3):
    P = make_params(s)
    up = np.zeros(9)
    keep = _Keep()
    lib().orc_genfullmat(C.byref(P), keep.d(L3), up.ctypes.data_as(_dp))
    return up


def rotate(s, L, Q):
    P = make_params(s)
    L = np.array(L, dtype=np.float64)
    Q = np.array(Q, dtype=np.float64)
    lib().orc_rotate(C.byref(P), len(L), L.ctypes.data_as(_dp), Q.ctypes.data_as(_dp))
    return L, Q


def advance_mom(s, p, F, R):
    P = make_params(s)
    p = np.array(p, dtype=np.float64)
    keep = _Keep()
    lib().orc_advance_mom(C.byref(P), len(p), p.ctypes.data_as(_dp), keep.d(F), keep.d(R))
    return p


def advance_pos(s, x, p):
    P = make_params(s)
    b = make_box(s["box"], s.get("periodic", True))
    x = np.array(x, dtype=np.float64)
    p = np.array(p, dtype=np.float64)
    lib().orc_advance_pos(C.byref(P), C.byref(b), len(x), x.ctypes.data_as(_dp),
                          p.ctypes.data_as(_dp))
    return x, p


def pair_forces(box, x, ij, kind, p0, p1, p2=0.0, kk=None, r0=None, periodic=True):
    b = make_box(box, periodic)
    keep = _Keep()
    F = np.zeros((len(x), 3))
    lib().orc_pair_forces(C.byref(b), len(x), keep.d(x), keep.i(ij), len(ij), int(kind),
                          float(p0), float(p1), float(p2),
                          keep.d(kk) if kk is not None else None,
                          keep.d(r0) if r0 is not None else None, F.ctypes.data_as(_dp))
    return F


def angle_forces(box, x, ijk, form, k, theta0, periodic=True):
    b = make_box(box, periodic)
    keep = _Keep()
    F = np.zeros((len(x), 3))
    lib().orc_angle_forces(C.byref(b), len(x), keep.d(x), keep.i(ijk), len(ijk), int(form),
                           keep.d(k), keep.d(theta0), F.ctypes.data_as(_dp))
    return F


def active_forces(box, x, prev, nxt, fa, periodic=True):
    b = make_box(box, periodic)
    keep = _Keep()
    F = np.zeros((len(x), 3))
    lib().orc_active_forces(C.byref(b), len(x), keep.d(x), keep.i(prev), keep.i(nxt), float(fa),
                            F.ctypes.data_as(_dp))
    return F


def measured_temperature(s, p, L):
    P = make_params(s)
    out = np.zeros(6)
    keep = _Keep()
    lib().orc_measured_temperature(C.byref(P), len(p), keep.d(p), keep.d(L),
                                   out.ctypes.data_as(_dp))
    return out


# ------------------------------------------------------- reference harness


def have_ref():
    return os.path.exists(REF_BIN)


def run_ref(cmd, workdir, job: dict, arrays: dict, threads=1, timeout=3600):
    """Write a job directory, run oracle/_ref/mtref, return its stdout."""
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "job.txt"), "w") as f:
        for k, v in job.items():
            f.write(f"{k} {v}\n")
    for name, arr in arrays.items():
        np.ascontiguousarray(arr).tofile(os.path.join(workdir, name + ".bin"))
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run([REF_BIN, cmd, workdir], stdin=subprocess.DEVNULL, capture_output=True,
                       text=True, env=env, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"mtref {cmd} failed: {r.stderr[-2000:]}")
    return r.stdout


def read_ref(workdir, name, dtype=np.float64):
    return np.fromfile(os.path.join(workdir, name + ".bin"), dtype=dtype)
