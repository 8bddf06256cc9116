"""The C-ABI library loads without a GPU and exports every function that
include/mdx.h declares (no compute calls here)."""
import ctypes
import os
import subprocess

import pytest

from mtcode_amd import mdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_the_boundary():
    names = mdx.declared_symbols()
    for required in ["mdx_create", "mdx_destroy", "mdx_last_error", "mdx_upload_state",
                     "mdx_download_state", "mdx_set_pair_wca", "mdx_set_patches",
                     "mdx_set_bonds", "mdx_set_angles", "mdx_set_active", "mdx_cell_assign",
                     "mdx_neighbor_pairs", "mdx_compute_forces", "mdx_prime", "mdx_step",
                     "mdx_observables"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(mdx.LIB_PATH):
        pytest.skip("libmdx.so not built (python -m mtcode_amd.build)")
    lib = ctypes.CDLL(mdx.LIB_PATH)
    missing = [n for n in mdx.declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # and nothing mdx_* beyond them is exported
    out = subprocess.run(["nm", "-D", "--defined-only", mdx.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if " T mdx_" in ln})
    assert exported == mdx.declared_symbols()


def test_errors_are_returned_not_raised_as_exits():
    if not os.path.exists(mdx.LIB_PATH):
        pytest.skip("libmdx.so not built")
    L = mdx.lib()
    # a null descriptor is an argument error with a message, never an exit
    rc = L.mdx_create(None, ctypes.byref(ctypes.c_void_p()))
    assert rc == mdx.MDX_ERR_ARG
    assert b"descriptor" in L.mdx_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(mdx, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(mdx, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        mdx.lib()


def _desc(n, box=10.0, dt=0.005, precision=None, integrator=None, kT=1.0, nrep=None):
    nn = (ctypes.c_int * len(n))(*n)
    bx = (ctypes.c_double * (3 * len(n)))(*([box] * 3 * len(n)))
    kt = (ctypes.c_double * len(n))(*([kT] * len(n)))
    d = mdx.Desc()
    d.nreplicas = len(n) if nrep is None else nrep
    d.n_per_replica = ctypes.cast(nn, ctypes.POINTER(ctypes.c_int))
    d.box = ctypes.cast(bx, ctypes.POINTER(ctypes.c_double))
    d.periodic = 1
    d.kT = ctypes.cast(kt, ctypes.POINTER(ctypes.c_double))
    d.dt, d.m, d.gamma, d.gammar = dt, 1.0, 1.0, 1.0
    for k in range(3):
        d.inertia[k] = 0.1
    d.precision = mdx.PREC_FP64 if precision is None else precision
    d.integrator = mdx.INTEG_ROTOR if integrator is None else integrator
    return d, (nn, bx, kt)


@pytest.mark.parametrize("kw,msg", [
    (dict(n=[0]), b"no particles"),
    (dict(n=[10, 0]), b"replica 1 has no particles"),
    (dict(n=[10], nrep=0), b"nreplicas"),
    (dict(n=[10], box=0.0), b"box edge"),
    (dict(n=[10], box=-2.0), b"box edge"),
    (dict(n=[10], dt=0.0), b"dt"),
    (dict(n=[10], kT=-1.0), b"kT"),
    (dict(n=[10], precision=7), b"precision"),
    (dict(n=[10], integrator=9), b"integrator"),
])
def test_descriptor_errors_before_any_device_call(kw, msg):
    """Empty systems and replicas, non-positive boxes and steps, unknown
    modes: rejected with MDX_ERR_ARG and a message, validated before the
    device is touched (so this runs without a GPU)"""
    if not os.path.exists(mdx.LIB_PATH):
        pytest.skip("libmdx.so not built")
    L = mdx.lib()
    d, keep = _desc(**kw)
    h = ctypes.c_void_p()
    rc = L.mdx_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == mdx.MDX_ERR_ARG
    assert msg in L.mdx_last_error()
    assert not h.value


def test_null_engine_is_an_argument_error():
    if not os.path.exists(mdx.LIB_PATH):
        pytest.skip("libmdx.so not built")
    L = mdx.lib()
    assert L.mdx_step(None, 1, None) == mdx.MDX_ERR_ARG
    assert L.mdx_prime(None, None) == mdx.MDX_ERR_ARG
    assert L.mdx_download_state(None, None, None, None, None) == mdx.MDX_ERR_ARG
    assert L.mdx_upload_state(None, None, None, None, None) == mdx.MDX_ERR_ARG
