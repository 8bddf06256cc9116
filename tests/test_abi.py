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
