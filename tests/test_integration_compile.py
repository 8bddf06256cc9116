"""INTEGRATION.md's reference-side binding compiled against the reference's
own classes: tests/cpp/integration_run_mdx.cpp is a LangevinNVTR subclass
(the reference's Code/MDBase/Rotational/LangevinR.h) whose run_mdx hands its
state to libmdx.so through include/mdx.h and copies it back.  Built by
`make -C oracle ref` into oracle/_ref/run_mdx where the reference sources are
present (the binary travels with the in-tree library)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "run_mdx")
REF = "/root/reference/Code"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_binding_compiles_against_the_reference_and_reports_without_a_gpu():
    """compiles and links against the real classes; on a host without a GPU
    run_mdx returns the engine's status and error text (no crash, no exit)"""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/run_mdx"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the GPU test runs the binding for real")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0
    assert out.stdout.startswith("run_mdx rc 2: "), out.stdout
    assert "device" in out.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/run_mdx not built here")
def test_binding_steps_the_reference_class_on_the_gpu():
    """the reference's LangevinNVTR state handed to the engine, 5 mixed steps
    there and back: the call returns MDX_OK"""
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "run_mdx rc 0", out.stdout
