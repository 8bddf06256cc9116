"""Host build of the integrator's position wrap (mdx::wrap1 in
mtcode_amd/csrc/mdx_device.h, cube::correct_position_and_momentum of the
reference's MDBase/geometry.h:403-445), fed NaN, +-inf and huge values: every
call returns, non-finite inputs are reported (the kernels then keep the
particle's previous state), finite results lie in the box, and a coordinate
one box length out rounds as the reference's loop does.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_wrap_bounded_on_host(tmp_path):
    exe = str(tmp_path / "wrap_check")
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
                    "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'mtcode_amd', 'csrc')}",
                    os.path.join(ROOT, "tests", "cpp", "wrap_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("ok")
    assert "FAIL" not in r.stdout
