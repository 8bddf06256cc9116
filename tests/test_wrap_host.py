"""Host builds of device helpers.  The integrator's position wrap (mdx::wrap1 in
mtcode_amd/csrc/mdx_device.h, cube::correct_position_and_momentum of the
reference's MDBase/geometry.h:403-445), fed NaN, +-inf and huge values: every
call returns, non-finite inputs are reported (the kernels then keep the
particle's previous state), finite results lie in the box, and a coordinate
one box length out rounds as the reference's loop does.  Also the drop-in's
patch-list types (patchint, adjacency) on the reference's patch lists.  CPU
only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def host_build_and_run(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
                    "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'mtcode_amd', 'csrc')}",
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe],
                   check=True, capture_output=True)
    return subprocess.run([exe], capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_wrap_bounded_on_host(tmp_path):
    r = host_build_and_run(tmp_path, "wrap_check")
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("ok")
    assert "FAIL" not in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_orientation_store_on_host(tmp_path):
    """The mixed engine's quaternion orientation store (quat_to_mat /
    mat_to_quat of mdx_device.h): matrix -> quaternion -> matrix on 200,000
    random rotations (all branches of Shepperd's method), unit norm, and
    the download / upload round trip within FP32 rounding"""
    r = host_build_and_run(tmp_path, "orient_check")
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("ok")


def test_patch_list_types_on_host(tmp_path):
    """the drop-in's patchint accessors and adjacency (host code of
    mtcode_amd/host/mtcode_amd.hpp) on the compiled reference's patch lists
    (tests/golden/plist.npz): the run starts equal the reference's adjacency"""
    import numpy as np

    exe = str(tmp_path / "plist_check")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'mtcode_amd', 'host')}",
                    os.path.join(ROOT, "tests", "cpp", "plist_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    PL = np.load(os.path.join(ROOT, "tests", "golden", "plist.npz"))
    for case in ("ab", "wide", "wall"):
        rows = np.ascontiguousarray(PL[case + "_plist"], dtype=np.int32)
        src, dst = tmp_path / f"{case}.bin", tmp_path / f"{case}_adj.bin"
        rows.tofile(src)
        r = subprocess.run([exe, str(src), str(dst)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.strip().endswith("ok")
        assert np.array_equal(np.fromfile(dst, dtype=np.int32), PL[case + "_adjacency"])
