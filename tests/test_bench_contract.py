"""CPU-side pieces of bench.py: the one-GPU refusal, the committed HBM-traffic
record the roofline entry reads, and the reference CPU-baseline leg on a small
sample (when the reference harness is built)."""
import os
import subprocess
import sys

import pytest

import pyoracle as po
from mtcode_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_traffic_record_matches_the_step_phases():
    t = bench.measured_traffic()
    assert t is not None
    assert t["N"] == 1 << 20
    assert set(t["per_phase_bytes"]) == set(bench.STEP_PHASES)
    total = sum(t["per_phase_bytes"].values())
    assert abs(total - t["bytes_per_step"]) <= 1e-9 * total
    # the counters are scaled by a plain 256 MiB device copy: FETCH_SIZE read
    # half the bytes that copy moved (scale 2.0), WRITE_SIZE all of them (1.0)
    assert 0.5 <= t["fetch_scale"] <= 4 and 0.5 <= t["write_scale"] <= 4
    # HBM bytes of a whole step cannot undercut the state the step must
    # read and write once (172 B per particle in mixed precision)
    assert bench.ALGO_BYTES["mixed"] == 172
    assert t["bytes_per_step"] / t["N"] >= bench.ALGO_BYTES["mixed"]


@pytest.mark.skipif(not po.have_ref(), reason="reference harness (oracle/_ref/mtref) not built")
def test_cpu_baseline_leg_on_a_small_sample():
    s = wl.condensate_ab(4096, phi=0.05)
    cb = bench.cpu_baseline(s, 2, 1)
    assert cb["kind"] == "reference"
    assert cb["unit"] == "particle-steps/s"
    assert cb["cores"] == 1
    assert cb["value"] > 0
    assert cb["sample"].startswith("4096 particles, 2 steps")
