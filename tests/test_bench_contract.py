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
    """the committed record's shape (whether or not it is current)"""
    import json

    with open(bench.TRAFFIC_FILE) as f:
        t = json.load(f)
    assert t["N"] == 1 << 20
    assert set(t["per_phase_bytes"]) == set(bench.STEP_PHASES)
    total = sum(t["per_phase_bytes"].values())
    assert abs(total - t["bytes_per_step"]) <= 1e-9 * total
    # the counters are scaled by a plain 256 MiB device copy: FETCH_SIZE read
    # half the bytes that copy moved (scale 2.0), WRITE_SIZE all of them (1.0)
    assert 0.5 <= t["fetch_scale"] <= 4 and 0.5 <= t["write_scale"] <= 4
    # HBM bytes of a whole step cannot undercut the state the engine itself
    # must read and write once: x 24 + p 12 + Q 16 + L 12 B, and the type word
    assert t["bytes_per_step"] / t["N"] >= 2 * (24 + 12 + 16 + 12) + 4
    # the workload's algorithmic figure (x split 36 + p 12 + Q 36 + L 12 per
    # rotor, read and written, and the type word)
    assert bench.ALGO_BYTES["mixed"] == 2 * (36 + 12 + 36 + 12) + 4


def test_traffic_record_is_reported_only_for_its_kernel_sources(tmp_path, monkeypatch):
    """a record measured on other kernel sources is stale: bench reports no
    traffic for it"""
    import json

    from mtcode_amd import profiling

    rec = {"N": 1 << 20, "bytes_per_step": 1.0, "per_phase_bytes": {}}
    f = tmp_path / "t.json"
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    f.write_text(json.dumps(dict(rec, source="0" * 16)))
    assert bench.measured_traffic() is None
    f.write_text(json.dumps(dict(rec, source=profiling.source_stamp())))
    assert bench.measured_traffic()["bytes_per_step"] == 1.0


def test_profiling_summaries_of_rocprof_csv(tmp_path):
    """the kernel-trace table and the copy-probe-scaled traffic of synthetic
    rocprofv3 CSVs"""
    from mtcode_amd import profiling

    tr = tmp_path / "trace.csv"
    tr.write_text('"Kernel_Name","Start_Timestamp","End_Timestamp"\n'
                  '"void mdx::k_decide(mdx::StepCtl*)",1000,5000\n'
                  '"void mdx::k_decide(mdx::StepCtl*)",9000,11000\n')
    rows = profiling.trace_summary(str(tr))
    assert rows == [("k_decide", 2, 3.0, 2.0, 4.0, 0.006)]
    head = '"Kernel_Name","Counter_Name","Counter_Value"\n'
    fe = tmp_path / "f.csv"
    wr = tmp_path / "w.csv"
    # the copy: 256 MiB read counted as 131072 KiB (scale 2), 262144 KiB written
    fe.write_text(head + '"__amd_rocclr_copyBuffer","FETCH_SIZE",131072\n'
                  '"void mdx::k_integrate<float, 3>(mdx::KArgs<float>)","FETCH_SIZE",1000\n'
                  '"void mdx::k_integrate<float, 3>(mdx::KArgs<float>)","FETCH_SIZE",3000\n')
    wr.write_text(head + '"__amd_rocclr_copyBuffer","WRITE_SIZE",262144\n'
                  '"void mdx::k_integrate<float, 3>(mdx::KArgs<float>)","WRITE_SIZE",500\n')
    t = profiling.traffic(str(fe), str(wr), 1024)
    assert t["fetch_scale"] == 2.0 and t["write_scale"] == 1.0
    k = t["per_kernel"]["k_integrate<float, 3>"]
    assert k["fetch_bytes"] == 2.0 * 2000 * 1024 and k["write_bytes"] == 500 * 1024
    assert t["per_phase_bytes"]["finish"] == k["fetch_bytes"] + k["write_bytes"]
    assert t["source"] == profiling.source_stamp()


@pytest.mark.skipif(not po.have_ref(), reason="reference harness (oracle/_ref/mtref) not built")
def test_cpu_baseline_leg_on_a_small_sample():
    s = wl.condensate_ab(4096, phi=0.05)
    cb = bench.cpu_baseline(s, 2, 1)
    assert cb["kind"] == "reference"
    assert cb["unit"] == "particle-steps/s"
    assert cb["cores"] == 1
    assert cb["value"] > 0
    assert cb["sample"].startswith("4096 particles, 2 steps")
