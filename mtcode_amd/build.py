"""Builds libmdx.so (the HIP engine behind include/mdx.h) in-tree for gfx950.

    python -m mtcode_amd.build          # or: from mtcode_amd.build import build; build()

hipcc cross-compiles without a GPU.  Packed FP32 instructions are disabled
(-packed-fp32-ops): some of their forms have given wrong results on these
machines, and hipcc would otherwise form them from ordinary float code.
Floating-point contraction is off by default so that FP64 results (and every
FP64 cut-off or cone decision) round as the reference's own arithmetic does.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmdx.so")
SOURCES = [os.path.join(HERE, "csrc", "mdx_engine.hip")]
HEADERS = [os.path.join(HERE, "csrc", "mdx_device.h"), os.path.join(HERE, "csrc", "mdx_kernels.h"),
           os.path.join(ROOT, "include", "mdx.h")]
ARCH = os.environ.get("MDX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build libmdx.so)")


def compile_cmd(out: str = LIB) -> list:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # FP64 paths round like the reference's x86 code (no fused multiply-add);
        # the FP32 pair math re-enables contraction locally
        "-ffp-contract=off",
        "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}",
        *SOURCES, "-o", out,
    ]


def build(verbose: bool = False) -> str:
    """Compile libmdx.so (always; the call is the 'does it build' check)."""
    tmp = LIB + ".tmp"
    cmd = compile_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed to build libmdx.so")
    os.replace(tmp, LIB)
    return LIB


HOST_HEADER = os.path.join(HERE, "host", "mtcode_amd.hpp")
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


def build_driver(verbose: bool = False) -> str:
    """Compile the C++ drop-in test driver (host classes over libmdx.so) with
    the host compiler; no HIP headers are involved."""
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(HERE, 'host')}", DRIVER_SRC, f"-L{HERE}", "-lmdx",
           "-Wl,-rpath,$ORIGIN/../../mtcode_amd", "-o", DRIVER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("failed to build the drop-in driver")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_driver(verbose=True))
