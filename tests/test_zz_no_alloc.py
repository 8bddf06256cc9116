"""mdx_stress_noalloc (the stress entry point that allocates nothing, its
kernels reducing into a static device scratch) against mdx_stress, through
the C ABI on an MI355X: two calls in a row, the second on a smaller input."""
import sys

import numpy as np
import pytest

from mtcode_amd import mdx
from mtcode_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_stress_noalloc_matches_stress():
    """FP64 sums in another order: within 1e-12 of the largest component;
    a larger system, then a smaller one (and a single replica of a batch),
    so the second call reads only what it wrote itself"""
    big = wl.condensate_ab(1 << 17, phi=0.1)
    small = wl.patchy_tetra(3000)
    reps = wl.replicas(5, seed=8, nmin=300, nmax=2000)
    cases = [("big", big, None, -1), ("small", small, None, -1), ("replica 3", None, reps, 3)]
    for name, s, r, rep in cases:
        with mdx.engine_from_system(s, precision=mdx.PREC_FP64, replicas=r) as e:
            with pytest.raises(mdx.MdxError):
                e.stress_noalloc(rep)  # no cell grid yet
            ref = e.stress(rep)
            got = e.stress_noalloc(rep)
            say(name, "stress", ref[0, 0], "noalloc", got[0, 0], "err", relmax(got, ref))
            assert relmax(got, ref) < 1e-12
            e.prime()
            e.step(5)
            got2 = e.stress_noalloc(rep)
            ref2 = e.stress(rep)
            assert relmax(got2, ref2) < 1e-12
