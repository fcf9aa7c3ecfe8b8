"""The gather-ceiling probe (pffrg_gather_ceiling): random trilinear q3 lookups on a slot's
gather copy, over the whole q3 block or confined to a footprint.  It measures rather than
computes, so the checks are that it runs on every lattice shape the vertex pass supports (one
site group per wave up to 16 groups) and footprint (whole block, the minimum of two ν' planes,
2 MiB), reports a positive rate, and rejects arguments outside its documented ranges."""
import math

import pytest

from pffrg import action as A
from pffrg import native
from pffrg.native import FP32_MIXED, FP64

from helpers import small_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,size,N", [("square", 2, 6), ("square", 3, 8), ("pyrochlore", 3, 6)])
def test_gather_ceiling_runs(name, size, N):
    prob = small_problem(name, size, N)
    ctx = prob.context(n_slots=1)
    try:
        ctx.upload(0, A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=1))
        for mode in (FP32_MIXED, FP64):
            for w in (0, 3):
                for fp in (0, 1, 1 << 21):
                    gbps, ms = ctx.gather_ceiling(0, mode, w, fp, n_iter=16, n_rep=2)
                    print(f"{name} L={size} ns={prob.r.n_sites} mode={mode} waves={w} footprint={fp}: "
                          f"{gbps:.0f} GB/s, {ms:.3f} ms", flush=True)
                    assert math.isfinite(gbps) and gbps > 0 and ms > 0
    finally:
        ctx.close()


def test_gather_ceiling_rejects_bad_arguments():
    prob = small_problem()
    ctx = prob.context(n_slots=1)
    try:
        for args in ((1, FP64, 0, 0, 16, 1), (0, 7, 0, 0, 16, 1), (0, FP64, 9, 0, 16, 1), (0, FP64, 0, -1, 16, 1),
                     (0, FP64, 0, 0, 0, 1), (0, FP64, 0, 0, 16, 0)):
            with pytest.raises(native.PffrgError):
                ctx.gather_ceiling(*args)
    finally:
        ctx.close()
