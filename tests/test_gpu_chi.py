"""Correlations on the GPU (pffrg_compute_chi) against the oracle's compute_χ! restatement.

Both run the same adaptive cubature with the same arithmetic (the GPU file is compiled without
FMA contraction), so every box ends with the same region count and χ agrees to rounding:
relative max-norm error <= 1e-12 (FP64)."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A

from helpers import rel_err, small_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,size,L", [("square", 2, 1.0), ("kagome", 3, 0.5), ("square", 3, 2.0)])
def test_chi_matches_oracle(name, size, L):
    prob = small_problem(name, size, 6, 1.0)
    a = A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=11)
    orc = O.Oracle(prob.r, prob.mesh)
    ref, nr_ref = orc.compute_chi(L, a, with_regions=True)
    with prob.context(n_slots=1) as ctx:
        ctx.upload(0, a)
        got, nr = ctx.compute_chi(L, 0, with_regions=True)
    e = rel_err(got, ref)
    print(f"chi {name}{size} L={L}: rel err {e:.3e}, regions {nr.sum()} (oracle {nr_ref.sum()}, max {nr.max()})",
          flush=True)
    assert np.array_equal(nr, nr_ref)
    assert e <= 1e-12
