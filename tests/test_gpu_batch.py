"""Batched 1-loop right-hand sides (pffrg_batch_rhs_1l): one vertex-pass launch over the
elements of several problems of different lattices, sizes, couplings, meshes and Λ must give,
per problem, exactly the result of that problem's own compute_dΣ! + compute_dΓ_1l! (the same
kernel body runs on the same arguments), and the FP64 oracle's to rounding."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg.mesh import get_eval
from pffrg.native import FP32_MIXED, FP64, Batch

from helpers import rel_err, small_problem

pytestmark = pytest.mark.gpu

PROBLEMS = [  # (lattice, size, couplings, Λ, mesh size, seed)
    ("square", 2, ((1.0,),), 1.0, 6, 3),
    ("kagome", 3, ((1.0,),), 0.7, 7, 4),
    ("pyrochlore", 3, ((1.0,),), 2.0, 6, 5),
    ("square", 3, ((1.0,), (0.5,)), 0.3, 8, 6),
]


def _problems():
    out = []
    for name, size, J, L, N, seed in PROBLEMS:
        prob = small_problem(name, size, N, 1.0, J)
        a = A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=seed)
        out.append((prob, a, L, get_eval(L)))
    return out


@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_batch_equals_single(precision):
    probs = _problems()
    ctxs = [p.context(n_slots=3, precision=precision) for p, _, _, _ in probs]
    try:
        for ctx, (p, a, L, ev) in zip(ctxs, probs):
            ctx.upload(0, a)
            ctx.compute_dsigma(L, 0, 1)
            ctx.compute_dgamma_1l(L, 0, 1, ev)
        single = [ctx.download(1) for ctx in ctxs]
        with Batch(ctxs) as b:
            b.rhs_1l([L for _, _, L, _ in probs], 0, 2, [ev for _, _, _, ev in probs])
            b.rhs_1l([L for _, _, L, _ in probs], 0, 2, [ev for _, _, _, ev in probs])  # reuse of the batch
        for k, ctx in enumerate(ctxs):
            got = ctx.download(2)
            print(f"problem {k}: max |batch - single| = {np.max(np.abs(got.gamma - single[k].gamma)):.3e}", flush=True)
            np.testing.assert_array_equal(got.gamma, single[k].gamma)
            np.testing.assert_array_equal(got.Sigma, single[k].Sigma)
    finally:
        for ctx in ctxs:
            ctx.close()


def test_batch_matches_oracle():
    probs = _problems()
    ctxs = [p.context(n_slots=2) for p, _, _, _ in probs]
    try:
        for ctx, (p, a, L, ev) in zip(ctxs, probs):
            ctx.upload(0, a)
        with Batch(ctxs) as b:
            b.rhs_1l([L for _, _, L, _ in probs], 0, 1, [ev for _, _, _, ev in probs])
        for ctx, (p, a, L, ev) in zip(ctxs, probs):
            orc = O.Oracle(p.r, p.mesh)
            ref = A.Action.empty(p.r.n_sites, p.mesh)
            orc.compute_dsigma(L, a, ref)
            orc.compute_dgamma_1l(L, a, ref, ev)
            got = ctx.download(1)
            e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
            print(f"{p.name} L={L}: {e:.3e}", flush=True)
            assert e < 1e-12
    finally:
        for ctx in ctxs:
            ctx.close()


def test_batch_chi_equals_single():
    # pffrg_batch_compute_chi: every problem's χ equals its own compute_chi bitwise, region counts
    # included; a problem left out returns nothing
    probs = _problems()
    ctxs = [p.context(n_slots=2) for p, _, _, _ in probs]
    try:
        single = []
        for ctx, (p, a, L, ev) in zip(ctxs, probs):
            ctx.upload(0, a)
            single.append(ctx.compute_chi(L, 0))
        active = [k != 1 for k in range(len(probs))]
        with Batch(ctxs) as b:
            got = b.compute_chi([L for _, _, L, _ in probs], 0, active=active)
        for k, (g, s) in enumerate(zip(got, single)):
            if not active[k]:
                assert g is None
                continue
            print(f"problem {k}: chi {g.shape}, max |batch - single| = {np.max(np.abs(g - s)):.3e}", flush=True)
            np.testing.assert_array_equal(g, s)
    finally:
        for ctx in ctxs:
            ctx.close()
