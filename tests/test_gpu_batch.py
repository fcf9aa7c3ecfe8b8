"""Batched 1-loop right-hand sides (pffrg_batch_rhs_1l): one vertex-pass launch over the
elements of several problems of different lattices, sizes, couplings, meshes and Λ must give,
per problem, exactly the result of that problem's own compute_dΣ! + compute_dΓ_1l! (the same
kernel body runs on the same arguments), and the FP64 oracle's to rounding."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import flow as F
from pffrg.mesh import get_eval
from pffrg.native import FP32_MIXED, FP64, Batch

from helpers import entries, half_elements, rel_err, small_problem

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


def test_batch_workload_size_equals_single_and_oracle_sample():
    # the batched sweep at its real shape: 16 square L=5 J1-J2 problems (J2/J1 = k/15) and square
    # L=3..7 at J2/J1 = 0.5, N=32, Λ=1, eval=10.  The problems fall into both LDS launch classes;
    # each problem's batched RHS equals its own in both precisions, and the largest truncation
    # (L=7, 20 sites) matches the oracle on a strided sample of Katanin elements
    specs = [(5, k / 15.0) for k in range(16)] + [(L, 0.5) for L in (3, 4, 5, 6, 7)]
    N = 32
    probs = [F.Problem.build("square", L, [[1.0], [j2]], initial=1.0, num_sigma=N, num_Omega=N, num_nu=N,
                             num_chi=10) for L, j2 in specs]
    acts = [A.synthetic_action(p.r.n_sites, p.mesh, p.bare, seed=20261203 + k) for k, p in enumerate(probs)]
    n = len(probs)
    for precision in (FP64, FP32_MIXED):
        ctxs = [p.context(n_slots=3, precision=precision) for p in probs]
        try:
            for c, a in zip(ctxs, acts):
                c.upload(0, a)
                F.rhs(c, 1.0, 1, 0, 1, 10)
            with Batch(ctxs) as b:
                b.rhs_1l([1.0] * n, 0, 2, [10] * n)
            for k, c in enumerate(ctxs):
                single, got = c.download(1), c.download(2)
                np.testing.assert_array_equal(got.gamma, single.gamma, err_msg=f"problem {k} {specs[k]}")
                np.testing.assert_array_equal(got.Sigma, single.Sigma, err_msg=f"problem {k} {specs[k]}")
            print(f"precision {precision}: {n} batched RHS equal their single runs bitwise", flush=True)
            if precision != FP64:
                continue
            p, a, got = probs[-1], acts[-1], ctxs[-1].download(2)
            orc = O.Oracle(p.r, p.mesh)
            dd = A.Action.empty(p.r.n_sites, p.mesh)
            dd.Sigma[:] = got.Sigma
            orc.compute_corrs(1.0, a, dd, True)
            out = A.Action.empty(p.r.n_sites, p.mesh)
            sample = half_elements(N, N)[::613]
            for e in sample:
                orc.compute_element(O.P_KAT, 1.0, a, dd, *e, 10, out)
            scale = float(np.max(np.abs(got.gamma)))
            ev = max(float(np.max(np.abs(entries(got, *e) - entries(out, *e)))) for e in sample) / scale
            ref_s = A.Action.empty(p.r.n_sites, p.mesh)
            orc.compute_dsigma(1.0, a, ref_s)
            es = rel_err(got.Sigma, ref_s.Sigma)
            print(f"square L=7 (20 sites) in the batch: dSigma {es:.2e}, {len(sample)} sampled elements {ev:.2e}",
                  flush=True)
            assert es < 1e-12 and ev < 1e-12
        finally:
            for c in ctxs:
                c.close()


def test_batch_stage_algebra_equals_per_context():
    # pffrg_batch_replace_with / _mult_with_add_to / _abs_max (the BS3 stage algebra of all flows
    # of a batch, Action.jl:181-359) against the per-context calls: bitwise, with each problem's
    # own factor; a bad slot is refused before any problem is touched
    probs = _problems()
    ctxs = [p.context(n_slots=4) for p, _, _, _ in probs]
    ref = [p.context(n_slots=4) for p, _, _, _ in probs]
    try:
        for k, (p, a, _, _) in enumerate(probs):
            b2 = A.synthetic_action(p.r.n_sites, p.mesh, p.bare, seed=90 + k)
            for c in (ctxs[k], ref[k]):
                c.upload(0, a)
                c.upload(1, b2)
        facs = [0.37, -1.25, 2.5e-3, -0.8]
        for c, f in zip(ref, facs):
            c.replace_with(2, 0)
            c.mult_with_add_to(1, f, 2)
            c.mult_with_add_to(0, -f, 1, vertex_only=True)
        want_max = [c.abs_max(2) for c in ref]
        with Batch(ctxs) as b:
            b.replace_with(2, 0)
            b.mult_with_add_to(1, facs, 2)
            b.mult_with_add_to(0, [-f for f in facs], 1, vertex_only=True)
            got_max = b.abs_max(2)
            with pytest.raises(Exception):
                b.mult_with_add_to(1, facs, 9)
        for k in range(len(probs)):
            for s in (1, 2):
                g, w = ctxs[k].download(s), ref[k].download(s)
                np.testing.assert_array_equal(g.gamma, w.gamma)
                np.testing.assert_array_equal(g.Sigma, w.Sigma)
        np.testing.assert_array_equal(got_max, np.array(want_max))
        print(f"batched stage algebra = per-context, abs max {got_max}", flush=True)
    finally:
        for c in ctxs + ref:
            c.close()
