"""The shared-node vertex pass (the FP32-mixed mode) against the FP64 GPU path and the FP32
per-element pass, through the C ABI.

The shared-node pass integrates a tile of (ν, ν') elements of one (channel, Ω) together on the
nodes they share (quadrature.jl:2-139: every element of a block has the same sub-intervals and
levels); each element keeps its own Richardson test and stops at its own level.  It computes
in FP32 what the per-element FP32-mixed pass computes in FP32, with FP64 node sums, so:
 - against the FP64 path, every array within the FP32-mixed mode's bound, 1e-4 relative
   (max-norm per array, the reference's norm);
 - against the per-element FP32 pass, within 2e-5: the two differ only in summation order
   and in the reweighted nodes, which the shared pass takes from the previous level.
Each wavefront layout by site count is covered: one thread per element (2-8 sites), two
(9-16), four (17-32) and eight (33-64 sites), with the tile shapes that go with them."""
import numpy as np
import pytest

from pffrg import action as A
from pffrg import flow as F
from pffrg.native import FP32_ELEMENT, FP32_MIXED, FP64

from helpers import per_array_err, rel_err, small_problem

pytestmark = pytest.mark.gpu

TOL32 = 1e-4
TOL_ELEM = 2e-5


def _rhs_modes(prob, a, loops, modes, L=1.0, eval_=10):
    ns, m = prob.r.n_sites, prob.mesh
    out = {}
    for mode in modes:
        with prob.context(n_slots=3, precision=mode) as ctx:
            ctx.upload(0, a)
            ctx.upload(1, A.Action.empty(ns, m))
            ctx.upload(2, A.Action.empty(ns, m))
            F.rhs(ctx, L, loops, 0, 1, eval_, da_l=2)
            out[mode] = (ctx.download(1), ctx.stats())
    return out


@pytest.mark.parametrize("name,size,N,loops", [("square", 2, 8, 1), ("square", 3, 8, 2), ("kagome", 3, 8, 1),
                                               ("square", 6, 6, 1), ("pyrochlore", 4, 6, 1), ("square", 9, 5, 1),
                                               ("pyrochlore", 6, 4, 1)])
def test_shared_rhs_against_fp64(name, size, N, loops):
    prob = small_problem(name, size, N, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=40 + size)
    print(f"shared pass {name} L={size} sites {ns} N={N} {loops}-loop", flush=True)
    out = _rhs_modes(prob, a, loops, (FP64, FP32_ELEMENT, FP32_MIXED))
    g64, s64 = out[FP64]
    gel, _ = out[FP32_ELEMENT]
    gsh, ssh = out[FP32_MIXED]
    e64 = per_array_err(gsh, g64)
    eel = per_array_err(gsh, gel)
    w64 = max(e64.items(), key=lambda kv: kv[1])
    wel = max(eel.items(), key=lambda kv: kv[1])
    print(f"   vs FP64 worst {w64[0]} {w64[1]:.2e}; vs FP32 per-element worst {wel[0]} {wel[1]:.2e}; "
          f"element-nodes {ssh['nodes']} (FP64 pass {s64['nodes']})", flush=True)
    assert w64[1] < TOL32
    assert wel[1] < TOL_ELEM
    assert rel_err(gsh.Sigma, g64.Sigma) < 1e-12


def test_shared_multiloop_central_pass():
    prob = small_problem("pyrochlore", 3, 6, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=51)
    print(f"shared pass pyrochlore L=3 sites {ns} 3-loop with Σ corrections", flush=True)
    res = {}
    for mode in (FP64, FP32_MIXED):
        with prob.context(n_slots=7, precision=mode) as ctx:
            ctx.upload(0, a)
            for k in range(1, 7):
                ctx.upload(k, A.Action.empty(ns, m))
            F.rhs(ctx, 1.0, 3, 0, 1, 10, da_l=2, da_c=3, da_temp=4, da_sigma=5, sigma_corr=True)
            res[mode] = ctx.download(1)
    e = per_array_err(res[FP32_MIXED], res[FP64])
    w = max(e.items(), key=lambda kv: kv[1])
    print(f"   vs FP64 worst {w[0]} {w[1]:.2e}", flush=True)
    assert w[1] < TOL32


def test_shared_parquet_passes():
    # compute_Σ! (reduced s bubble) and compute_Γ! (BSE) of the parquet loop
    prob = small_problem("kagome", 3, 6, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=52)
    print(f"shared pass kagome L=3 sites {ns}: SDE and BSE", flush=True)
    res = {}
    for mode in (FP64, FP32_MIXED):
        with prob.context(n_slots=3, precision=mode) as ctx:
            ctx.upload(0, a)
            ctx.upload(1, A.Action.empty(ns, m))
            ctx.upload(2, A.Action.empty(ns, m))
            ctx.compute_sigma(1.0, 0, 1, 10)
            ctx.compute_gamma(1.0, 0, 2, 10)
            res[mode] = (ctx.download(1).Sigma, ctx.download(2))
    es = rel_err(res[FP32_MIXED][0], res[FP64][0])
    e = per_array_err(res[FP32_MIXED][1], res[FP64][1])
    w = max(e.items(), key=lambda kv: kv[1])
    print(f"   SDE Σ vs FP64 {es:.2e}; BSE worst {w[0]} {w[1]:.2e}", flush=True)
    assert es < TOL32
    assert w[1] < TOL32


SHAPES = [(8, 8, 64), (8, 8, 16), (8, 4, 32), (4, 8, 64), (4, 4, 16)]


def test_shared_unit_order_changes_no_result():
    # every launch of a pass dispatches its units in the order of the previous launch's unit run
    # times; each partial sum has one owner unit, so repeated right-hand sides on one context are
    # bitwise equal whatever order each launch ran in
    prob = small_problem("square", 4, 8, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=77)
    with prob.context(n_slots=3, precision=FP32_MIXED) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, A.Action.empty(ns, m))
        ctx.upload(2, A.Action.empty(ns, m))
        outs = []
        for _ in range(3):
            F.rhs(ctx, 1.0, 2, 0, 1, 10, da_l=2)
            outs.append(ctx.download(1))
    print(f"square L=4 sites {ns} N=8: three 2-loop right-hand sides in measured unit order", flush=True)
    for o in outs[1:]:
        assert np.array_equal(o.gamma, outs[0].gamma) and np.array_equal(o.Sigma, outs[0].Sigma)


def test_shared_shapes_agree():
    # every launch shape of the shared-node pass (waves per block, batch capacity, most nodes per
    # batch) computes the same integrals: within FP32 rounding of each other, within the mode's
    # bound of FP64
    prob = small_problem("square", 3, 8, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=61)
    ref = _rhs_modes(prob, a, 2, (FP64,))[FP64][0]
    outs = {}
    for shape in SHAPES:
        with prob.context(n_slots=3, precision=FP32_MIXED) as ctx:
            ctx.set_shared_shape(*shape)
            assert ctx.shared_shape() == shape
            ctx.upload(0, a)
            ctx.upload(1, A.Action.empty(ns, m))
            ctx.upload(2, A.Action.empty(ns, m))
            F.rhs(ctx, 1.0, 2, 0, 1, 10, da_l=2)
            outs[shape] = ctx.download(1)
    for shape, g in outs.items():
        e64 = max(per_array_err(g, ref).values())
        e0 = max(per_array_err(g, outs[SHAPES[0]]).values())
        print(f"shape {shape}: vs FP64 {e64:.2e}, vs default shape {e0:.2e}", flush=True)
        assert e64 < TOL32
        assert e0 < TOL_ELEM


def test_shared_shape_refused_and_tuned():
    prob = small_problem("square", 6, 6, 1.0)  # 16 sites: 4 waves would need 4 sites per wave
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=62)
    with prob.context(n_slots=3, precision=FP32_MIXED) as ctx:
        with pytest.raises(Exception):
            ctx.set_shared_shape(4, 8, 64)
        ctx.upload(0, a)
        ctx.upload(1, A.Action.empty(ns, m))
        best, ms = ctx.tune_shared(1.0, 0, 1, 10)
        print(f"tuned shape {best}, candidate ms {['%.2f' % x for x in ms]}", flush=True)
        assert best[0] == 8 and ctx.shared_shape() == best
        assert all(x < 0 for x in ms[5:]) and all(x > 0 for x in ms[:3])
        F.rhs(ctx, 1.0, 1, 0, 1, 10)
        tuned = ctx.download(1)
    with prob.context(n_slots=3, precision=FP32_MIXED) as ctx:
        ctx.set_shared_shape(*best)
        ctx.upload(0, a)
        ctx.upload(1, A.Action.empty(ns, m))
        F.rhs(ctx, 1.0, 1, 0, 1, 10)
        np.testing.assert_array_equal(ctx.download(1).gamma, tuned.gamma)


def test_batch_of_mixed_shapes_equals_single():
    # a batched launch runs every problem in its own shape (the block has the threads of the
    # widest): each problem's result equals its own run bitwise
    from pffrg.mesh import get_eval
    from pffrg.native import Batch
    specs = [("square", 3, 8, (4, 4, 16)), ("square", 5, 6, (8, 8, 64)), ("kagome", 3, 6, (4, 8, 32))]
    probs, acts, ctxs, single = [], [], [], []
    try:
        for k, (name, size, N, shape) in enumerate(specs):
            p = small_problem(name, size, N, 1.0)
            a = A.synthetic_action(p.r.n_sites, p.mesh, p.bare, seed=70 + k)
            c = p.context(n_slots=3, precision=FP32_MIXED)
            c.set_shared_shape(*shape)
            c.upload(0, a)
            F.rhs(c, 0.8, 1, 0, 1, get_eval(0.8))
            single.append(c.download(1))
            probs.append(p)
            ctxs.append(c)
        with Batch(ctxs) as b:
            b.rhs_1l([0.8] * len(ctxs), 0, 2, [get_eval(0.8)] * len(ctxs))
        for k, c in enumerate(ctxs):
            got = c.download(2)
            np.testing.assert_array_equal(got.gamma, single[k].gamma)
            np.testing.assert_array_equal(got.Sigma, single[k].Sigma)
        print("batched mixed shapes = single runs", flush=True)
    finally:
        for c in ctxs:
            c.close()
