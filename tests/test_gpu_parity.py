"""GPU parity: the HIP library through the C ABI against the FP64 CPU oracle.

Tolerances
 - FP64 mode: the kernels restate the oracle's arithmetic in the same order and
   take the same adaptive decisions (node counts are compared exactly), so results
   agree to rounding: rel. max-norm error <= 1e-12 per array (observed <= 6e-15).
 - FP32-mixed mode: FP32 vertex copy, interpolation and products with FP64
   quadrature; the target of the path is <= 1e-4 relative error per array
   against FP64, tested here with that bound.
"""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import flow as F
from pffrg.native import FP32_MIXED

from helpers import per_array_err, rel_err, small_problem

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-4


def _setup(name="square", size=2, N=6, seed=3, L=1.0, initial=1.0):
    prob = small_problem(name, size, N, initial)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=seed)
    return prob, ns, m, a


def _empty(ns, m):
    return A.Action.empty(ns, m)


def _print(tag, errs):
    worst = max(errs.items(), key=lambda kv: kv[1])
    print(f"{tag}: worst {worst[0]} {worst[1]:.3e}", flush=True)


@pytest.mark.parametrize("name,size", [("square", 2), ("kagome", 3), ("pyrochlore", 3)])
def test_dsigma(name, size):
    prob, ns, m, a = _setup(name, size)
    orc = O.Oracle(prob.r, m)
    ref = _empty(ns, m)
    orc.compute_dsigma(1.0, a, ref)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.compute_dsigma(1.0, 0, 1)
        got = ctx.get_sigma(1)
    e = rel_err(got, ref.Sigma)
    print(f"dsigma {name}: {e:.3e}", flush=True)
    assert e < TOL64


@pytest.mark.parametrize("name,size,L", [("square", 2, 1.0), ("kagome", 3, 1.0), ("pyrochlore", 3, 0.7),
                                         ("square", 3, 0.3)])
def test_rhs_1l(name, size, L):
    prob, ns, m, a = _setup(name, size, N=6)
    ev = F.M.get_eval(L)
    orc = O.Oracle(prob.r, m)
    ref = _empty(ns, m)
    orc.compute_dsigma(L, a, ref)
    nodes_ref = orc.compute_dgamma_1l(L, a, ref, ev)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.reset_stats()
        F.rhs(ctx, L, 1, 0, 1, ev)
        st = ctx.stats()
        got = ctx.download(1)
    errs = per_array_err(got, ref)
    _print(f"1l {name} L={L}", errs)
    print(f"   nodes gpu {st['nodes']} oracle {nodes_ref}", flush=True)
    assert rel_err(got.Sigma, ref.Sigma) < TOL64
    assert max(errs.values()) < TOL64
    # same adaptive decisions -> same node count (symmetrisation excluded on both sides)
    assert st["nodes"] == nodes_ref


@pytest.mark.parametrize("model,J", [("breathing", [[1.0, 0.6]]), ("pyrochlore-breathing-c3", [[1.0, 0.6, 0.1, 0.05]])])
def test_rhs_1l_breathing_models(model, J):
    # lattice tables of the lower-symmetry SU(2) models (model_breathing.jl) through the same passes
    name = "kagome" if model == "breathing" else "pyrochlore"
    prob = F.Problem.build(name, 2, J, initial=1.0, num_sigma=8, num_Omega=6, num_nu=6, normalize=False, model=model)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=9)
    orc = O.Oracle(prob.r, m)
    ref = _empty(ns, m)
    orc.compute_dsigma(1.0, a, ref)
    nodes_ref = orc.compute_dgamma_1l(1.0, a, ref, 10)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.reset_stats()
        F.rhs(ctx, 1.0, 1, 0, 1, 10)
        st = ctx.stats()
        got = ctx.download(1)
    errs = per_array_err(got, ref)
    _print(f"1l {model} {name} L=2 ({ns} sites)", errs)
    assert rel_err(got.Sigma, ref.Sigma) < TOL64 and max(errs.values()) < TOL64
    assert st["nodes"] == nodes_ref


@pytest.mark.parametrize("name,size", [("square", 2), ("kagome", 3)])
def test_rhs_2l(name, size):
    prob, ns, m, a = _setup(name, size, N=6, seed=5)
    orc = O.Oracle(prob.r, m)
    ref, ref_l = _empty(ns, m), _empty(ns, m)
    orc.compute_dsigma(1.0, a, ref)
    nodes_ref = orc.compute_dgamma_2l(1.0, a, ref, ref_l, 10)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.upload(2, _empty(ns, m))
        ctx.reset_stats()
        F.rhs(ctx, 1.0, 2, 0, 1, 10, da_l=2)
        st = ctx.stats()
        got, got_l = ctx.download(1), ctx.download(2)
    assert st["nodes"] == nodes_ref
    e, el = per_array_err(got, ref), per_array_err(got_l, ref_l)
    _print(f"2l {name} da", e)
    _print(f"2l {name} da_l", el)
    assert max(e.values()) < TOL64 and max(el.values()) < TOL64


def test_rhs_ml_sigma_corr():
    prob, ns, m, a = _setup("kagome", 3, N=5, seed=9)
    orc = O.Oracle(prob.r, m)
    acts = [_empty(ns, m) for _ in range(5)]
    orc.compute_dsigma(1.0, a, acts[0])
    orc.compute_dgamma_ml(1.0, 3, a, *acts, 10)
    orc.compute_dsigma_corr(1.0, a, acts[0], acts[4])
    with prob.context(n_slots=6) as ctx:
        ctx.upload(0, a)
        for s in range(2, 6):
            ctx.upload(s, _empty(ns, m))
        F.rhs(ctx, 1.0, 3, 0, 1, 10, da_l=2, da_c=3, da_temp=4, da_sigma=5)
        got = [ctx.download(s) for s in range(1, 6)]
    for k, (g, r) in enumerate(zip(got, acts)):
        e = per_array_err(g, r)
        _print(f"ml slot {k}", e)
        assert max(e.values()) < TOL64
    assert rel_err(got[0].Sigma, acts[0].Sigma) < TOL64
    assert rel_err(got[4].Sigma, acts[4].Sigma) < TOL64


@pytest.mark.parametrize("name,size", [("square", 2), ("pyrochlore", 3)])
def test_parquet_sde_bse(name, size):
    prob, ns, m, a = _setup(name, size, N=5, seed=13)
    orc = O.Oracle(prob.r, m)
    r2 = _empty(ns, m)
    orc.compute_sigma(1.0, a, r2, 10)
    r3 = _empty(ns, m)
    orc.compute_gamma(1.0, a, r3, 10)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.compute_sigma(1.0, 0, 1, 10)
        g2 = ctx.download(1)
        ctx.compute_gamma(1.0, 0, 2, 10)
        g3 = ctx.download(2)
    e2 = max(rel_err(g2.view(c, 0, q), r2.view(c, 0, q)) for c in range(2) for q in ("q1", "q2_2"))
    print(f"SDE {name}: bubble {e2:.3e} sigma {rel_err(g2.Sigma, r2.Sigma):.3e}", flush=True)
    assert e2 < TOL64 and rel_err(g2.Sigma, r2.Sigma) < TOL64
    e3 = per_array_err(g3, r3)
    _print(f"BSE {name}", e3)
    assert max(e3.values()) < TOL64


def test_skipped_log_domains():
    """Λ far above the mesh scale: 2(Λ+Ω/2) >= Ω_max+ν_max for every element, so the
    logarithmic domains are empty (quadrature.jl:128-138 `if b > a`)."""
    prob, ns, m, a = _setup("square", 2, N=5, seed=17)
    L = 400.0
    orc = O.Oracle(prob.r, m)
    ref = _empty(ns, m)
    orc.compute_dsigma(L, a, ref)
    orc.compute_dgamma_1l(L, a, ref, 10)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        F.rhs(ctx, L, 1, 0, 1, 10)
        got = ctx.download(1)
    e = per_array_err(got, ref)
    _print("1l L=400", e)
    assert max(e.values()) < TOL64


@pytest.mark.parametrize("loops", [1, 2])
def test_fp32_mixed(loops):
    prob, ns, m, a = _setup("kagome", 3, N=6, seed=21)
    orc = O.Oracle(prob.r, m)
    ref, ref_l = _empty(ns, m), _empty(ns, m)
    orc.compute_dsigma(1.0, a, ref)
    if loops == 1:
        orc.compute_dgamma_1l(1.0, a, ref, 10)
    else:
        orc.compute_dgamma_2l(1.0, a, ref, ref_l, 10)
    with prob.context(n_slots=3, precision=FP32_MIXED) as ctx:
        ctx.upload(0, a)
        ctx.upload(2, _empty(ns, m))
        F.rhs(ctx, 1.0, loops, 0, 1, 10, da_l=2)
        got = ctx.download(1)
    e = per_array_err(got, ref)
    _print(f"fp32 {loops}l", e)
    assert max(e.values()) < TOL32


def test_algebra():
    prob, ns, m, a = _setup("kagome", 3, N=5, seed=23)
    b = A.synthetic_action(ns, m, prob.bare, seed=24)
    orc = O.Oracle(prob.r, m)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, b)
        ctx.mult_with_add_to(0, -0.37, 1)
        g = ctx.download(1)
        np.testing.assert_allclose(g.gamma, b.gamma + (-0.37) * a.gamma, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(g.Sigma, b.Sigma + (-0.37) * a.Sigma, rtol=1e-15, atol=1e-15)
        assert ctx.abs_max(1) == g.abs_max()
        ctx.mult_with(1, 0.5, vertex_only=True)
        g2 = ctx.download(1)
        np.testing.assert_array_equal(g2.gamma, g.gamma * 0.5)
        np.testing.assert_array_equal(g2.Sigma, g.Sigma)
        ctx.replace_with(2, 0)
        g3 = ctx.download(2)
        np.testing.assert_array_equal(g3.gamma, a.gamma)
        # symmetrize! and symmetrize_add_to! against the oracle
        ctx.symmetrize(2)
        s_ref = a.copy()
        orc.symmetrize(s_ref)
        np.testing.assert_array_equal(ctx.download(2).gamma, s_ref.gamma)
        ctx.upload(2, b)
        ctx.symmetrize_add_to(0, 2)
        s2 = b.copy()
        orc.symmetrize_add_to(a, s2)
        np.testing.assert_array_equal(ctx.download(2).gamma, s2.gamma)
