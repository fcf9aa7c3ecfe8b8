"""Parity at the CPU-runnable flow's size (square L=3, 6 sites, nearest-neighbour Heisenberg,
N=24 per axis, 1-loop, Λ from 50 to 0.3; launch! with launch_1l!, launcher_1l.jl:72-198) and at
the parquet loop's (kagome L=3, parquet.jl:1-100), every output element against the FP64 oracle:
 - one right-hand side (compute_dΣ! and compute_dΓ_1l!, compute.jl:57-112) at Λ = 50 and at
   Λ = 0.3, every vertex entry, FP64 within 1e-12 of the oracle (the oracle's vertex pass reads
   the GPU's dΣ, so that an ulp of difference there cannot move an adaptive decision), and the
   FP32-mixed form within 1e-4 of FP64 per array;
 - the first three BS3 steps of that flow (χ measured and meshes rescanned after every step)
   against the same driver on the oracle, as test_gpu_flow does at N=6;
 - five parquet iterations (compute_Σ! and compute_Γ! with damping) against the oracle."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import flow as F
from pffrg.native import FP32_MIXED, FP64

from helpers import per_array_err, rel_err
from oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu

N = 24


def _c1(initial):
    return F.Problem.build("square", 3, [[1.0]], initial=initial, num_sigma=N, num_Omega=N, num_nu=N, num_chi=10)


@pytest.mark.parametrize("Lam", [50.0, 0.3])
def test_c1_rhs_every_element_matches_oracle(Lam):
    prob = _c1(Lam)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=24)
    ev = 10 if Lam >= 1.0 else int(np.ceil(10 / np.sqrt(Lam)))
    print(f"square L=3 N={N} 1-loop RHS at Λ={Lam}, eval={ev}", flush=True)
    out = {}
    for mode in (FP64, FP32_MIXED):
        with prob.context(n_slots=2, precision=mode) as ctx:
            ctx.upload(0, a)
            ctx.upload(1, A.Action.empty(ns, m))
            F.rhs(ctx, Lam, 1, 0, 1, ev)
            out[mode] = ctx.download(1)
    got = out[FP64]
    orc = O.Oracle(prob.r, m)
    ref_s = A.Action.empty(ns, m)
    orc.compute_dsigma(Lam, a, ref_s)
    ref = A.Action.empty(ns, m)
    ref.Sigma[:] = got.Sigma
    orc.compute_dgamma_1l(Lam, a, ref, ev)
    es = rel_err(got.Sigma, ref_s.Sigma)
    eg = rel_err(got.gamma, ref.gamma)
    e32 = per_array_err(out[FP32_MIXED], got)
    w32 = max(e32.items(), key=lambda kv: kv[1])
    print(f"   dΣ {es:.2e}; every vertex entry {eg:.2e} ({got.gamma.size} entries); FP32-mixed worst "
          f"{w32[0]} {w32[1]:.2e}", flush=True)
    assert es < 1e-12 and eg < 1e-12
    assert w32[1] < 1e-4


def test_c1_flow_first_steps_match_oracle_flow():
    prob = _c1(50.0)
    a0 = prob.bare_action()
    with prob.context(n_slots=9) as ctx:
        res = F.launch_flow(ctx, prob, a0, 50.0, 0.3, max_steps=3)
        got = ctx.download(F.A)
    octx = OracleContext(prob)
    ref_res = F.launch_flow(octx, prob, a0, 50.0, 0.3, max_steps=3)
    ref = octx.download(F.A)
    print(f"square L=3 N={N} flow: GPU Λ {res.Lambdas} ({res.accepted}/{res.rejected}); oracle Λ {ref_res.Lambdas}",
          flush=True)
    assert res.accepted == ref_res.accepted == 3 and res.rejected == ref_res.rejected
    np.testing.assert_allclose(res.Lambdas, ref_res.Lambdas, rtol=1e-12)
    e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
    ec = max(rel_err(c, r) for c, r in zip(res.chi, ref_res.chi))
    print(f"   action after 3 steps {e:.2e}; χ along the flow {ec:.2e}", flush=True)
    assert e < 1e-10 and ec < 1e-10
    for mg, mr in zip(res.meshes, ref_res.meshes):
        for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
            np.testing.assert_allclose(getattr(mg, f), getattr(mr, f), rtol=1e-12)


def test_parquet_iterations_match_oracle():
    prob = F.Problem.build("kagome", 3, [[1.0]], initial=1.0, num_sigma=10, num_Omega=8, num_nu=8, num_chi=5)
    a0 = prob.bare_action()
    L = 1.0
    with prob.context(n_slots=9) as ctx:
        res = F.launch_parquet(ctx, L, a0, max_iter=5, parquet_tol=(1e-30, 1e-30))
        got = ctx.download(F.A)
    octx = OracleContext(prob)
    ref_res = F.launch_parquet(octx, L, a0, max_iter=5, parquet_tol=(1e-30, 1e-30))
    ref = octx.download(F.A)
    print(f"kagome L=3 parquet: GPU {res.iterations} iterations (SDE {res.sde_iterations}), errors "
          f"{[f'{e:.3e}' for e, _ in res.history]}; oracle SDE {ref_res.sde_iterations}", flush=True)
    assert res.iterations == ref_res.iterations == 5 and res.status == "max_iter"
    assert res.sde_iterations == ref_res.sde_iterations
    np.testing.assert_allclose([e for e, _ in res.history], [e for e, _ in ref_res.history], rtol=1e-9)
    e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
    print(f"   solution after 5 iterations {e:.2e}", flush=True)
    assert e < 1e-10
