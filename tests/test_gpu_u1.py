"""The U(1)-symmetric DM action (six vertex components, action_u1_dm.jl) on the GPU against the
FP64 oracle, through the C ABI: the triangular lattice with the C3 DM model
(model_triangular_dm_c3.jl), every pass of the flow and parquet right-hand sides, the
self-energy kernels, the symmetrisation and algebra kernels, the correlations χxx, χzz, χxy,
mesh rescanning and a short BS3 flow.

Tolerances as for SU(2): FP64 ≤1e-12 relative to the largest entry (bitwise where the GPU
restates the oracle's arithmetic: symmetrisation, χ cubature regions, rescanning); the
FP32-mixed mode (the per-element FP32 pass for this action) ≤1e-4 per q3 array."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import flow as F
from pffrg import mesh as M
from pffrg.native import FP32_MIXED, FP64

from helpers import rel_err
from oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu

QS = ("q1", "q2_1", "q2_2", "q3")


def _dm_problem(size=2, N=5, J=((1.0, 0.6, 0.4),), initial=1.0):
    return F.Problem.build("triangular", size, [list(j) for j in J], initial=initial, num_sigma=N + 2, num_Omega=N,
                           num_nu=N, normalize=False, model="triangular-dm-c3")


def _worst_q3(got, ref):
    return max(rel_err(got.view(c, ch, "q3"), ref.view(c, ch, "q3")) for c in range(6) for ch in range(3)
               if np.max(np.abs(ref.view(c, ch, "q3"))) > 0)


@pytest.mark.parametrize("size,loops", [(2, 1), (2, 2), (3, 3), (4, 1)])
def test_u1_rhs_matches_oracle(size, loops):
    prob = _dm_problem(size, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=50 + size)
    print(f"U(1)-DM triangular L={size} ({ns} sites) {loops}-loop right-hand side", flush=True)
    outs = {}
    for prec in (FP64, FP32_MIXED):
        with prob.context(n_slots=7, precision=prec) as ctx:
            ctx.upload(0, a)
            for k in range(1, 7):
                ctx.upload(k, A.Action.empty(ns, m, nc=6))
            F.rhs(ctx, 1.0, loops, 0, 1, 8, da_l=2, da_c=3, da_temp=4, da_sigma=5)
            outs[prec] = ctx.download(1)
    octx = OracleContext(prob, n_slots=7)
    octx.upload(0, a)
    F.rhs(octx, 1.0, loops, 0, 1, 8, da_l=2, da_c=3, da_temp=4, da_sigma=5)
    ref = octx.download(1)
    got = outs[FP64]
    eg, es = rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma)
    e32 = _worst_q3(outs[FP32_MIXED], got)
    print(f"   FP64 vertex {eg:.2e}, Σ {es:.2e}; FP32-mixed worst q3 {e32:.2e}", flush=True)
    assert eg < 1e-12 and es < 1e-12
    assert e32 < 1e-4


def test_u1_parquet_passes_match_oracle():
    prob = _dm_problem(2, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=61)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, a)
        ctx.upload(2, A.Action.empty(ns, m, nc=6))
        ctx.compute_sigma(1.0, 0, 1, 8)
        ctx.compute_gamma(1.0, 0, 2, 8)
        sde, bse = ctx.download(1), ctx.download(2)
    orc = O.Oracle(prob.r, m, nc=6)
    r1, r2 = a.copy(), A.Action.empty(ns, m, nc=6)
    orc.compute_sigma(1.0, a, r1, 8)
    orc.compute_gamma(1.0, a, r2, 8)
    e1 = max(rel_err(sde.Sigma, r1.Sigma), rel_err(sde.gamma, r1.gamma))
    e2 = rel_err(bse.gamma, r2.gamma)
    print(f"U(1)-DM SDE (reduced bubble + Σ) {e1:.2e}; BSE {e2:.2e}", flush=True)
    assert e1 < 1e-12 and e2 < 1e-12


def test_u1_integrand_probe_matches_oracle():
    # the integrand of single elements at chosen loop frequencies, every pass and channel
    prob = _dm_problem(3, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=62)
    d = A.synthetic_action(ns, m, prob.bare, seed=63)
    orc = O.Oracle(prob.r, m, nc=6)
    vs = np.array([-3.1, -0.7, 0.0, 0.45, 2.2, 9.0])
    worst = 0.0
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, d)
        for p in (O.P_KAT, O.P_LEFT, O.P_CENTRAL, O.P_BSE):
            for ch in range(3):
                for (k, w2, w3) in ((1, -1, -1), (2, 1, -1), (3, -1, 2), (4, 2, 1)):
                    got = ctx.probe_integrand(p, 1.0, 0, 1, ch, k, 2, w2, w3, vs)
                    for j, v in enumerate(vs):
                        ref = orc.integrand(p, 1.0, a, d, ch, k, 2, w2, w3, v)
                        s = max(float(np.max(np.abs(ref))), 1e-300)
                        worst = max(worst, float(np.max(np.abs(got[j] - ref))) / s)
    print(f"U(1)-DM integrand probe: worst {worst:.2e}", flush=True)
    assert worst < 1e-12


def test_u1_symmetrize_and_algebra_bitwise():
    prob = _dm_problem(3, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=64)
    b = A.synthetic_action(ns, m, prob.bare, seed=65)
    orc = O.Oracle(prob.r, m, nc=6)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, b)
        ctx.symmetrize(0)
        ctx.symmetrize_add_to(1, 0)
        got = ctx.download(0)
        gmax = ctx.abs_max(0)
        ctx.mult_with_add_to(1, -0.5, 0)
        ax = ctx.download(0)
    ref = a.copy()
    orc.symmetrize(ref)
    orc.symmetrize_add_to(b, ref)
    np.testing.assert_array_equal(got.gamma, ref.gamma)
    assert gmax == ref.abs_max()
    np.testing.assert_array_equal(ax.gamma, ref.gamma + -0.5 * b.gamma)


def test_u1_chi_matches_oracle_with_equal_regions():
    prob = _dm_problem(2, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=66)
    tol = (1e-7, 1e-3)
    with prob.context(n_slots=1) as ctx:
        ctx.upload(0, a)
        chi, nr = ctx.compute_chi(1.0, 0, tol, m.chi[:3], with_regions=True)
    orc = O.Oracle(prob.r, m, nc=6)
    ref, rnr = orc.compute_chi(1.0, a, tol, m.chi[:3], with_regions=True)
    e = max(rel_err(chi[c], ref[c]) for c in range(3))
    print(f"U(1)-DM χxx, χzz, χxy: rel err {e:.2e}; regions equal: {np.array_equal(nr, rnr)}", flush=True)
    assert chi.shape == (3, ns, 3)
    np.testing.assert_array_equal(nr, rnr)
    assert e < 1e-12


def test_u1_rescan_bitwise():
    prob = _dm_problem(2, 5)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=67)
    orc = O.Oracle(prob.r, m, nc=6)
    base = M.launch_mesh(0.7, m.num_sigma - 1, m.num_Omega - 1, m.num_nu - 1, m.num_chi - 1)
    m2 = M.Mesh(base.sigma * 1.2, base.Os * 1.2, base.ns * 0.9, base.Ot * 1.1, base.nt * 1.3, base.chi)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        idx, cuts = ctx.scan_cuts(0)
        ridx, rcuts = orc.scan_cuts(a)
        np.testing.assert_array_equal(idx, ridx)
        np.testing.assert_array_equal(cuts, rcuts)
        assert idx.shape == (18, 4)
        ctx.resample(0, 1, m2)
        got = ctx.download(1)
    ref = A.Action.empty(ns, m2, nc=6)
    orc.resample(a, m2, ref)
    np.testing.assert_array_equal(got.gamma, ref.gamma)
    np.testing.assert_array_equal(got.Sigma, ref.Sigma)
    np.testing.assert_array_equal(got.bare, ref.bare)


def _flow_pair(prob, **kw):
    a0 = prob.bare_action()
    with prob.context(n_slots=9) as ctx:
        res = F.launch_flow(ctx, prob, a0, 5.0, 1.0, max_steps=4, **kw)
        got = ctx.download(F.A)
    octx = OracleContext(prob)
    ref_res = F.launch_flow(octx, prob, a0, 5.0, 1.0, max_steps=4, **kw)
    ref = octx.download(F.A)
    print(f"   GPU Λ {res.Lambdas}; oracle Λ {ref_res.Lambdas}", flush=True)
    assert res.accepted == ref_res.accepted and res.rejected == ref_res.rejected
    np.testing.assert_allclose(res.Lambdas, ref_res.Lambdas, rtol=1e-12)
    e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
    ec = max(rel_err(c, r) for c, r in zip(res.chi, ref_res.chi))
    print(f"   final action {e:.2e}; χ along the flow {ec:.2e}", flush=True)
    assert e < 1e-10 and ec < 1e-10
    assert np.asarray(res.chi[-1]).shape[0] == 3
    return res, ref_res


def test_u1_flow_matches_oracle_flow():
    # launch_1l! for the U(1)-DM action (launcher_1l.jl:72-198) with the DM coupling, χ after
    # every step (three components; the monotonicity stop reads the largest), fixed meshes: from
    # bare DM couplings the zd and dz channels hold only rounding noise (their exact value is 0),
    # and the mesh scan (scan_channel, Action.jl:449-480) of such a channel follows the noise, so
    # rescanned meshes are not reproducible between two implementations that round differently
    print("U(1)-DM flow with DM coupling, fixed meshes", flush=True)
    _flow_pair(_dm_problem(2, 5, initial=5.0), rescan=False)


def test_u1_xxz_flow_with_rescanning_matches_oracle_flow():
    # the same model without DM (an XXZ model in the U(1)-DM action): DM, zd and dz stay exactly 0,
    # and the flow rescans its meshes after every step (scan over all 18 component channels)
    print("U(1)-DM flow, XXZ couplings, rescanned meshes", flush=True)
    res, ref_res = _flow_pair(_dm_problem(2, 5, J=((1.0, 0.6, 0.0),), initial=5.0))
    assert len(res.meshes) == len(ref_res.meshes) == res.accepted
    for mg, mr in zip(res.meshes, ref_res.meshes):
        for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
            np.testing.assert_allclose(getattr(mg, f), getattr(mr, f), rtol=1e-12)
