"""Parity at the single-RHS bench size (square L=6, 16 sites, N=40, Λ=1, eval=10, the seeded
synthetic vertex): the GPU's compute_dΣ!, its Katanin pass and the left pass of the 2-loop
right-hand side against the FP64 oracle on strided samples of the output elements (with every
(channel, kernel) class at the ends of its Ω and ν ranges), and
FP32-mixed against FP64 on every element; at the pyrochlore 3-loop flow's shape (L=4, N=32),
the central pass of the multiloop right-hand side; at the warm-start shape (kagome L=5, N=32),
the parquet SDE and BSE passes; at the largest meshes the library takes (64 points per axis), the
Katanin pass; at each wavefront layout by site count (2 to 49 sites), the Katanin and left passes.

Tolerances: FP64 ≤1e-12 relative to the largest entry (the pass restates the oracle's arithmetic
and takes the same adaptive decisions; the oracle's elements read the GPU's dΣ so that they test
the vertex pass alone); FP32-mixed ≤1e-4 per q3 array, the path's stated tolerance."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import flow as F
from pffrg.native import FP32_MIXED

from helpers import class_sample, entries, full_elements, half_elements, rel_err

pytestmark = pytest.mark.gpu

# every 997th element of the half pass (117 over all channels and kernels), plus each (channel,
# kernel) class at the ends of its Ω and ν ranges
STRIDE = 997


def test_bench_size_katanin_rhs_matches_oracle_sample():
    prob = F.Problem.build("square", 6, [[1.0]], initial=1.0, num_sigma=40, num_Omega=40, num_nu=40, num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=20261203)
    outs = {}
    for prec in (0, FP32_MIXED):
        with prob.context(n_slots=2, precision=prec) as ctx:
            ctx.upload(0, a)
            ctx.compute_dsigma(1.0, 0, 1)
            ctx.compute_dgamma_1l(1.0, 0, 1, 10)
            outs[prec] = ctx.download(1)
    got = outs[0]
    orc = O.Oracle(prob.r, m)
    ref_s = A.Action.empty(ns, m)
    orc.compute_dsigma(1.0, a, ref_s)
    es = rel_err(got.Sigma, ref_s.Sigma)
    dd = A.Action.empty(ns, m)
    dd.Sigma[:] = got.Sigma
    orc.compute_corrs(1.0, a, dd, True)
    out = A.Action.empty(ns, m)
    elems = half_elements(m.num_Omega, m.num_nu)
    sample = class_sample(elems, m.num_Omega, m.num_nu, STRIDE)
    for e in sample:
        orc.compute_element(O.P_KAT, 1.0, a, dd, *e, 10, out)
    scale = float(np.max(np.abs(got.gamma)))
    ev = max(float(np.max(np.abs(entries(got, *e) - entries(out, *e)))) for e in sample) / scale
    e32 = max(rel_err(outs[FP32_MIXED].view(c, ch, "q3"), got.view(c, ch, "q3")) for c in range(2) for ch in range(3))
    print(f"square L=6 N=40: dSigma {es:.2e}; {len(sample)} sampled elements {ev:.2e}; FP32-mixed q3 {e32:.2e}",
          flush=True)
    assert es < 1e-12 and ev < 1e-12
    assert e32 < 1e-4


def test_bench_size_left_pass_matches_oracle_sample():
    # compute_dΓ_2l! (compute.jl:115-174): the left pass reads the action and its 1-loop
    # derivative and writes every element of slot da_l, which the oracle's elements match
    prob = F.Problem.build("square", 6, [[1.0]], initial=1.0, num_sigma=40, num_Omega=40, num_nu=40, num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=20261203)
    with prob.context(n_slots=4) as ctx:
        ctx.upload(0, a)
        ctx.compute_dsigma(1.0, 0, 1)
        ctx.compute_dgamma_1l(1.0, 0, 1, 10)
        da1 = ctx.download(1)
        ctx.compute_dsigma(1.0, 0, 2)
        ctx.compute_dgamma_2l(1.0, 0, 2, 3, 10)
        left = ctx.download(3)
    orc = O.Oracle(prob.r, m)
    orc.compute_corrs(1.0, a, a, False)
    out = A.Action.empty(ns, m)
    sample = class_sample(full_elements(m.num_Omega, m.num_nu), m.num_Omega, m.num_nu, STRIDE)
    for e in sample:
        orc.compute_element(O.P_LEFT, 1.0, a, da1, *e, 10, out)
    scale = float(np.max(np.abs(left.gamma)))
    ev = max(float(np.max(np.abs(entries(left, *e) - entries(out, *e)))) for e in sample) / scale
    print(f"square L=6 N=40 left pass: {len(sample)} sampled elements {ev:.2e}", flush=True)
    assert ev < 1e-12


def test_pyrochlore_flow_size_central_pass_matches_oracle_sample():
    # compute_dΓ_ml! (compute.jl:177-286) at the 3-loop flow's shape, pyrochlore L=4 (18 sites),
    # N=32: the central pass of loop 3 reads the 2-loop left pass and fills slot da_c (its half
    # of the elements, then symmetrized), which the oracle's elements match
    prob = F.Problem.build("pyrochlore", 4, [[1.0]], initial=1.0, num_sigma=32, num_Omega=32, num_nu=32, num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=20261203)
    with prob.context(n_slots=9) as ctx:
        ctx.upload(0, a)
        ctx.compute_dsigma(1.0, 0, 1)
        ctx.compute_dgamma_2l(1.0, 0, 1, 2, 10)
        left2 = ctx.download(2)
        ctx.compute_dsigma(1.0, 0, 3)
        ctx.compute_dgamma_ml(1.0, 3, 0, 3, 4, 5, 6, 7, 10)
        central = ctx.download(5)
    orc = O.Oracle(prob.r, m)
    orc.compute_corrs(1.0, a, a, False)
    out = A.Action.empty(ns, m)
    sample = class_sample(half_elements(m.num_Omega, m.num_nu), m.num_Omega, m.num_nu, 499)
    for e in sample:
        orc.compute_element(O.P_CENTRAL, 1.0, a, left2, *e, 10, out)
    scale = float(np.max(np.abs(central.gamma)))
    ev = max(float(np.max(np.abs(entries(central, *e) - entries(out, *e)))) for e in sample) / scale
    print(f"pyrochlore L=4 N=32 central pass: {len(sample)} sampled elements {ev:.2e}", flush=True)
    assert ev < 1e-12


def test_warmstart_size_parquet_passes_match_oracle():
    # compute_Σ! (SDE.jl:58-96) and compute_Γ! (compute.jl:2-54) at the warm-start shape,
    # kagome L=5 (19 sites), N=32: the reduced s bubble and Σ in full, the BSE pass on a sample
    prob = F.Problem.build("kagome", 5, [[1.0], [0.2]], initial=1.0, num_sigma=32, num_Omega=32, num_nu=32,
                           num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=20261203)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        ctx.compute_sigma(1.0, 0, 1, 10)
        sde = ctx.download(1)
        ctx.compute_gamma(1.0, 0, 2, 10)
        bse = ctx.download(2)
    orc = O.Oracle(prob.r, m)
    ref = A.Action.empty(ns, m)
    orc.compute_sigma(1.0, a, ref, 10)  # also leaves the tail factors the BSE pass uses
    e_sde = max(rel_err(sde.view(c, 0, q), ref.view(c, 0, q)) for c in range(2) for q in ("q1", "q2_2"))
    e_sig = rel_err(sde.Sigma, ref.Sigma)
    out = A.Action.empty(ns, m)
    sample = class_sample(half_elements(m.num_Omega, m.num_nu), m.num_Omega, m.num_nu, 499)
    for e in sample:
        orc.compute_element(O.P_BSE, 1.0, a, a, *e, 10, out)
    scale = float(np.max(np.abs(bse.gamma)))
    e_bse = max(float(np.max(np.abs(entries(bse, *e) - entries(out, *e)))) for e in sample) / scale
    print(f"kagome L=5 N=32: reduced bubble {e_sde:.2e}, SDE sigma {e_sig:.2e}, {len(sample)} BSE elements "
          f"{e_bse:.2e}", flush=True)
    assert e_sde < 1e-12 and e_sig < 1e-12 and e_bse < 1e-12


def test_largest_meshes_katanin_rhs_matches_oracle_sample():
    # the largest meshes the library takes: 64 points per axis (N=63), the LDS tables full, on
    # square L=3 (6 sites, 10 groups per wavefront): dΣ in full and a strided sample of the
    # Katanin pass against the oracle, FP32-mixed q3 against FP64
    prob = F.Problem.build("square", 3, [[1.0]], initial=1.0, num_sigma=63, num_Omega=63, num_nu=63, num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    assert len(m.Os) == len(m.ns) == len(m.sigma) == 64
    a = A.synthetic_action(ns, m, prob.bare, seed=20261204)
    outs = {}
    for prec in (0, FP32_MIXED):
        with prob.context(n_slots=2, precision=prec) as ctx:
            ctx.upload(0, a)
            ctx.compute_dsigma(0.5, 0, 1)
            ctx.compute_dgamma_1l(0.5, 0, 1, 10)
            outs[prec] = ctx.download(1)
    got = outs[0]
    orc = O.Oracle(prob.r, m)
    ref_s = A.Action.empty(ns, m)
    orc.compute_dsigma(0.5, a, ref_s)
    es = rel_err(got.Sigma, ref_s.Sigma)
    dd = A.Action.empty(ns, m)
    dd.Sigma[:] = got.Sigma
    orc.compute_corrs(0.5, a, dd, True)
    out = A.Action.empty(ns, m)
    sample = half_elements(m.num_Omega, m.num_nu)[::2999]
    for e in sample:
        orc.compute_element(O.P_KAT, 0.5, a, dd, *e, 10, out)
    scale = float(np.max(np.abs(got.gamma)))
    ev = max(float(np.max(np.abs(entries(got, *e) - entries(out, *e)))) for e in sample) / scale
    e32 = max(rel_err(outs[FP32_MIXED].view(c, ch, "q3"), got.view(c, ch, "q3")) for c in range(2) for ch in range(3))
    print(f"square L=3 N=63 (64 points): dSigma {es:.2e}; {len(sample)} sampled elements {ev:.2e}; "
          f"FP32-mixed q3 {e32:.2e}", flush=True)
    assert es < 1e-12 and ev < 1e-12
    assert e32 < 1e-4


@pytest.mark.parametrize("name,size,groups", [("square", 1, 16), ("square", 9, 2), ("pyrochlore", 5, 2),
                                               ("square", 12, 1), ("pyrochlore", 6, 1)])
def test_site_layouts_match_oracle_sample(name, size, groups):
    # the wavefront layouts by site count: 2 sites (16 groups of 2 lanes, half the wavefront
    # idle), 30 and 31 sites (2 groups), 49 and 48 sites (one group with 15 or 16 idle lanes,
    # overlap tables of 7000-8000 rows in LDS).  dΣ in full, sampled Katanin elements and
    # (square) sampled left-pass elements against the oracle, FP32-mixed q3 against FP64
    prob = F.Problem.build(name, size, [[1.0]], initial=1.0, num_sigma=10, num_Omega=8, num_nu=8, num_chi=10)
    ns, m = prob.r.n_sites, prob.mesh
    assert min(64 // ns, 16) == groups
    a = A.synthetic_action(ns, m, prob.bare, seed=20261205)
    outs = {}
    for prec in (0, FP32_MIXED):
        with prob.context(n_slots=4, precision=prec) as ctx:
            ctx.upload(0, a)
            ctx.compute_dsigma(1.0, 0, 1)
            ctx.compute_dgamma_1l(1.0, 0, 1, 10)
            kat = ctx.download(1)
            ctx.compute_dsigma(1.0, 0, 2)
            ctx.compute_dgamma_2l(1.0, 0, 2, 3, 10)
            outs[prec] = (kat, ctx.download(3))
    got, left = outs[0]
    orc = O.Oracle(prob.r, m)
    ref_s = A.Action.empty(ns, m)
    orc.compute_dsigma(1.0, a, ref_s)
    es = rel_err(got.Sigma, ref_s.Sigma)
    dd = A.Action.empty(ns, m)
    dd.Sigma[:] = got.Sigma
    orc.compute_corrs(1.0, a, dd, True)
    out = A.Action.empty(ns, m)
    sample = half_elements(m.num_Omega, m.num_nu)[::11]
    for e in sample:
        orc.compute_element(O.P_KAT, 1.0, a, dd, *e, 10, out)
    ev = max(float(np.max(np.abs(entries(got, *e) - entries(out, *e)))) for e in sample) / float(np.max(np.abs(got.gamma)))
    el = 0.0
    if name == "square":
        orc.compute_corrs(1.0, a, a, False)
        out_l = A.Action.empty(ns, m)
        sample_l = full_elements(m.num_Omega, m.num_nu)[::23]
        for e in sample_l:
            orc.compute_element(O.P_LEFT, 1.0, a, got, *e, 10, out_l)
        el = max(float(np.max(np.abs(entries(left, *e) - entries(out_l, *e)))) for e in sample_l) / \
            float(np.max(np.abs(left.gamma)))
    e32 = max(rel_err(outs[FP32_MIXED][0].view(c, ch, "q3"), got.view(c, ch, "q3")) for c in range(2) for ch in range(3))
    print(f"{name} L={size} ({ns} sites): dSigma {es:.2e}; {len(sample)} Katanin elements {ev:.2e}; left {el:.2e}; "
          f"FP32-mixed q3 {e32:.2e}", flush=True)
    assert es < 1e-12 and ev < 1e-12 and el < 1e-12
    assert e32 < 1e-4
