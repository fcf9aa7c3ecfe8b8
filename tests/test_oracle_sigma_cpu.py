"""The oracle's self-energy integrals against scipy: compute_dΣ_kernel (flow_su2_self.jl:2-33),
its corrections compute_dΣ_kernel_corr1/corr2 (flow_su2_self.jl:36-101) and compute_Σ_kernel
(SDE_su2.jl:45-77), restated here on top of the oracle's vertex evaluation get_Γ_avx!
(Action.jl:104-167; pinned by the interpolation tests of test_oracle_cpu) and integrated with
scipy's adaptive quadrature over the reference's intervals (-∞, -2Λ, 0, 2Λ, ∞).  The oracle runs
its QuadGK restatement at a tight tolerance (1e-12, 1e-10), so the two agree to the quadrature
error of the integrands' jumps (measured up to 6e-8), tested at 5e-7 relative.  A wrong frequency argument, site weight
or propagator factor in the oracle's kernels moves the result by far more (planted in
tests/test_oracle_mutations)."""
import warnings

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from pffrg import action as A

from helpers import small_problem

TIGHT = (1e-12, 1e-10)
TOL = 5e-7


def _setup(seed):
    prob = small_problem("square", 3, 5, 1.0)
    a = A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=seed)
    a.Sigma[:] = -0.3 * prob.mesh.sigma / (1.0 + prob.mesh.sigma ** 2)
    return prob, a, O.Oracle(prob.r, prob.mesh)


def _gamma(orc, a, s, t, u, mask=7):
    return orc.get_gamma_avx(a, [*s, *t, *u], mask)


def _sigma_loop(orc, a, w, v, mult, S, mask_loc=7, mask_nl=7):
    # local vertex on site 0 (buffers b2*) and non-local density on every site (buffers b1*)
    loc = _gamma(orc, a, (v + w, 0.5 * (-v + w), 0.5 * (-v + w)), (v - w, 0.5 * (v + w), 0.5 * (v + w)), (0.0, w, v),
                 mask_loc)
    nl = _gamma(orc, a, (v + w, 0.5 * (-v + w), 0.5 * (v - w)), (0.0, w, v), (-v + w, 0.5 * (v + w), 0.5 * (v + w)),
                mask_nl)
    val = 3.0 * loc[0, 0] + loc[1, 0]
    for j in range(len(mult)):
        val -= 2.0 * mult[j] * (2.0 * S) * nl[1, j]
    return val


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        yield


def _quad(f, L):
    pts = [-np.inf, -2.0 * L, 0.0, 2.0 * L, np.inf]
    tot = 0.0
    for lo, hi in zip(pts[:-1], pts[1:]):
        val, _ = integrate.quad(f, lo, hi, epsabs=1e-13, epsrel=1e-11, limit=2000)
        tot += val
    return tot


def _check(got, ref):
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"rel err {err:.2e} (max |Σ| {np.max(np.abs(ref)):.3e})", flush=True)
    assert err < TOL


def test_dsigma_kernel_against_scipy():
    prob, a, orc = _setup(41)
    L, m = 0.8, prob.mesh
    mult = np.asarray(prob.r.mult, dtype=float)
    da = A.Action.empty(a.ns, m)
    orc.compute_dsigma(L, a, da, TIGHT)
    ref = np.zeros(m.num_sigma)
    for i in range(1, m.num_sigma, 2):
        w = m.sigma[i]
        ref[i] = _quad(lambda v: _sigma_loop(orc, a, w, v, mult, a.S) * orc.S(L, v, a.Sigma) / (2.0 * np.pi), L)
    _check(da.Sigma[1::2], ref[1::2])


def test_dsigma_corrections_against_scipy():
    prob, a, orc = _setup(43)
    L, m = 0.8, prob.mesh
    mult = np.asarray(prob.r.mult, dtype=float)
    das = A.synthetic_action(a.ns, m, prob.bare, seed=44)
    da = A.Action.empty(a.ns, m)
    d2 = das.copy()
    orc.compute_dsigma_corr(L, a, da, d2, TIGHT)
    # corr1: da_Σ's vertex without u locally and without t non-locally, times -G
    ref1 = np.zeros(m.num_sigma)
    ref2 = np.zeros(m.num_sigma)
    for i in range(1, m.num_sigma, 2):
        w = m.sigma[i]
        ref1[i] = _quad(lambda v: _sigma_loop(orc, das, w, v, mult, a.S, 3, 5) * -orc.G(L, v, a.Sigma) / (2.0 * np.pi), L)
    _check(d2.Sigma[1::2], ref1[1::2])
    # corr2: a's full vertex times G² Σ_daΣ, with Σ_daΣ = the first correction just computed
    for i in range(1, m.num_sigma, 2):
        w = m.sigma[i]
        ref2[i] = _quad(lambda v: _sigma_loop(orc, a, w, v, mult, a.S) * orc.G(L, v, a.Sigma) ** 2
                        * orc.get_sigma(d2.Sigma, v) / (2.0 * np.pi), L)
    _check(da.Sigma[1::2], (d2.Sigma + ref2)[1::2])


def test_sde_sigma_kernel_against_scipy():
    prob, a, orc = _setup(45)
    L, m = 0.8, prob.mesh
    mult = np.asarray(prob.r.mult, dtype=float)
    a2 = A.synthetic_action(a.ns, m, prob.bare, seed=46)
    orc.compute_sigma(L, a, a2, 10, gtol=(1e-8, 1e-4), stol=TIGHT)
    inf = np.inf

    def kern(w, v):
        # s buffers only, the fermionic argument of the s channel at infinity (SDE_su2.jl:54-60)
        loc = _gamma(orc, a2, (v + w, inf, 0.5 * (-v + w)), (inf, inf, inf), (inf, inf, inf), 1)
        nl = _gamma(orc, a2, (v + w, inf, 0.5 * (v - w)), (inf, inf, inf), (inf, inf, inf), 1)
        val = 3.0 * loc[0, 0] + loc[1, 0]
        for j in range(len(mult)):
            val -= 2.0 * mult[j] * (2.0 * a2.S) * nl[1, j]
        return val * -orc.G(L, v, a.Sigma) / (2.0 * np.pi)

    ref = np.zeros(m.num_sigma)
    for i in range(1, m.num_sigma, 2):
        ref[i] = _quad(lambda v: kern(m.sigma[i], v), L)
    _check(a2.Sigma[1::2], ref[1::2])
