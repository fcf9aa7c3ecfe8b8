"""Mesh rescanning on the GPU (pffrg_scan_cuts, pffrg_resample; Action.jl:449-557,
channel.jl:416-470) against the FP64 oracle's restatement.

Tolerance: the argmax positions and cuts are exact (comparisons and one subtraction); the
resampled action is bit-identical (same weights, same sequence of products and sums, no FMA
contraction); a right-hand side on the new meshes agrees with the oracle's to 1e-12, like
every pass."""
import numpy as np
import pytest

import oracle as O
from pffrg import flow as F
from pffrg import mesh as M
from pffrg import native
from pffrg.action import Action

from helpers import rel_err, small_problem

pytestmark = pytest.mark.gpu


def _random_action(prob, seed):
    rng = np.random.default_rng(seed)
    a = Action.empty(prob.r.n_sites, prob.mesh, prob.S)
    a.gamma[:] = rng.normal(size=a.gamma.shape) / (1.0 + np.arange(a.gamma.size) % 7)
    a.Sigma[:] = -0.4 * prob.mesh.sigma / (1.0 + prob.mesh.sigma ** 2)
    a.bare[:] = prob.bare
    return a


@pytest.mark.parametrize("name,size,N", [("square", 3, 8), ("kagome", 2, 10)])
def test_scan_cuts_match_oracle(name, size, N):
    prob = small_problem(name, size, N, 1.0)
    a = _random_action(prob, 21)
    q3 = a.view(0, 1, "q3")
    big = float(np.max(np.abs(a.gamma))) + 1.0
    q3[1, 2, 3, 4] = big   # a tie: the earlier column-major position wins
    q3[0, 2, 3, 5] = -big
    print(f"scan_cuts {name} L={size} N={N}", flush=True)
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        idx, cuts = ctx.scan_cuts(0)
    ridx, rcuts = O.Oracle(prob.r, prob.mesh).scan_cuts(a)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(cuts, rcuts)
    assert tuple(idx[1]) == (1, 2, 3, 4)


@pytest.mark.parametrize("precision", [native.FP64, native.FP32_MIXED])
def test_resample_matches_oracle_and_next_rhs(precision):
    prob = small_problem("square", 3, 8, 1.0)
    a = _random_action(prob, 4)
    m_old = prob.mesh
    base = M.launch_mesh(0.7, m_old.num_sigma - 1, m_old.num_Omega - 1, m_old.num_nu - 1, m_old.num_chi - 1)
    m_new = M.Mesh(base.sigma * 1.2, base.Os * 1.2, base.ns * 0.9, base.Ot * 1.1, base.nt * 1.3, base.chi)
    print(f"resample square L=3 N=8 precision={precision}", flush=True)
    ref = Action.empty(prob.r.n_sites, m_new, prob.S)
    O.Oracle(prob.r, m_old).resample(a, m_new, ref)
    with prob.context(n_slots=3, precision=precision) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, prob.bare_action())
        F.rhs(ctx, 1.0, 1, 1, 2, 10)  # builds the gather copies and tables on the old meshes
        ctx.resample(0, 1, m_new)
        assert ctx.m is m_new
        got = ctx.download(1)
        assert np.array_equal(got.gamma, ref.gamma)
        assert np.array_equal(got.Sigma, ref.Sigma)
        assert np.array_equal(got.bare, ref.bare)
        # the next right-hand side runs on the new meshes
        F.rhs(ctx, 0.8, 1, 1, 2, 10)
        da = ctx.download(2)
    if precision == native.FP64:
        orc = O.Oracle(prob.r, m_new)
        rda = Action.empty(prob.r.n_sites, m_new, prob.S)
        orc.compute_dsigma(0.8, ref, rda)
        orc.compute_dgamma_1l(0.8, ref, rda, 10)
        e = max(rel_err(da.gamma, rda.gamma), rel_err(da.Sigma, rda.Sigma))
        print(f"RHS on the new meshes vs oracle: {e:.3e}", flush=True)
        assert e < 1e-12


def test_resample_identity_and_errors():
    prob = small_problem("square", 2, 6, 1.0)
    a = _random_action(prob, 9)
    a.Sigma[0] = 0.0
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.resample(0, 1, prob.mesh)
        got = ctx.download(1)
        assert np.array_equal(got.gamma, a.gamma) and np.array_equal(got.Sigma, a.Sigma)
        with pytest.raises(native.PffrgError):
            ctx.resample(0, 0, prob.mesh)
        bad = M.Mesh(prob.mesh.sigma, prob.mesh.Os[::-1].copy(), prob.mesh.ns, prob.mesh.Ot, prob.mesh.nt,
                     prob.mesh.chi)
        with pytest.raises(native.PffrgError):
            ctx.resample(0, 1, bad)
        assert ctx.m is prob.mesh
