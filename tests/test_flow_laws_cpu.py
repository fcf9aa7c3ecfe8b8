"""The host flow drivers against laws the mathematics fixes, on a fake context whose
right-hand sides are chosen in closed form (no vertex code runs).

BS3 (launcher_1l.jl:72-198): the Bogacki-Shampine tableau integrates a right-hand side that is
a polynomial of degree <= 2 in Λ exactly, whatever the steps; for dΓ/dΛ = c Λ² its embedded
second-order solution is off by exactly c dΛ³/24 (the weights (7/24, 1/4, 1/3, 1/8) at the nodes
(0, 1/2, 3/4, 1) integrate s² to 3/8 instead of 1/3), so the error norm, the acceptance and the
step law dΛ' = max(bmin, min(bmax Λ, 0.85 err^(-1/3) dΛ)) with the checkpoint clamp
(launcher_1l.jl:113-198) can be predicted for every attempted step.

Parquet (parquet.jl:43-84): for affine maps Σ -> m_s Σ + c_s and Γ -> m_g Γ + c_g the damped
iteration x' = (1 - β) x + β f(x), β = min(Λ, 1), contracts the distance to the fixed point
c / (1 - m) by exactly r = 1 - β (1 - m) per iteration, so successive parquet errors have the
ratio r_g; the loop ends on the first error below either tolerance, or at max_iter."""
import numpy as np
import pytest

from pffrg import action as A
from pffrg import flow as F

from helpers import small_problem


class FakeContext:
    """The slot API of pffrg.native.Context on host actions; the right-hand side is dΓ/dΛ =
    rhs(Λ) for every vertex entry and dΣ/dΛ = 0, the parquet maps are affine."""

    def __init__(self, prob, rhs=None, maps=None):
        self.m, self.ns, self.S = prob.mesh, prob.r.n_sites, prob.S
        self.slots = {}
        self.rhs_fn, self.maps = rhs, maps
        self.calls = []

    def set_mesh(self, m):
        self.m = m

    def upload(self, s, a):
        self.slots[s] = a.copy()

    def download(self, s):
        return self.slots[s].copy()

    def replace_with(self, dst, src, vertex_only=False):
        self.slots[dst].gamma[:] = self.slots[src].gamma
        if not vertex_only:
            self.slots[dst].Sigma[:] = self.slots[src].Sigma

    def mult_with_add_to(self, src, fac, dst, vertex_only=False):
        self.slots[dst].gamma[:] += fac * self.slots[src].gamma
        if not vertex_only:
            self.slots[dst].Sigma[:] += fac * self.slots[src].Sigma

    def mult_with(self, s, fac, vertex_only=False):
        self.slots[s].gamma[:] *= fac
        if not vertex_only:
            self.slots[s].Sigma[:] *= fac

    def abs_max(self, s) -> float:
        return float(np.max(np.abs(self.slots[s].gamma)))

    def get_sigma(self, s):
        return self.slots[s].Sigma.copy()

    def set_sigma(self, s, sig):
        self.slots[s].Sigma[:] = sig

    def synchronize(self):
        pass

    def compute_dsigma(self, L, a, da, tol=None):
        self.slots[da].Sigma[:] = 0.0

    def compute_dgamma_1l(self, L, a, da, eval_, tol=None):
        self.calls.append(L)
        self.slots[da].gamma[:] = self.rhs_fn(L)

    def compute_sigma(self, L, a, ap, eval_, gtol=None, stol=None):
        ms, cs = self.maps[0]
        self.slots[ap].Sigma[:] = ms * self.slots[a].Sigma + cs

    def compute_gamma(self, L, a, ap, eval_, tol=None):
        mg, cg = self.maps[1]
        self.slots[ap].gamma[:] = mg * self.slots[a].gamma + cg


def _prob():
    return small_problem("square", 2, 4, 5.0)


def _run(flow):
    while flow.begin_step():
        for k in range(4):
            flow.before(k)
            flow.rhs(k)
            flow.after(k)
        flow.finish_step_a()
        flow.finish_step_b(None)
    return flow.result()


class _WrongWeight(F.BS3Flow):
    # a planted mistake: the third stage weighted 5/9 instead of 4/9
    def after(self, k):
        if k == 2:
            self.ctx.mult_with_add_to(F.A_STAGE, -5.0 * self.dL / 9.0, F.DA)
            self.ctx.mult_with_add_to(F.A_STAGE, -1.0 * self.dL / 3.0, F.A_ERR)
        else:
            super().after(k)


@pytest.mark.parametrize("cls,exact", [(F.BS3Flow, True), (_WrongWeight, False)])
def test_bs3_integrates_quadratic_rhs_exactly(cls, exact):
    prob = _prob()
    rng = np.random.default_rng(5)
    a0 = prob.bare_action()
    n = a0.gamma.size
    c0, c1, c2 = (rng.standard_normal(n) for _ in range(3))
    a0.gamma[:] = rng.standard_normal(n)
    ctx = FakeContext(prob, rhs=lambda L: c0 + c1 * L + c2 * L * L)
    res = _run(cls(ctx, prob, a0, 5.0, 1.0, chi_tol=None, rescan=False, ODE_tol=(1e-8, 1e-3)))
    Lend = res.Lambdas[-1] if res.Lambdas else 5.0
    F_ = lambda L: c0 * L + c1 * L ** 2 / 2 + c2 * L ** 3 / 3
    want = a0.gamma + F_(Lend) - F_(5.0)
    err = float(np.max(np.abs(ctx.download(F.A).gamma - want)) / np.max(np.abs(want)))
    print(f"{cls.__name__}: {res.status}, {res.accepted} steps ({res.rejected} rejected) to Λ={Lend}; "
          f"rel err {err:.2e}", flush=True)
    if exact:
        assert res.status == "finished" and Lend == 1.0 and res.accepted >= 5
        assert err < 1e-13
    else:  # the embedded error flags the flow as unstable, or the result is off
        assert res.status != "finished" or err > 1e-6


def _attempts(calls):
    # every attempted step evaluates its stages at L, L - dL/2, L - 3dL/4, L - dL
    assert len(calls) % 4 == 0
    return [(calls[i], calls[i] - calls[i + 3]) for i in range(0, len(calls), 4)]


@pytest.mark.parametrize("bmin,cps", [(1e-4, (3.7, 2.2)), (0.45, ())])
def test_bs3_step_law_from_closed_form_error(bmin, cps):
    prob = _prob()
    c, g0, Li, Lf, tol, bmax = 3.0, 1.0, 5.0, 1.0, (1e-8, 1e-4), 0.1
    a0 = prob.bare_action()
    a0.gamma[:] = g0
    ctx = FakeContext(prob, rhs=lambda L: np.full(a0.gamma.size, c * L * L))
    res = _run(F.BS3Flow(ctx, prob, a0, Li, Lf, bmin=bmin, bmax=bmax, chi_tol=None, rescan=False, ODE_tol=tol,
                         cps=cps))
    att = _attempts(ctx.calls)
    exact = lambda L: g0 + c * (L ** 3 - Li ** 3) / 3.0
    allcps = sorted(set(cps) | {Li, Lf}, reverse=True)
    # the first step: bmax Li, clamped to 0.85 of the distance to the next required cutoff
    below = [x for x in allcps if x < Li]
    assert att[0] == (Li, pytest.approx(min(bmax * Li, 0.85 * (Li - below[0])), rel=1e-12))
    n_acc = n_rej = 0
    for (L, dL), nxt in zip(att, att[1:] + [None]):
        err = (abs(c) * dL ** 3 / 24.0) / (tol[0] + max(abs(exact(L - dL)), abs(exact(L))) * tol[1])
        assert err < 10.0
        accept = err <= 1.0 or dL <= bmin * (1.0 + 1e-12)  # dL read back from the stage Λs
        L2 = L - dL if accept else L
        dL2 = max(bmin, min(bmax * L2, 0.85 * (1.0 / err) ** (1.0 / 3.0) * dL))
        lower = [x for x in allcps if x < L2]
        dL2 = min(dL2, L2 - (lower[0] if lower else 0.0))
        n_acc += accept
        n_rej += not accept
        if nxt is not None:
            assert nxt[0] == pytest.approx(L2, rel=1e-13, abs=1e-13)
            assert nxt[1] == pytest.approx(dL2, rel=1e-9)
    print(f"bmin {bmin} cps {cps}: {len(att)} attempts, {n_rej} rejected; Λ {res.Lambdas}", flush=True)
    assert (n_acc, n_rej) == (res.accepted, res.rejected)
    assert n_rej >= 1 and res.Lambdas[-1] == Lf
    for x in cps:  # every required cutoff is hit
        assert any(abs(L - x) <= 1e-12 for L in res.Lambdas)
    if bmin > 0.1:  # a step at the minimal size is accepted whatever its error
        assert any(dL <= bmin * (1.0 + 1e-12) and err > 1.0 for (L, dL) in att
                   for err in [(abs(c) * dL ** 3 / 24.0) / (tol[0] + max(abs(exact(L - dL)), abs(exact(L))) * tol[1])])


def test_bs3_zero_error_takes_the_largest_step():
    # a vanishing embedded error (err = 0): 0.85 err^(-1/3) dΛ is unbounded and the step is bmax Λ
    prob = _prob()
    a0 = prob.bare_action()
    ctx = FakeContext(prob, rhs=lambda L: np.zeros(a0.gamma.size))
    res = _run(F.BS3Flow(ctx, prob, a0, 5.0, 1.0, chi_tol=None, rescan=False))
    att = _attempts(ctx.calls)
    assert res.rejected == 0 and res.status == "finished"
    for (L, dL) in att[1:-1]:
        assert dL == pytest.approx(0.1 * L, rel=1e-14)


@pytest.mark.parametrize("Lam,tol,max_iter,stop", [(0.5, (1e-9, 1e-30), 200, "abs"), (2.0, (1e-30, 1e-7), 200, "rel"),
                                                   (0.3, (1e-12, 1e-12), 7, "max_iter")])
def test_parquet_damping_and_stops(Lam, tol, max_iter, stop):
    prob = _prob()
    rng = np.random.default_rng(9)
    a0 = prob.bare_action()
    ms, mg = 0.4, -0.3
    cs = rng.standard_normal(a0.Sigma.size)
    cg = rng.standard_normal(a0.gamma.size)
    a0.gamma[:] = rng.standard_normal(a0.gamma.size)
    ctx = FakeContext(prob, maps=((ms, cs), (mg, cg)))
    res = F.launch_parquet(ctx, Lam, a0, max_iter=max_iter, parquet_tol=tol)
    beta = min(Lam, 1.0)
    rg = 1.0 - beta * (1.0 - mg)
    sig_star, gam_star = cs / (1.0 - ms), cg / (1.0 - mg)
    errs = [e for e, _ in res.history]
    ratios = [errs[k + 1] / errs[k] for k in range(len(errs) - 1)]
    print(f"Λ={Lam} β={beta}: {res.iterations} iterations ({res.status}), SDE {res.sde_iterations}; "
          f"ratios {ratios[:3]} (r_g = {rg})", flush=True)
    # the first error: |f(Γ0) - Γ0| = (1 - m_g) |Γ0 - Γ*|
    assert errs[0] == pytest.approx((1.0 - mg) * np.max(np.abs(a0.gamma - gam_star)), rel=1e-12)
    for q in ratios:
        assert q == pytest.approx(abs(rg), rel=1e-6)
    got = ctx.download(F.A)
    if stop == "max_iter":
        assert res.status == "max_iter" and res.iterations == max_iter == len(errs)
        assert errs[-1] >= tol[0]
    else:
        assert res.status == "converged"
        last_abs, last_rel = res.history[-1]
        prev_abs, prev_rel = res.history[-2]
        assert (last_abs < tol[0]) if stop == "abs" else (last_rel < tol[1])
        assert prev_abs >= tol[0] and prev_rel >= tol[1]
        # the result sits within the stopping error of the fixed point
        assert np.max(np.abs(got.gamma - gam_star)) <= last_abs / (1.0 - mg) + 1e-14
        assert np.max(np.abs(got.Sigma - sig_star)) <= 1e-6 * np.max(np.abs(sig_star))
    # Σ is converged by the first SDE loop; every later parquet iteration takes one SDE step
    n_first = res.sde_iterations - (res.iterations - 1)
    rs = 1.0 - beta * (1.0 - ms)
    d0 = (1.0 - ms) * np.max(np.abs(a0.Sigma - sig_star))
    # the SDE loop stops on the first (1 - m_s) |Σ_k - Σ*| below a tolerance, k = n_first - 1
    abs_k = lambda k: d0 * abs(rs) ** k
    assert abs_k(n_first - 2) >= tol[0] or n_first == 1
