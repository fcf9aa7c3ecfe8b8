"""The channel masks of the oracle's multiloop, parquet and self-energy-correction integrands,
pinned by invariances the reference's keyword masks imply (no integrand is restated here).

The reference drops channel buffers with get_Γ's ch_s/ch_t/ch_u keywords or empty buffers:
 - left pass: flow_su2_s.jl:86 (da without s), flow_su2_u.jl:86 (da without u),
   flow_su2_t.jl:108-134 (non-local da without t, local da without u);
 - central pass: flow_su2_s.jl:141-142 (da_l s only), flow_su2_u.jl:141-142 (u only),
   flow_su2_t.jl:216-221 (non-local da_l t only, local da_l u only);
 - BSE: BSE_su2_{s,t,u}.jl: the left pass's masks on a itself;
 - reduced s bubble (SDE_su2.jl:2-43): the left vertex is the bare coupling;
 - SDE self energy (SDE_su2.jl:45-77): s buffers only;
 - first dΣ correction (flow_su2_self.jl:36-70): local vertex without u, non-local without t.
A mask drops a buffer, and an s or u buffer whose fold sets the map flag reads the other
channel's data (get_Γ_comp, Action.jl:29-103), so the s and u masks do not drop data channels.
The t channel is never mapped: a t mask drops the t data.  The tests therefore use
 - t data: the left and central passes are linear in the derivative action, so a derivative
   holding t data only gives exactly zero where its t buffers are dropped (the adaptive
   quadrature of an integrand that is identically zero stops at its first level with zero);
 - BSE = left pass with the action as its own derivative, bitwise (same masks, same bubble);
 - the reduced bubble vanishes with the bare couplings, whatever the channels hold.
An error in a mask shows as a non-zero output or a difference (tests/test_oracle_mutations
plants such errors)."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A

from helpers import small_problem

LAM, EVAL = 1.0, 4


def _channel(a, ch, comps=(0, 1)):
    n = A.channel_len(a.ns, a.nO, a.nn)
    return [slice((c * 3 + ch) * n, (c * 3 + ch + 1) * n) for c in comps]


def _only(a, keep, comps=(0, 1), bare=False):
    """a copy of a holding the channels in keep (components comps) and no others."""
    b = a.copy()
    for ch in range(3):
        for c in range(2):
            if ch not in keep or c not in comps:
                b.gamma[_channel(b, ch, (c,))[0]] = 0.0
    if not bare:
        b.bare[:] = 0.0
    return b


def _setup(name="square", size=3, N=5, seed=17):
    prob = small_problem(name, size, N, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=seed)
    orc = O.Oracle(prob.r, m)
    orc.corrs[:] = orc.compute_corrs(LAM, a, a, False)
    return prob, a, orc


def _pass(orc, p, a, d, full=True):
    out = A.Action.empty(a.ns, orc.m)
    orc.run_pass(p, LAM, a, d, out, full, EVAL)
    return out


def _chan_out(out, ch):
    return np.concatenate([out.gamma[s] for s in _channel(out, ch)])


def _site_out(out, ch, i):
    return np.concatenate([out.view(c, ch, q)[i].ravel() for c in range(2) for q in ("q1", "q2_1", "q2_2", "q3")])


def test_left_pass_masks():
    # t channel: the derivative's t data reaches the t-channel output only through the local
    # (site-0) vertex, whose product with the right vertex at site i vanishes when the action
    # vanishes on i's orbit under the site exchange; its u data reaches site i (non-local vertex)
    prob, a, orc = _setup(seed=17)
    d = A.synthetic_action(a.ns, orc.m, prob.bare, seed=18)
    x = np.asarray(prob.r.exchange)
    i = next(k for k in range(1, a.ns) if x[k] != 0)
    a0 = a.copy()
    for k in {i, int(x[i])}:
        a0.bare[:, k] = 0.0
        for c in range(2):
            for ch in range(3):
                for q in ("q1", "q2_1", "q2_2", "q3"):
                    a0.view(c, ch, q)[k] = 0.0
    out_t = _pass(orc, O.P_LEFT, a0, _only(d, {1}))
    assert np.all(_site_out(out_t, 1, i) == 0.0)
    assert np.max(np.abs(_site_out(out_t, 1, 0))) > 1e-6
    out_u = _pass(orc, O.P_LEFT, a0, _only(d, {2}))
    assert np.max(np.abs(_site_out(out_u, 1, i))) > 1e-6
    # the s and u outputs read the derivative's t data (non-local vertex)
    for ch in (0, 2):
        assert np.max(np.abs(_chan_out(out_t, ch))) > 1e-6


def test_central_pass_masks():
    # the s and u outputs take the derivative through an s (u) buffer only: its t data never
    # reaches them; the t output takes it through a t buffer
    prob, a, orc = _setup(seed=21)
    dl = A.synthetic_action(a.ns, orc.m, prob.bare, seed=22)
    out = _pass(orc, O.P_CENTRAL, a, _only(dl, {1}), full=False)
    assert np.all(_chan_out(out, 0) == 0.0)
    assert np.all(_chan_out(out, 2) == 0.0)
    assert np.max(np.abs(_chan_out(out, 1))) > 1e-6
    out = _pass(orc, O.P_CENTRAL, a, _only(dl, {0, 2}), full=False)
    assert min(np.max(np.abs(_chan_out(out, ch))) for ch in range(3)) > 1e-6


def test_bse_is_left_pass_with_the_action_as_derivative():
    # BSE_su2_{s,t,u}.jl use the left pass's masks on a; the bubble is the plain one in both
    prob, a, orc = _setup(seed=23)
    bse = _pass(orc, O.P_BSE, a, a, full=False)
    left = _pass(orc, O.P_LEFT, a, a, full=False)
    np.testing.assert_array_equal(bse.gamma, left.gamma)


def test_reduced_bubble_left_vertex_is_bare():
    prob, a, orc = _setup(seed=25)
    nobare = a.copy()
    nobare.bare[:] = 0.0
    out = A.Action.empty(a.ns, orc.m)
    orc.run_pass(O.P_REDUCED, LAM, nobare, nobare, out, False, EVAL)
    assert np.all(out.gamma == 0.0)
    orc.run_pass(O.P_REDUCED, LAM, a, a, out, False, EVAL)
    assert np.max(np.abs(out.view(0, 0, "q2_2"))) > 1e-6


def test_sde_self_energy_drops_the_t_channel():
    # compute_Σ! (SDE.jl:58-96) overwrites a2's s-channel q1 and q2_2 with the reduced bubble of a1
    # and integrates Σ through s buffers only (SDE_su2.jl:54-60): a2's t data must not matter
    prob, a, orc = _setup(seed=27)
    s_only = _only(a, {0, 2}, bare=True)
    sig = []
    for a2 in (a, s_only):
        tgt = a2.copy()
        orc.compute_sigma(LAM, a, tgt, EVAL)
        sig.append(tgt.Sigma.copy())
    np.testing.assert_array_equal(sig[0], sig[1])
    assert np.max(np.abs(sig[0])) > 1e-8


def test_first_sigma_correction_masks():
    # corr1: local vertex without u buffer, non-local (density only) without t buffer, so da_Σ's
    # t data enters through the local vertex (site 0) only
    prob, a, orc = _setup(seed=29)
    das = A.synthetic_action(a.ns, orc.m, prob.bare, seed=30)

    def corr1(daS):
        da = A.Action.empty(a.ns, orc.m)
        d2 = daS.copy()
        orc.compute_dsigma_corr(LAM, a, da, d2)
        return d2.Sigma.copy()

    t_only = _only(das, {1})
    off_site0 = t_only.copy()
    for c in range(2):
        for q in ("q1", "q2_1", "q2_2", "q3"):
            off_site0.view(c, 1, q)[0] = 0.0
    assert np.all(corr1(off_site0) == 0.0)
    assert np.max(np.abs(corr1(t_only))) > 1e-8
    assert np.max(np.abs(corr1(_only(das, {2}, comps=(1,))))) > 1e-8
