"""Closed forms for the oracle's left, central, BSE and reduced passes (flow_su2_{s,t,u}.jl
left/central kernels, BSE_su2_{s,t,u}.jl, SDE_su2.jl:2-43) on bare-only actions.

With Σ = 0 and every channel zero, each interpolated vertex is the bare coupling vector, so the
site and spin contraction of every pass reduces to that of the Katanin pass (the same kernels,
flow_su2_t.jl:54-80, with the left vertex taken from a different action) times the integral of
the pass's own bubble.  On the square lattice with nearest-neighbour coupling b = 1/4 on the
bonds and none on site 0 (so the local t terms vanish):
    s: spin 2 bl br, density -3 bl br;  u: spin -2 bl br, density -3 bl br;
    t: spin 2 (2S) P_i bl br, density 0  (P_i two-step nearest-neighbour paths to site i),
with bl, br the left and right vertex's bare couplings and the plain bubble
p = -G0(x1) G0(x2) / 2π (bubbles.jl:77-94) in place of the Katanin one:
 - left pass: left vertex from da (bare couplings β b), right from a;
 - central pass: left from a, right from da_l (β b);
 - BSE and reduced passes: both from a (the reduced pass's left vertex is a's bare coupling).
The Katanin closed form itself is test_bare_only_rhs_closed_form of test_oracle_cpu.
Tolerance 1e-4 relative, the adaptive quadrature's own tolerance.  A wrong bubble sign, a bubble
of the wrong kind or the wrong action on either side fails (planted in
tests/test_oracle_mutations)."""
import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O
from pffrg import action as A
from pffrg import lattice as LT
from pffrg import mesh as M

L = 1.0
BETA = 0.7


def _G0(w):
    return 0.0 if abs(w) <= 1e-8 else -np.expm1(-w * w / L ** 2) / w


def _plain_integral(ch, W):
    def p(v):
        x1, x2 = {0: (v + W / 2, W / 2 - v), 1: (v + W / 2, v - W / 2), 2: (v - W / 2, v + W / 2)}[ch]
        return -_G0(x1) * _G0(x2) / (2 * np.pi)
    return sum(quad(p, lo, hi, epsabs=1e-14, epsrel=1e-12, limit=400)[0]
               for lo, hi in ((-np.inf, -W / 2 - 1), (-W / 2 - 1, W / 2 + 1), (W / 2 + 1, np.inf)))


def _setup():
    m = M.launch_mesh(1.0, 8, 6, 6, 5)
    l, r = LT.build("square", 3, [[1.0]])
    orc = O.Oracle(r, m)
    a = A.Action.empty(r.n_sites, m)
    a.bare[:] = LT.bare_couplings(l, r)
    da = A.Action.empty(r.n_sites, m)
    da.bare[:] = BETA * a.bare
    nb = [np.array(v) for v in ([1, 0], [-1, 0], [0, 1], [0, -1])]
    P = np.array([sum(1 for k in nb if np.abs(k - v[:2]).sum() == 1) for v in r.vecs])
    return m, r, orc, a, da, P


def _expected(ch, bare, scale, P, S=0.5):
    # (spin, density) coefficients of the bare contraction per site: bl br = scale b_i² in the s
    # and u channels; the t channel sums bl br = scale / 16 over the P_i two-step paths
    b2 = scale * bare * bare
    if ch == 0:
        return 2 * b2, -3 * b2
    if ch == 2:
        return -2 * b2, -3 * b2
    return 2 * (2 * S) * P * scale / 16.0, 0.0 * P


@pytest.mark.parametrize("pass_,scale", [(O.P_LEFT, BETA), (O.P_CENTRAL, BETA), (O.P_BSE, 1.0)])
def test_bare_only_pass_closed_form(pass_, scale):
    m, r, orc, a, da, P = _setup()
    b = a.bare[0]
    assert np.allclose(np.sort(np.abs(b))[-1], 0.25) and b[0] == 0.0
    orc.compute_corrs(L, a, a, False)
    n = 0
    for ch in range(3):
        Om = m.Ot if ch == 1 else m.Os
        for w1 in (0, 2, 4):
            W = Om[w1]
            out = A.Action.empty(r.n_sites, m)
            orc.compute_element(pass_, L, a, da if pass_ != O.P_BSE else a, ch, 4, w1, 1, 2, 10, out)
            qs, qd = out.view(0, ch, "q3")[:, w1, 1, 2], out.view(1, ch, "q3")[:, w1, 1, 2]
            Ip = _plain_integral(ch, W)
            es, ed = _expected(ch, b, scale, P)
            es, ed = es * Ip, ed * Ip
            sc = max(float(np.max(np.abs(es))), 1e-30)
            assert np.max(np.abs(qs - es)) <= 1e-4 * sc, (ch, w1, qs, es)
            assert np.max(np.abs(qd - ed)) <= 1e-4 * max(float(np.max(np.abs(ed))), sc), (ch, w1, qd, ed)
            n += 1
    assert n == 9


def test_bare_only_reduced_closed_form():
    m, r, orc, a, da, P = _setup()
    orc.compute_corrs(L, a, a, False)
    for w1 in (0, 2, 4):
        W = m.Os[w1]
        Ip = _plain_integral(0, W)
        es, ed = _expected(0, a.bare[0], 1.0, P)
        for kernel, w3, q, idx in ((1, -1, "q1", (slice(None), w1)), (3, 2, "q2_2", (slice(None), w1, 2))):
            out = A.Action.empty(r.n_sites, m)
            orc.compute_element(O.P_REDUCED, L, a, a, 0, kernel, w1, -1, w3, 10, out)
            qs, qd = out.view(0, 0, q)[idx], out.view(1, 0, q)[idx]
            sc = max(float(np.max(np.abs(es * Ip))), 1e-30)
            assert np.max(np.abs(qs - es * Ip)) <= 1e-4 * sc, (w1, kernel, qs, es * Ip)
            assert np.max(np.abs(qd - ed * Ip)) <= 1e-4 * sc, (w1, kernel, qd, ed * Ip)
