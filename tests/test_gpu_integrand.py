"""Integrand parity at hand-picked loop frequencies: the GPU integrand of single output elements
(pffrg_probe_integrand, no quadrature) against the FP64 oracle's at loop frequencies that put
the interpolation on its edge cases -- v = 0, v on mesh points (the frequency transfers then
land exactly on mesh points of other buffers), the boundary Ω_max + ν_max where the integrand
jumps, points beyond every mesh, and tiny values -- for every pass, channel and kernel type.
The GPU brackets mesh values through a hint table indexed by the value's bits; the oracle
searches the mesh linearly as get_param does (param.jl:26-86), so any difference in a bracket
shows here as an O(1) error.  Tolerance: 1e-12 relative max-norm (same arithmetic, FP64)."""
import numpy as np
import pytest

import oracle as O
from pffrg import action as A

from helpers import small_problem

pytestmark = pytest.mark.gpu

PASSES = [(O.P_KAT, "kat"), (O.P_LEFT, "left"), (O.P_CENTRAL, "central"), (O.P_BSE, "bse"), (O.P_REDUCED, "reduced")]


def _loop_freqs(m, L, W):
    nu = np.concatenate([m.ns, m.nt])
    om = np.concatenate([m.Os, m.Ot])
    ref = L + 0.5 * W
    val = max(m.Os[-1] + m.ns[-1], m.Ot[-1] + m.nt[-1])
    pts = [0.0, 1e-12, 1e-9, 2 * ref, -2 * ref, val, -val, 1.5 * val, -3 * val, m.sigma[-1], 2 * m.sigma[-1]]
    pts += list(nu[::2]) + list(-nu[1::3]) + list(om[::3]) + list(-om[1::4]) + list(0.5 * nu[::5])
    pts += list(nu[1:6] - W) + list(W - nu[1:6]) + list(2 * nu[1:4] + W)
    rng = np.random.default_rng(7)
    pts += list(rng.uniform(-1.2 * val, 1.2 * val, 12))
    return np.array(pts)


@pytest.mark.parametrize("name,size,N", [("square", 2, 6), ("kagome", 3, 7), ("pyrochlore", 3, 6)])
def test_integrand_matches_oracle_at_edge_frequencies(name, size, N):
    prob = small_problem(name, size, N)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=11)
    da = A.synthetic_action(ns, m, prob.bare, seed=12)
    da.Sigma[:] = 0.1 * np.sin(np.arange(len(da.Sigma)))
    L = 0.8
    orc = O.Oracle(prob.r, m)
    nO, nn = m.num_Omega, m.num_nu
    worst = 0.0
    with prob.context(n_slots=2) as ctx:
        ctx.upload(0, a)
        ctx.upload(1, da)
        for ps, pname in PASSES:
            chans = [0] if ps == O.P_REDUCED else [0, 1, 2]
            for ch in chans:
                for (kernel, w1, w2, w3) in [(1, 0, 0, 0), (2, nO - 1, 1, 0), (3, 1, 0, nn - 1), (4, 0, 0, 0),
                                             (4, nO // 2, nn - 1, 1), (4, nO - 1, 2, nn - 2)]:
                    if ps == O.P_REDUCED and kernel in (2, 4):
                        continue
                    Om = m.Ot if ch == 1 else m.Os
                    v = _loop_freqs(m, L, Om[w1])
                    got = ctx.probe_integrand(ps, L, 0, 1, ch, kernel, w1, w2, w3, v)
                    ref = np.array([orc.integrand(ps, L, a, da, ch, kernel, w1, w2, w3, float(x)) for x in v])
                    scale = max(np.max(np.abs(ref)), 1e-300)
                    err = float(np.max(np.abs(got - ref)) / scale)
                    worst = max(worst, err)
                    assert err < 1e-12, (pname, ch, kernel, w1, w2, w3, err)
    print(f"{name}: worst integrand rel err {worst:.3e}", flush=True)
