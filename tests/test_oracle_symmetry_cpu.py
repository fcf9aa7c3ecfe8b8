"""symmetrize_add_to! (action_su2.jl:176-250) of the oracle: the left part of a multiloop
correction plus its right part, the left part under the ν <-> ν' exchange T.  T maps
q1[i] -> q1[x(i)], (q2_1, q2_2)[i] -> (q2_2, q2_1)[x(i)] and q3[i, w, v, v'] -> q3[x(i), w, v', v],
with x the reduced lattice's site exchange in the s and t channels and the identity in the u
channel (the same map that symmetrize!, action_su2.jl:136-170, uses to fill the upper triangle of
q3).  T is an involution, so for every left part L the sum L + T(L):
 - is T-symmetric, and symmetrize! leaves it unchanged;
 - vanishes exactly when L is T-antisymmetric (L = B - T(B)) and doubles L when it is symmetric;
 - adds linearly to the target.
A dropped exchange or an unswapped q2 fails these (planted in tests/test_oracle_mutations)."""
import numpy as np

import oracle as O
from pffrg import action as A

from helpers import small_problem

QS = ("q1", "q2_1", "q2_2", "q3")


def _T(a, x):
    """The exchange T applied to every array of a's vertex."""
    out = a.copy()
    for c in range(2):
        for ch in range(3):
            ex = x if ch != 2 else np.arange(len(x))
            out.view(c, ch, "q1")[:] = a.view(c, ch, "q1")[ex]
            out.view(c, ch, "q2_1")[:] = a.view(c, ch, "q2_2")[ex]
            out.view(c, ch, "q2_2")[:] = a.view(c, ch, "q2_1")[ex]
            out.view(c, ch, "q3")[:] = a.view(c, ch, "q3")[ex].transpose(0, 1, 3, 2)
    return out


def _setup(name, size, seed):
    prob = small_problem(name, size, 5, 1.0)
    ns, m = prob.r.n_sites, prob.mesh
    orc = O.Oracle(prob.r, m)
    return prob, orc, A.synthetic_action(ns, m, prob.bare, seed=seed), np.asarray(prob.r.exchange)


def _sym_add(orc, al, a):
    out = a.copy()
    orc.symmetrize_add_to(al, out)
    return out


def test_symmetrize_add_to_laws():
    # kagome and pyrochlore have sites that the exchange moves; square's are all fixed
    for name, size in (("square", 3), ("kagome", 3), ("pyrochlore", 3)):
        prob, orc, al, x = _setup(name, size, 61 + size)
        assert np.any(x != np.arange(len(x))) == (name != "square")
        zero = A.Action.empty(al.ns, orc.m)
        out = _sym_add(orc, al, zero)
        # T-symmetric, and unchanged by symmetrize!
        np.testing.assert_array_equal(_T(out, x).gamma, out.gamma)
        sym = out.copy()
        orc.symmetrize(sym)
        np.testing.assert_array_equal(sym.gamma, out.gamma)
        # T-antisymmetric left parts cancel exactly, symmetric ones double
        anti = al.copy()
        anti.gamma[:] = al.gamma - _T(al, x).gamma
        assert np.max(np.abs(anti.gamma)) > 0.1
        assert np.all(_sym_add(orc, anti, zero).gamma == 0.0)
        np.testing.assert_array_equal(_sym_add(orc, out, zero).gamma, 2.0 * out.gamma)
        # linear in the target
        a = A.synthetic_action(al.ns, orc.m, prob.bare, seed=99)
        np.testing.assert_array_equal(_sym_add(orc, al, a).gamma, a.gamma + out.gamma)
        print(f"{name} L={size}: {al.ns} sites, symmetrize_add_to laws hold", flush=True)
