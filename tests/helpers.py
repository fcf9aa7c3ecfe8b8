"""Shared fixtures-as-functions for the parity tests."""
from __future__ import annotations

import numpy as np

from pffrg import action as A
from pffrg import flow as F
from pffrg import lattice as LT
from pffrg import mesh as M


def small_problem(name="square", size=2, N=6, initial=1.0, J=((1.0,),)):
    return F.Problem.build(name, size, [list(j) for j in J], initial=initial, num_sigma=N + 2, num_Omega=N, num_nu=N,
                           normalize=False)


def rel_err(got: np.ndarray, ref: np.ndarray) -> float:
    """max |got - ref| / max |ref| (the reference's max-norm, Action.jl:346-359)."""
    d = float(np.max(np.abs(got - ref)))
    s = float(np.max(np.abs(ref)))
    return d / s if s > 0 else d


def per_array_err(got: A.Action, ref: A.Action) -> dict:
    out = {}
    for comp in range(2):
        for ch, cn in enumerate("stu"):
            for q in ("q1", "q2_1", "q2_2", "q3"):
                out[f"{'sd'[comp]}{cn}{q}"] = rel_err(got.view(comp, ch, q), ref.view(comp, ch, q))
    return out


def half_elements(nO, nn):
    """Output elements (channel, kernel, w1, w2, w3) of a half pass: q3 with w2 >= w3 (compute.jl:57-112)."""
    return [(ch, k, w1, w2, w3) for ch in range(3) for w1 in range(nO)
            for (k, w2, w3) in ([(1, -1, -1)] + [(2, w, -1) for w in range(nn)] + [(3, -1, w) for w in range(nn)]
                                + [(4, w2, w3) for w3 in range(nn) for w2 in range(w3, nn)])]


def entries(a: A.Action, ch, k, w1, w2, w3):
    """Both components of one output element of action a, concatenated."""
    q = {1: "q1", 2: "q2_1", 3: "q2_2", 4: "q3"}[k]
    idx = {1: (w1,), 2: (w1, w2), 3: (w1, w3), 4: (w1, w2, w3)}[k]
    return np.concatenate([a.view(c, ch, q)[(slice(None),) + idx] for c in range(2)])


def class_sample(elems, nO, nn, stride):
    """Every stride-th element, plus every element of each (channel, kernel) class at the ends of
    the Ω and ν ranges (w1 in {0, nO-1}; each used frequency index in {0, nn-1})."""
    pick = set(elems[::stride])
    eo, en = {0, nO - 1}, {0, nn - 1}
    for e in elems:
        _, _, w1, w2, w3 = e
        if w1 in eo and (w2 < 0 or w2 in en) and (w3 < 0 or w3 in en):
            pick.add(e)
    return sorted(pick)


def full_elements(nO, nn):
    """Output elements of a full pass: every q3 entry (compute.jl:115-174)."""
    return [(ch, k, w1, w2, w3) for ch in range(3) for w1 in range(nO)
            for (k, w2, w3) in ([(1, -1, -1)] + [(2, w, -1) for w in range(nn)] + [(3, -1, w) for w in range(nn)]
                                + [(4, w2, w3) for w3 in range(nn) for w2 in range(nn)])]
