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
