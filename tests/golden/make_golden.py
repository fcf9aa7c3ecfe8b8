"""Writes the golden fixtures of tests/golden/ from closed-form integrands with scipy.

Nothing from the reference is imported or run: the integrands are restated from the
reference's own test (Flow/test.jl:1-13) and from the propagator formulas with Σ = 0
(Flow/bubbles.jl:2-96), then integrated to high accuracy.

    python tests/golden/make_golden.py
"""
import json
from pathlib import Path

import numpy as np
from scipy.integrate import quad

HERE = Path(__file__).resolve().parent


def quadrature_golden():
    # ∫_1^5 x^i e^{-x²} sin(2πx) dx, i = 1..100 (Flow/test.jl:51-61)
    vals = []
    for i in range(1, 101):
        v, _ = quad(lambda x: x ** i * np.exp(-x * x) * np.sin(2 * np.pi * x), 1.0, 5.0, epsabs=0.0, epsrel=1e-13,
                    limit=1000)
        vals.append(v)
    return {"a": 1.0, "b": 5.0, "powers": list(range(1, 101)), "values": vals}


def G0(L, w):
    return 0.0 if abs(w) <= 1e-8 else -np.expm1(-w * w / (L * L)) / w


def S0(L, w):
    # single-scale propagator at Σ = 0: (G/G0)² e^{-w²/Λ²} 2w/Λ³ with G = G0
    return 0.0 if abs(w) <= 1e-8 else np.exp(-w * w / (L * L)) * 2.0 * w / L ** 3


def kat_bubble(L, ch, W, v):
    if ch == 0:
        x1, x2 = v + 0.5 * W, 0.5 * W - v
    elif ch == 1:
        x1, x2 = v + 0.5 * W, v - 0.5 * W
    else:
        x1, x2 = v - 0.5 * W, v + 0.5 * W
    return (S0(L, x1) * G0(L, x2) + S0(L, x2) * G0(L, x1)) / (2.0 * np.pi)


def bubble_integrals():
    # ∫_{-∞}^{∞} p_kat(v) dv at Σ = dΣ = 0 for a few (Λ, channel, W)
    out = []
    for L in (0.5, 1.0, 3.0):
        for ch in range(3):
            for W in (0.0, 0.7, 2.5, 9.0):
                pts = sorted({-0.5 * W, 0.5 * W, 0.0})
                v1, _ = quad(lambda v: kat_bubble(L, ch, W, v), -np.inf, pts[0], epsabs=1e-14, epsrel=1e-12, limit=500)
                v2, _ = quad(lambda v: kat_bubble(L, ch, W, v), pts[0], pts[-1], epsabs=1e-14, epsrel=1e-12,
                             limit=500, points=pts[1:-1] or None) if pts[-1] > pts[0] else (0.0, 0.0)
                v3, _ = quad(lambda v: kat_bubble(L, ch, W, v), pts[-1], np.inf, epsabs=1e-14, epsrel=1e-12, limit=500)
                out.append({"L": L, "ch": ch, "W": W, "integral": v1 + v2 + v3})
    return out


if __name__ == "__main__":
    (HERE / "quadrature.json").write_text(json.dumps(quadrature_golden(), indent=1))
    (HERE / "bubble_integrals.json").write_text(json.dumps(bubble_integrals(), indent=1))
    print("written")
