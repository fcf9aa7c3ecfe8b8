"""Planted mistakes: each test compiles a copy of the oracle with one plausible error in a
frequency-transfer formula of the flow integrands (flow_su2_{s,t,u}.jl:21-39 restated) and
checks that the vertex-symmetry invariant of test_oracle_cpu detects it.  The copies live in a
temporary directory and run in a subprocess; the oracle itself is not modified."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "pffrg_oracle.cpp"

PROBE = r'''
import sys
sys.path[:0] = [%(pkg)r, %(orc)r, %(tests)r]
import numpy as np
import oracle as O
from pffrg import action as A, lattice as LT, mesh as M
from test_oracle_cpu import _symmetric_action
m = M.launch_mesh(1.0, 6, 6, 6, 5)
l, r = LT.build("square", 2, [[1.0]])
a = _symmetric_action(r, m, LT.bare_couplings(l, r), seed=7)
orc = O.Oracle(r, m)
orc.symmetrize(a)
da = A.Action.empty(r.n_sites, m)
orc.compute_dsigma(1.0, a, da)
orc.compute_corrs(1.0, a, da, True)
out = A.Action.empty(r.n_sites, m)
orc.run_pass(O.P_KAT, 1.0, a, da, out, True, 10)
x = np.asarray(r.exchange)
worst = 0.0
for comp in range(2):
    for ch in range(3):
        ex = x if ch != 2 else np.arange(r.n_sites)
        q3 = out.view(comp, ch, "q3")
        worst = max(worst, float(np.abs(q3 - q3[ex].transpose(0, 1, 3, 2)).max() / np.abs(q3).max()))
print(worst)
'''

# (original expression, planted mistake): one per channel family of the integrands
MUTATIONS = [
    ("buffer_t(y, v - vs, 0.5 * (s + v + vs)", "buffer_t(y, v - vs, 0.5 * (s - v + vs)"),
    ("buffer_t(y, v - vu, 0.5 * (u + v + vu)", "buffer_t(y, v - vu, 0.5 * (u + v - vu)"),
    ("buffer_s(y, v + vt, 0.5 * (-t - v + vt)", "buffer_s(y, v - vt, 0.5 * (-t - v + vt)"),
]


def _worst(lib: Path | None, tmp: Path) -> float:
    env = dict(os.environ)
    if lib is not None:
        env["PFFRG_ORACLE_LIB"] = str(lib)
    code = PROBE % {"pkg": str(ROOT / "pffrgsolver.jl_amd"), "orc": str(ROOT / "oracle"), "tests": str(ROOT / "tests")}
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return float(out.stdout.strip().split()[-1])


@pytest.mark.parametrize("orig,bad", MUTATIONS)
def test_invariant_detects_planted_mistake(orig, bad, tmp_path):
    text = SRC.read_text()
    assert text.count(orig) == 1, orig
    (tmp_path / "m.cpp").write_text(text.replace(orig, bad))
    lib = tmp_path / "liboracle_m.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                    f"-I{ROOT / 'pffrgsolver.jl_amd' / 'csrc'}", "-shared", "-o", str(lib), str(tmp_path / "m.cpp")],
                   check=True, timeout=300)
    assert _worst(lib, tmp_path) > 1e-2


def test_invariant_holds_for_intact_oracle(tmp_path):
    assert _worst(None, tmp_path) < 2e-3
