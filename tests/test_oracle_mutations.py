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


# Planted mask errors (one per pass family): each must make the matching invariance of
# tests/test_oracle_masks_cpu.py fail on the mutated oracle.
MASK_MUTATIONS = [
    ("test_left_pass_masks",
     "gamma_vec(y, d, D, bs1, bt1, bu1, Mask{true, false, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());",
     "gamma_vec(y, d, D, bs1, bt1, bu1, Mask{true, true, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());"),
    ("test_central_pass_masks",
     "case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{true, false, false}, t2.data()); break;",
     "case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{true, true, false}, t2.data()); break;"),
    ("test_central_pass_masks",
     "case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{false, false, true}, t2.data()); break;",
     "case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{false, true, true}, t2.data()); break;"),
    ("test_bse_is_left_pass_with_the_action_as_derivative",
     "case P_BSE: gamma_vec(y, d, A, bs1, bt1, bu1, Mask{false, true, true}, t1.data());",
     "case P_BSE: gamma_vec(y, d, A, bs1, bt1, bu1, Mask{true, true, true}, t1.data());"),
    ("test_reduced_bubble_left_vertex_is_bare",
     "for (int k = 0; k < nc; ++k) a[k] = A.bare[k * ns + i];",
     "for (int k = 0; k < nc; ++k) a[k] = t2[k * ns + i];"),
    ("test_sde_self_energy_drops_the_t_channel",  # the dΣ kernel's local t buffer copied into the SDE
     "double val = sigma_loop2(*y, d, *a2, b2s, e, e, b1s, e, e, SONLY, SONLY);",
     "double val = sigma_loop2(*y, d, *a2, b2s, buffer_t(*y, v - w, 0.5 * (v + w), 0.5 * (v + w)), e, b1s, e, e, ALL, SONLY);"),
    ("test_first_sigma_correction_masks",
     "sigma_loop(*y, d, *daS, b, Mask{true, true, false}, Mask{true, false, true});",
     "sigma_loop(*y, d, *daS, b, Mask{true, true, false}, Mask{true, true, true});"),
]


# Planted errors in the self-energy kernels (frequency arguments, propagator factors): each must
# make the matching scipy comparison of tests/test_oracle_sigma_cpu.py fail.
SIGMA_MUTATIONS = [
    ("test_dsigma_kernel_against_scipy", "b.b1t = buffer_t(y, 0.0, w, v);", "b.b1t = buffer_t(y, 0.0, v, w);"),
    ("test_dsigma_kernel_against_scipy", "b.b2s = buffer_s(y, v + w, 0.5 * (-v + w), 0.5 * (-v + w));",
     "b.b2s = buffer_s(y, v + w, 0.5 * (-v + w), 0.5 * (v - w));"),
    ("test_dsigma_kernel_against_scipy", "b.b2t = buffer_t(y, v - w, 0.5 * (v + w), 0.5 * (v + w));",
     "b.b2t = buffer_t(y, v + w, 0.5 * (v + w), 0.5 * (v + w));"),
    ("test_dsigma_kernel_against_scipy", "val *= S_single(*y, L, v, a->Sigma) / (2.0 * PI);",
     "val *= -G_full(*y, L, v, a->Sigma) / (2.0 * PI);"),
    ("test_dsigma_kernel_against_scipy", "val -= 2.0 * y.mult[j] * (2.0 * y.S) * gamma_comp",
     "val -= 2.0 * y.mult[j] * y.S * gamma_comp"),
    ("test_dsigma_corrections_against_scipy", "val *= G * G * get_sigma(*y, daS->Sigma, v) / (2.0 * PI);",
     "val *= G * get_sigma(*y, daS->Sigma, v) / (2.0 * PI);"),
    ("test_sde_sigma_kernel_against_scipy", "Buffer b1s = buffer_s(*y, v + w, INFINITY, 0.5 * (v - w));",
     "Buffer b1s = buffer_s(*y, v + w, INFINITY, 0.5 * (-v + w));"),
]

# Planted errors in symmetrize_add_to (action_su2.jl:176-250): an unswapped q2 and a site
# exchange in the u channel must fail the laws of tests/test_oracle_symmetry_cpu.py.
SYMMETRY_MUTATIONS = [
    ("test_symmetrize_add_to_laws", "double a21 = pair(l[d.q21 + k], r[d.q22 + kx]);",
     "double a21 = pair(l[d.q21 + k], r[d.q21 + kx]);"),
    ("test_symmetrize_add_to_laws", "auto ex = [&](int i) { return (ch == 2) ? i : (int)y.exchange[i]; };",
     "auto ex = [&](int i) { return (int)y.exchange[i]; };"),
    ("test_symmetrize_add_to_laws",
     "o[d.q1 + i + (long)d.ns * w] += pair(l[d.q1 + i + (long)d.ns * w], r[d.q1 + ex(i) + (long)d.ns * w]);",
     "o[d.q1 + i + (long)d.ns * w] += pair(l[d.q1 + i + (long)d.ns * w], r[d.q1 + i + (long)d.ns * w]);"),
]

# Planted errors in the bubble and in the action each side of the left, central and reduced
# passes: each must fail the bare-only closed forms of tests/test_oracle_passes_cpu.py.
PASS_MUTATIONS = [
    ("test_bare_only_pass_closed_form", "    return -prop(y, c.L, x1, x2, c.a->Sigma);\n}",
     "    return prop(y, c.L, x1, x2, c.a->Sigma);\n}"),
    ("test_bare_only_pass_closed_form",
     "case P_LEFT: gamma_vec(y, d, D, bs1, bt1, bu1, Mask{false, true, true}, t1.data());",
     "case P_LEFT: gamma_vec(y, d, A, bs1, bt1, bu1, Mask{false, true, true}, t1.data());"),
    ("test_bare_only_pass_closed_form",
     "gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{false, true, false}, t2.data());",
     "gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, Mask{false, true, false}, t2.data());"),
    ("test_bare_only_reduced_closed_form", "for (int k = 0; k < nc; ++k) a[k] = A.bare[k * ns + i];",
     "for (int k = 0; k < nc; ++k) a[k] = 0.5 * A.bare[k * ns + i];"),
]


# Planted errors in the U(1)-DM flags, products, self-energy loop, correlations and symmetrisation:
# each must fail its pin in tests/test_oracle_u1_cpu.py.
U1_MUTATIONS = [
    ('test_katanin_pass_preserves_the_u1_vertex_symmetry',
     '        if (comp == 2) sgn *= -1.0;\n        else if (comp == 4) src = 5;',
     '        if (comp == 2) sgn *= 1.0;\n        else if (comp == 4) src = 5;'),
    ('test_su2_embedding_reproduces_the_su2_oracle',
     'g[ZZ] = a[ZZ] * b[DD] + a[DD] * b[ZZ] - a[DZ] * b[ZD] - 2.0 * a[XX] * b[XX]',
     'g[ZZ] = a[ZZ] * b[DD] + a[DD] * b[ZZ] - a[DZ] * b[ZD] + 2.0 * a[XX] * b[XX]'),
    ('test_pi_rotation_commutes_with_every_right_hand_side',
     'g[DM] = a[DM] * c4[DD] - a[XX] * c4[DZ]',
     'g[DM] = a[DM] * c4[DD] - a[XX] * c4[DD]'),
    ('test_pi_rotation_commutes_with_every_right_hand_side',
     'G[ZD] += -p * (-2.0) * m * (a[ZD] * b[DD] + a[ZZ] * b[ZD]);',
     'G[ZD] += -p * (-2.0) * m * (a[ZD] * b[DD] + a[ZZ] * b[DD]);'),
    ('test_su2_embedding_reproduces_the_su2_oracle',
     'val -= 2.0 * y.mult[j] * gamma_comp(y, d, a, DD, j, b1s, b1t, b1u, mn);',
     'val -= 2.0 * y.mult[j] * gamma_comp(y, d, a, ZZ, j, b1s, b1t, b1u, mn);'),
    ('test_su2_embedding_reproduces_the_su2_oracle',
     'val -= (2.0 * vxx - vzz - vdd) / (2.0 * tp2);',
     'val -= (vxx - vzz - vdd) / (2.0 * tp2);'),
    ('test_u1_symmetrize_add_to_laws',
     '    return (comp == ZD || comp == DZ) ? -1.0 : 1.0;\n}',
     '    return 1.0;\n}'),
    ('test_interpolation_and_flags_of_every_component',
     '    if ((comp == 3 || comp == 4) && b.sgn_nu) sgn *= -1.0;',
     '    if ((comp == 3 || comp == 5) && b.sgn_nu) sgn *= -1.0;'),
]


def _fails_on_mutant(test_file, test, orig, bad, tmp_path):
    text = SRC.read_text()
    assert text.count(orig) == 1, orig
    (tmp_path / "m.cpp").write_text(text.replace(orig, bad))
    lib = tmp_path / "liboracle_m.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                    f"-I{ROOT / 'pffrgsolver.jl_amd' / 'csrc'}", "-shared", "-o", str(lib), str(tmp_path / "m.cpp")],
                   check=True, timeout=300)
    env = dict(os.environ, PFFRG_ORACLE_LIB=str(lib))
    out = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                          f"{ROOT / 'tests' / test_file}::{test}"],
                         env=env, capture_output=True, text=True, timeout=600, cwd=str(ROOT))
    assert out.returncode != 0, f"{test} passed on the mutated oracle:\n{out.stdout[-2000:]}"
    assert "failed" in out.stdout


@pytest.mark.parametrize("test,orig,bad", MASK_MUTATIONS)
def test_mask_invariance_detects_planted_mistake(test, orig, bad, tmp_path):
    _fails_on_mutant("test_oracle_masks_cpu.py", test, orig, bad, tmp_path)


@pytest.mark.parametrize("test,orig,bad", SIGMA_MUTATIONS)
def test_sigma_pin_detects_planted_mistake(test, orig, bad, tmp_path):
    _fails_on_mutant("test_oracle_sigma_cpu.py", test, orig, bad, tmp_path)


@pytest.mark.parametrize("test,orig,bad", SYMMETRY_MUTATIONS)
def test_symmetry_law_detects_planted_mistake(test, orig, bad, tmp_path):
    _fails_on_mutant("test_oracle_symmetry_cpu.py", test, orig, bad, tmp_path)


@pytest.mark.parametrize("test,orig,bad", PASS_MUTATIONS)
def test_pass_closed_form_detects_planted_mistake(test, orig, bad, tmp_path):
    _fails_on_mutant("test_oracle_passes_cpu.py", test, orig, bad, tmp_path)


@pytest.mark.parametrize("test,orig,bad", U1_MUTATIONS)
def test_u1_pin_detects_planted_mistake(test, orig, bad, tmp_path):
    _fails_on_mutant("test_oracle_u1_cpu.py", test, orig, bad, tmp_path)
