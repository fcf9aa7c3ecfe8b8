"""CPU tests: the FP64 oracle and the host logic pinned against the reference's own
known-answer tests (Frequency/test.jl, Lattice/test.jl, Action/test.jl, Flow/test.jl),
committed golden vectors and closed forms."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from pffrg import action as A
from pffrg import lattice as LT
from pffrg import mesh as M

GOLD = Path(__file__).resolve().parent / "golden"


def fake_problem(m, name="square", size=2):
    l, r = LT.build(name, size, [[1.0]])
    return l, r, O.Oracle(r, m)


def uniform_mesh(lst):
    return M.Mesh(lst, lst, lst, lst, lst, lst)


# ------------------------------------------------------------------ Frequency/test.jl
def test_mesh_endpoints():
    rng = np.random.default_rng(1)
    for _ in range(5):
        lst = M.get_mesh(float(rng.random()) * 0.9 + 0.05, 1.0, 30, 0.4)
        assert lst[0] == 0.0 and abs(lst[-1] - 1.0) < 1e-12 and len(lst) == 31
        assert np.all(np.diff(lst) > 0)


def test_mesh_closed_form():
    m = M.get_mesh(2.0, 50.0, 10, 0.4)
    nl = 4  # ceil(0.4*10)
    np.testing.assert_allclose(m[1:nl + 1], 2.0 * np.arange(1, nl + 1) / nl, rtol=1e-15)
    xi = (50.0 / 2.0) ** (1.0 / 6)
    np.testing.assert_allclose(m[nl + 1:], 2.0 * xi ** np.arange(1, 7), rtol=1e-14)


@pytest.mark.parametrize("ch", [0, 1, 2])
def test_buffer_kernels(ch):
    rng = np.random.default_rng(2)
    lst = M.get_mesh(float(rng.random()) * 0.9 + 0.05, 1.0, 30, 0.4)
    _, _, orc = fake_problem(uniform_mesh(lst))
    w, v, vp = rng.random(3)
    inf = np.inf
    assert orc.get_buffer(ch, w, v, vp)[0][4] == 4
    assert orc.get_buffer(ch, w, inf, vp)[0][4] == 3
    assert orc.get_buffer(ch, w, v, inf)[0][4] == 2
    assert orc.get_buffer(ch, w, inf, inf)[0][4] == 1
    assert orc.get_buffer(ch, inf, inf, inf)[0][4] == 0


def test_buffer_flags():
    """Sign folding of get_buffer_{s,t,u} (buffer.jl:88-199), flags = (exchange, map, sgn_μ, sgn_ν)."""
    lst = M.get_mesh(0.5, 1.0, 30, 0.4)
    _, _, orc = fake_problem(uniform_mesh(lst))
    f = lambda ch, w, v, vp: tuple(orc.get_buffer(ch, w, v, vp)[0][:4])
    assert f(0, -0.1, 0.2, 0.3) == (1, 0, 0, 0)
    assert f(0, 0.1, -0.2, 0.3) == (1, 1, 1, 0)
    assert f(0, 0.1, 0.2, -0.3) == (0, 1, 0, 1)
    assert f(0, -0.1, -0.2, -0.3) == (0, 0, 1, 1)
    assert f(1, -0.1, 0.2, 0.3) == (0, 0, 0, 0)
    assert f(1, 0.1, -0.2, -0.3) == (0, 0, 1, 1)
    assert f(2, -0.1, 0.2, 0.3) == (1, 0, 0, 0)
    assert f(2, 0.1, -0.2, 0.3) == (1, 1, 0, 1)
    assert f(2, 0.1, -0.2, -0.3) == (1, 0, 0, 0)


def test_get_param_cases():
    lst = M.get_mesh(0.5, 1.0, 10, 0.4)
    assert O.get_param(0.0, lst) == (0, 0, 1.0, 0.0)
    assert O.get_param(lst[3], lst) == (3, 3, 1.0, 0.0)
    assert O.get_param(5.0, lst) == (10, 10, 1.0, 0.0)
    x = 0.3 * lst[4] + 0.7 * lst[5]
    lo, hi, wl, wh = O.get_param(x, lst)
    assert (lo, hi) == (4, 5)
    assert abs(wl - 0.3) < 1e-12 and abs(wh - 0.7) < 1e-12


# ------------------------------------------------------------------ Flow/test.jl
@pytest.mark.parametrize("mode", [0, 1])
def test_quadrature_golden(mode):
    g = json.loads((GOLD / "quadrature.json").read_text())
    ref = np.array(g["values"])
    got = O.test_integrate(mode, 100, g["a"], g["b"], 100, 1e-8, 1e-8)
    np.testing.assert_allclose(got, ref, rtol=1e-8)


def test_quadgk_closed_forms():
    assert abs(O.test_quadgk(0, 0.7) - np.sqrt(np.pi)) < 1e-9
    ref = 0.25 * (np.pi / 2 - np.arctan(0.5)) + 0.25 * (1 - 0.5 / (1 + 0.25) * 0 - 0.5 / 1.25)
    x = 0.5
    exact = 0.5 * (np.pi / 2 - np.arctan(x) - x / (1 + x * x))
    assert abs(O.test_quadgk(1, 0.5) - exact) < 1e-9
    assert abs(O.test_quadgk(2, -1.3) - np.exp(-1.3)) < 1e-9


# ------------------------------------------------------------------ Action/test.jl
def _multilinear_action(r, m, rng):
    a = A.Action.empty(r.n_sites, m)
    ns = r.n_sites
    coef = rng.random((8, 3, 2, ns))

    def f(ch, comp, w, v, vp):
        c = coef[:, ch, comp][:, :, None, None, None] if False else coef[:, ch, comp]
        return (c[0] * w * v * vp + c[1] * w * v + c[2] * w * vp + c[3] * v * vp + c[4] * w + c[5] * v + c[6] * vp
                + c[7])

    for comp in range(2):
        for ch in range(3):
            Om, nu = (m.Ot, m.nt) if ch == 1 else (m.Os, m.ns)
            vm = nu[-1]
            q1 = a.view(comp, ch, "q1")
            q21, q22, q3 = a.view(comp, ch, "q2_1"), a.view(comp, ch, "q2_2"), a.view(comp, ch, "q3")
            for iw, w in enumerate(Om):
                q1[:, iw] = f(ch, comp, w, vm, vm)
                for iv, v in enumerate(nu):
                    q21[:, iw, iv] = f(ch, comp, w, v, vm)
                    q22[:, iw, iv] = f(ch, comp, w, vm, v)
                    for ivp, vp in enumerate(nu):
                        q3[:, iw, iv, ivp] = f(ch, comp, w, v, vp)
    return a, f


def test_multilinear_interpolation():
    rng = np.random.default_rng(3)
    m = M.launch_mesh(1.0, 8, 7, 6, 5)
    l, r, orc = fake_problem(m, "square", 3)
    a, f = _multilinear_action(r, m, rng)
    for ch in range(3):
        Om, nu = (m.Ot, m.nt) if ch == 1 else (m.Os, m.ns)
        for _ in range(20):
            w, v, vp = rng.random(3) * np.array([Om[-1], nu[-1], nu[-1]])
            site = int(rng.integers(r.n_sites))
            comp = int(rng.integers(2))
            exact = f(ch, comp, w, v, vp)[site]
            assert abs(orc.get_vertex(a, comp, ch, w, v, vp, site) - exact) <= 1e-10 * max(1, abs(exact))
            close = lambda x, y: abs(x - y) <= 1e-12 * max(1.0, abs(y))
            assert close(orc.get_vertex(a, comp, ch, w, np.inf, vp, site), f(ch, comp, w, nu[-1], vp)[site])
            assert close(orc.get_vertex(a, comp, ch, w, v, np.inf, site), f(ch, comp, w, v, nu[-1])[site])
            assert close(orc.get_vertex(a, comp, ch, w, np.inf, np.inf, site), f(ch, comp, w, nu[-1], nu[-1])[site])
            assert orc.get_vertex(a, comp, ch, Om[-1] * 1.01, v, vp, site) == 0.0
        # on-mesh reproduction
        iw, iv, ivp = 3, 2, 4
        for comp in range(2):
            got = [orc.get_vertex(a, comp, ch, Om[iw], nu[iv], nu[ivp], s) for s in range(r.n_sites)]
            np.testing.assert_array_equal(got, a.view(comp, ch, "q3")[:, iw, iv, ivp])


def test_vectorised_equals_sequential():
    """get_Γ_avx! (all sites) equals bare + Σ channel interpolations with the fold flags."""
    rng = np.random.default_rng(4)
    m = M.launch_mesh(1.0, 8, 7, 6, 5)
    l, r, orc = fake_problem(m, "kagome", 3)
    a, f = _multilinear_action(r, m, rng)
    a.bare[:] = rng.random(a.bare.shape)
    # positive frequencies: no flags, Γ = bare + s + t + u
    fr = np.array([1.1, 2.2, 0.4, 0.9, 0.3, 1.7, 2.0, 0.2, 0.8])
    got = orc.get_gamma_avx(a, fr)
    for comp in range(2):
        exp = a.bare[comp] + f(0, comp, *fr[0:3]) + f(1, comp, *fr[3:6]) + f(2, comp, *fr[6:9])
        np.testing.assert_allclose(got[comp], exp, rtol=1e-12)
    # negative ν in the s channel: exchange + map to u + density sign (buffer.jl:108-114)
    fr2 = fr.copy()
    fr2[1] = -fr2[1]
    got2 = orc.get_gamma_avx(a, fr2, mask=1)
    x = r.exchange
    np.testing.assert_allclose(got2[0], a.bare[0] + f(2, 0, fr[0], fr[1], fr[2])[x], rtol=1e-12)
    np.testing.assert_allclose(got2[1], a.bare[1] - f(2, 1, fr[0], fr[1], fr[2])[x], rtol=1e-12)


def test_sigma_interpolation():
    rng = np.random.default_rng(5)
    m = M.launch_mesh(1.0, 10, 6, 6, 5)
    _, _, orc = fake_problem(m)
    Sig = rng.random(m.num_sigma)
    Sig[0] = 0.0
    for k in (0, 3, 7):
        assert orc.get_sigma(Sig, m.sigma[k]) == Sig[k]
        assert orc.get_sigma(Sig, -m.sigma[k]) == -Sig[k]
    outer = m.sigma[-1] + 0.37
    assert abs(orc.get_sigma(Sig, outer) - Sig[-1] * m.sigma[-1] / outer) < 1e-15


def test_bare_norm():
    # Action/test.jl:319-328: square L=5, J=1 -> |Γspin.bare| = 0.25, |Γdens.bare| = 0
    l, r = LT.build("square", 5, [[1.0]])
    bare = LT.bare_couplings(l, r)
    assert abs(np.linalg.norm(bare[0]) - 0.25) < 1e-14
    assert np.linalg.norm(bare[1]) == 0.0


# ------------------------------------------------------------------ Lattice/test.jl
@pytest.mark.parametrize("name,count", [("square", 145), ("honeycomb", 109), ("kagome", 163), ("triangular", 217),
                                        ("cubic", 833), ("fcc", 2057), ("bcc", 1241), ("hyperhoneycomb", 319),
                                        ("hyperkagome", 415), ("pyrochlore", 1029), ("diamond", 525)])
def test_lattice_site_counts(name, count):
    assert len(LT.get_sites(8, LT.get_unitcell(name))) == count


@pytest.mark.parametrize("name", ["square", "kagome", "cubic"])
def test_euclidean_truncation(name):
    # Lattice/test.jl:46-53: Euclidean cut equals the sites of a larger bond-cut lattice inside the sphere
    uc = LT.get_unitcell(name)
    nn = np.linalg.norm(LT.get_vec(LT.get_sites(1, uc)[1], uc))
    for j in (2, 3):
        e = LT.get_sites(j, uc, euclidean=True)
        big = LT.get_sites(int(np.ceil(2.5 * j)), uc)
        inside = [s for s in big if np.linalg.norm(LT.get_vec(s, uc)) <= nn * j + 1e-12]
        assert len(e) == len(inside)


@pytest.mark.parametrize("name,L,ns,rows", [("square", 3, 6, 55), ("square", 4, 9, 136), ("square", 5, 12, 286),
                                            ("square", 6, 16, 541), ("square", 7, 20, 940), ("kagome", 5, 19, 667),
                                            ("pyrochlore", 4, 18, 885)])
def test_reduced_counts(name, L, ns, rows):
    l, r = LT.build(name, L, [[1.0]])
    assert r.n_sites == ns
    assert len(r.overlap_csr()[1]) == rows
    assert r.mult.sum() == len(l.ints)
    assert r.exchange[0] == 0
    assert np.all(r.exchange[r.exchange] == np.arange(ns))


@pytest.mark.parametrize("L", [3, 5, 6])
def test_square_reduction_closed_form(L):
    """Square lattice: irreducible sites are x >= y >= 0, x + y <= L, with D4 orbit sizes."""
    l, r = LT.build("square", L, [[1.0]])
    pts = sorted((int(round(v[0])), int(round(v[1]))) for v in np.abs(r.vecs))
    exp = sorted((x, y) for x in range(L + 1) for y in range(x + 1) if x + y <= L)
    assert [tuple(sorted(p, reverse=True)) for p in pts] == [tuple(sorted(p, reverse=True)) for p in exp] or \
        sorted(tuple(sorted(p, reverse=True)) for p in pts) == exp
    for v, mu in zip(np.abs(r.vecs), r.mult):
        x, y = sorted((int(round(v[0])), int(round(v[1]))), reverse=True)
        orbit = 1 if x == 0 else (4 if (y == 0 or x == y) else 8)
        assert mu == orbit
    # overlap multiplicities count the lattice sites in range of both the origin and site i
    for i, irr in enumerate(r.irreducible):
        in_range = sum(1 for j in range(len(l.ints)) if np.abs(l.vecs[j] - l.vecs[irr]).sum() <= L + 1e-9)
        assert r.overlap[i][:, 2].sum() == in_range


# ------------------------------------------------------------------ flow RHS closed form
def test_bare_only_rhs_closed_form():
    """Σ = dΣ = 0 and all channels zero: every interpolated vertex is the bare vector, so each
    Katanin element is a product of bare couplings times ∫ p_kat dv (scipy golden values):
    s: spin 2b_i², dens -3b_i²;  u: spin -2b_i², dens -3b_i²;  t: spin 2(2S) P_i/16 with P_i
    the number of two-step nearest-neighbour paths from the origin to site i (b = 1/4)."""
    gold = json.loads((GOLD / "bubble_integrals.json").read_text())
    L = 1.0
    m = M.launch_mesh(1.0, 8, 6, 6, 5)
    l, r = LT.build("square", 3, [[1.0]])
    orc = O.Oracle(r, m)
    a = A.Action.empty(r.n_sites, m)
    a.bare[:] = LT.bare_couplings(l, r)
    da = A.Action.empty(r.n_sites, m)
    orc.compute_corrs(L, a, da, True)
    b = a.bare[0]
    nb = [np.array(v) for v in ([1, 0], [-1, 0], [0, 1], [0, -1])]
    P = np.array([sum(1 for k in nb if np.abs(k - v[:2]).sum() == 1) for v in r.vecs])
    # committed scipy values of ∫ p_kat dv at W = 0 (the first Ω mesh point) for three Λ
    n_gold = 0
    for g in gold:
        if g["W"] != 0.0:
            continue
        Lg, ch = g["L"], g["ch"]
        orc.compute_corrs(Lg, a, da, True)
        out = A.Action.empty(r.n_sites, m)
        orc.compute_element(O.P_KAT, Lg, a, da, ch, 4, 0, 1, 2, 10, out)
        qs = out.view(0, ch, "q3")[:, 0, 1, 2]
        ref = {0: 2 * b * b, 1: 2 * (2 * 0.5) * P / 16.0, 2: -2 * b * b}[ch] * g["integral"]
        np.testing.assert_allclose(qs, ref, rtol=1e-4, atol=1e-4 * np.max(np.abs(ref)))
        n_gold += 1
    assert n_gold == 9
    orc.compute_corrs(L, a, da, True)
    # evaluate at mesh bosonic frequencies and compare with a direct scipy integral of p
    from scipy.integrate import quad
    sys_ok = 0
    for ch in range(3):
        Om = m.Ot if ch == 1 else m.Os
        for w1 in (0, 2, 4):
            W = Om[w1]
            out = A.Action.empty(r.n_sites, m)
            orc.compute_element(O.P_KAT, L, a, da, ch, 4, w1, 1, 2, 10, out)
            q = out.view(0, ch, "q3")[:, w1, 1, 2], out.view(1, ch, "q3")[:, w1, 1, 2]

            def p(v):
                x1, x2 = {0: (v + W / 2, W / 2 - v), 1: (v + W / 2, v - W / 2), 2: (v - W / 2, v + W / 2)}[ch]
                G0 = lambda w: 0.0 if abs(w) <= 1e-8 else -np.expm1(-w * w / L ** 2) / w
                S0 = lambda w: 0.0 if abs(w) <= 1e-8 else np.exp(-w * w / L ** 2) * 2 * w / L ** 3
                return (S0(x1) * G0(x2) + S0(x2) * G0(x1)) / (2 * np.pi)

            Ip = sum(quad(p, lo, hi, epsabs=1e-14, epsrel=1e-12, limit=400)[0]
                     for lo, hi in ((-np.inf, -W / 2 - 1), (-W / 2 - 1, W / 2 + 1), (W / 2 + 1, np.inf)))
            if ch == 0:
                exp_s, exp_d = 2 * b * b * Ip, -3 * b * b * Ip
            elif ch == 2:
                exp_s, exp_d = -2 * b * b * Ip, -3 * b * b * Ip
            else:
                exp_s, exp_d = 2 * (2 * 0.5) * P / 16.0 * Ip, 0 * b
            scale = max(np.max(np.abs(exp_s)), 1e-30)
            assert np.max(np.abs(q[0] - exp_s)) <= 1e-4 * scale
            assert np.max(np.abs(q[1] - exp_d)) <= 1e-4 * max(np.max(np.abs(exp_d)), scale)
            sys_ok += 1
    assert sys_ok == 9


def test_symmetrize_invariants():
    rng = np.random.default_rng(6)
    m = M.launch_mesh(1.0, 6, 5, 5, 5)
    l, r, orc = fake_problem(m, "kagome", 3)
    a = A.Action.empty(r.n_sites, m)
    a.gamma[:] = rng.standard_normal(a.gamma.size)
    lower = a.copy()
    orc.symmetrize(a)
    x = r.exchange
    nn = m.num_nu
    for comp in range(2):
        for ch in range(3):
            q, q0 = a.view(comp, ch, "q3"), lower.view(comp, ch, "q3")
            for v in range(nn):
                for vp in range(nn):
                    if v < vp:
                        src = q0[:, :, vp, v] if ch == 2 else q0[x][:, :, vp, v]
                        np.testing.assert_array_equal(q[:, :, v, vp], src)
                    else:
                        np.testing.assert_array_equal(q[:, :, v, vp], q0[:, :, v, vp])


def test_synthetic_generator_deterministic():
    m = M.launch_mesh(1.0, 6, 5, 5, 5)
    bare = np.zeros((2, 3))
    a1 = A.synthetic_action(3, m, bare, seed=20261203)
    a2 = A.synthetic_action(3, m, bare, seed=20261203)
    np.testing.assert_array_equal(a1.gamma, a2.gamma)
    z = A.normals(1, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def _symmetric_action(r, m, bare, seed):
    """A vertex with the stored ν <-> ν' symmetry the reference's symmetrize!/symmetrize_add_to!
    assume (action_su2.jl:141-250): q3[i,w,v,v'] = q3[x(i),w,v',v], q2_2[i] = q2_1[x(i)],
    q1[i] = q1[x(i)] in the s and t channels (x = exchange), the same without x in u."""
    ns = r.n_sites
    x = np.asarray(r.exchange)
    a = A.synthetic_action(ns, m, bare, seed=seed)
    for comp in range(2):
        for ch in range(3):
            ex = x if ch != 2 else np.arange(ns)
            q1 = a.view(comp, ch, "q1")
            q1[:] = 0.5 * (q1 + q1[ex])
            a.view(comp, ch, "q2_2")[:] = a.view(comp, ch, "q2_1")[ex]
            q3 = a.view(comp, ch, "q3")
            for v in range(m.num_nu):
                q3[:, :, v, v] = 0.5 * (q3[:, :, v, v] + q3[ex][:, :, v, v])
    return a


@pytest.mark.parametrize("name,size", [("square", 2), ("pyrochlore", 3)])
def test_katanin_pass_preserves_vertex_symmetry(name, size):
    """Invariant of the one-loop flow: for an input vertex with the stored ν <-> ν' symmetry,
    the Katanin derivative computed on the full (v, v') grid has the same symmetry, up to the
    quadrature tolerance (the reference computes half the grid and copies the rest).  A sign
    or argument error in any frequency-transfer formula of flow_su2_{s,t,u}.jl breaks it at the
    1e-2..1 level; the intact oracle meets it to ~1e-4."""
    m = M.launch_mesh(1.0, 6, 6, 6, 5)
    l, r = LT.build(name, size, [[1.0]])
    bare = LT.bare_couplings(l, r)
    a = _symmetric_action(r, m, bare, seed=7)
    orc = O.Oracle(r, m)
    orc.symmetrize(a)
    da = A.Action.empty(r.n_sites, m)
    orc.compute_dsigma(1.0, a, da)
    orc.compute_corrs(1.0, a, da, True)
    out = A.Action.empty(r.n_sites, m)
    orc.run_pass(O.P_KAT, 1.0, a, da, out, True, 10)
    x = np.asarray(r.exchange)
    for comp in range(2):
        for ch in range(3):
            ex = x if ch != 2 else np.arange(r.n_sites)
            q3 = out.view(comp, ch, "q3")
            assert np.abs(q3 - q3[ex].transpose(0, 1, 3, 2)).max() <= 2e-3 * np.abs(q3).max()
            q21, q22 = out.view(comp, ch, "q2_1"), out.view(comp, ch, "q2_2")
            assert np.abs(q22 - q21[ex]).max() <= 2e-3 * np.abs(q21).max()


# ---------------------------------------------------------------- correlations (compute_χ!)
def test_hcubature_polynomial_exactness():
    """The degree-7 Genz-Malik rule integrates every x^p y^q with p + q <= 7 exactly (and the
    embedded degree-5 rule agrees below degree 6, so those converge on one region)."""
    lo, hi = [-0.3, 0.2], [1.1, 0.9]
    for p in range(8):
        for q in range(8 - p):
            I, nr = O.test_hcubature(0, p, q, lo, hi, 1e-14, 1e-14)
            ex = (hi[0] ** (p + 1) - lo[0] ** (p + 1)) / (p + 1) * (hi[1] ** (q + 1) - lo[1] ** (q + 1)) / (q + 1)
            assert abs(I - ex) <= 1e-13 * max(1.0, abs(ex)), (p, q, I, ex)
            if p + q <= 5:
                assert nr == 1


def test_hcubature_gaussian():
    import math
    I, _ = O.test_hcubature(1, 0, 0, [-1.0, 0.0], [2.0, 1.0], 1e-12, 1e-10)
    ex = math.pi / 4 * (math.erf(2.0) + math.erf(1.0)) * math.erf(1.0)
    assert abs(I - ex) <= 1e-10 * ex


def _chi_bare_expected(prob, L, W):
    """χ of a frequency-independent (bare) vertex with Σ = 0: the double integral factorises
    into (∫ G0(v-W/2) G0(v+W/2) dv)^2 (correlation_su2.jl:14-55), plus the site-0 propagator
    term (2S)/(4π) ∫ G0 G0 dv (correlation_su2.jl:78-82)."""
    import math
    from scipy.integrate import quad
    G0 = lambda w: 0.0 if abs(w) <= 1e-8 else -math.expm1(-w * w / L ** 2) / w
    P = lambda v: G0(v - 0.5 * W) * G0(v + 0.5 * W)
    IP = sum(quad(P, lo, hi, epsabs=1e-14, epsrel=1e-12, limit=400)[0]
             for lo, hi in ((-np.inf, -W / 2 - 1), (-W / 2 - 1, W / 2 + 1), (W / 2 + 1, np.inf)))
    S2 = 2.0 * prob.S
    tp2 = (2.0 * math.pi) ** 2
    b = prob.bare
    out = -(S2 ** 2) * b[0] / tp2 * IP * IP
    out[0] += -S2 * (b[0, 0] - b[1, 0]) / (2.0 * tp2) * IP * IP + S2 * IP / (4.0 * math.pi)
    return out


@pytest.mark.parametrize("tol,bound", [((1e-8, 1e-4), 2e-5), ((1e-12, 1e-6), 1e-7)])
def test_chi_bare_vertex_closed_form(tol, bound):
    from pffrg import flow as F
    prob = F.Problem.build("square", 2, [[1.0]], initial=1.0, num_sigma=8, num_Omega=6, num_nu=6, num_chi=5,
                           normalize=False)
    a = A.Action.empty(prob.r.n_sites, prob.mesh)
    a.bare[:] = prob.bare
    orc = O.Oracle(prob.r, prob.mesh)
    chi = orc.compute_chi(1.0, a, tol=tol)
    for k, W in enumerate(prob.mesh.chi):
        exp = _chi_bare_expected(prob, 1.0, W)
        assert np.abs(chi[:, k] - exp).max() <= bound * np.abs(exp).max(), (W, chi[:, k], exp)


def test_mlp_reference_backprop_matches_finite_differences():
    """The numpy restatement of the warm-start network's back-propagation (FP64, no rounding)
    against central finite differences of its loss."""
    import mlp_ref as R
    rng = np.random.default_rng(3)
    dims = R.dims_of(2, 7, 4, 5)
    n = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(len(dims) - 1))
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((6, 2))
    t = rng.standard_normal((6, 5))
    _, g = R.loss_grad(p, dims, x, t, bf=False)
    fd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1e-6
        fd[i] = (R.loss_grad(p + e, dims, x, t, bf=False)[0] - R.loss_grad(p - e, dims, x, t, bf=False)[0]) / 2e-6
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_coupling_samples_deterministic():
    from pffrg import learn
    s = learn.coupling_samples(80, seed=5)
    assert s.shape == (80, 2)
    assert np.all((s[:, 0] >= -0.5) & (s[:, 0] <= 0.5)) and np.all((s[:, 1] >= 0.5) & (s[:, 1] <= 5.0))
    np.testing.assert_array_equal(s, learn.coupling_samples(80, seed=5))


def test_shared_exp_expm1_accuracy():
    """pffrg_math.h's exp/expm1 (used by the device code and the oracle alike, so that both
    produce the same bits) against the 80-bit long-double library: within 3 ulp."""
    lib = O.lib()
    rng = np.random.default_rng(17)
    xs = np.concatenate([rng.uniform(-45, 5, 20000), rng.uniform(-1e-3, 1e-3, 5000), rng.uniform(-1, 1, 5000),
                         rng.uniform(-700, 700, 5000), [0.0, 1e-300, -1e-20, -40.0, -40.5, 0.34657359, -0.34657359]])
    worst = 0.0
    for x in xs:
        for f, ref in ((lib.orc_pf_expm1, np.expm1), (lib.orc_pf_exp, np.exp)):
            got = f(float(x))
            r = ref(np.longdouble(x))
            if r == 0:
                assert got == 0.0
                continue
            ulp = np.spacing(np.abs(np.float64(r)))
            worst = max(worst, float(abs(np.longdouble(got) - r) / np.longdouble(ulp)))
    assert worst <= 3.0, worst
    assert lib.orc_pf_expm1(-50.0) == -1.0 and lib.orc_pf_exp(-800.0) == 0.0
