"""Mesh rescanning on the CPU: the host scan heuristic of the library (pffrg_scan, a host-only
entry point) against the oracle's restatement and against closed forms, and the oracle's
resampling and cut extraction against what the mathematics fixes (Action.jl:369-557,
channel.jl:416-470).

Pins:
- scan, peak heuristic: a triangular cut whose 75 % point is known in closed form;
- scan, origin heuristic: the returned extent puts the relative first step Δ inside (p1, p2)
  while the previous 2 % iterate did not; an in-window cut keeps the current extent;
- resampling: exact for functions affine in each frequency (multilinear interpolation),
  clamped beyond the old mesh, Σ by get_Σ's 1/w tail, and the identity on unchanged meshes;
- cuts: numpy's first argmax in column-major order and slicing of the same array.
Tolerances: exact where the arithmetic is the same (scan, identity, cuts), 1e-12 relative
for the affine reproductions (rounding of the weights)."""
import numpy as np
import pytest

import oracle as O
from pffrg import flow as F
from pffrg import mesh as M
from pffrg import native
from pffrg.action import Action

from helpers import small_problem
from oracle_ctx import OracleContext

P_OM = (M.P_OMEGA_S[0], M.P_OMEGA_S[1], M.P_OMEGA_S[2], M.P_OMEGA_S[3], M.P_OMEGA_S[4])

needs_lib = pytest.mark.skipif(not native.LIB_PATH.exists(), reason="libpffrg.so not built")


def _mesh(linear, upper=50.0, num=24, p=0.4):
    return M.get_mesh(linear, upper, num, p)


def _delta(x, y, p):
    """Δ of get_mesh(δ, ...) as the origin heuristic measures it (linear interpolation)."""
    return lambda d: (lambda xp: abs(np.interp(xp[1], x, y) - y[0]) / max(abs(np.interp(xp[1], x, y)), abs(y[0])))(
        M.get_mesh(d, x[-1], len(x) - 1, p[0]))


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for k in range(40):
        x = _mesh(float(rng.uniform(0.5, 10.0)), float(rng.uniform(40.0, 200.0)), int(rng.integers(8, 40)))
        kind = k % 4
        if kind == 0:
            y = np.exp(-x / rng.uniform(0.2, 5.0))
        elif kind == 1:
            x0 = rng.uniform(0.5, 8.0)
            y = x * np.exp(-x / x0) + rng.uniform(0.0, 0.3)
        elif kind == 2:
            y = 1.0 / (1.0 + (x / rng.uniform(0.3, 4.0)) ** 2) * rng.choice([-1.0, 1.0])
        else:
            y = rng.normal(size=len(x)).cumsum()
        p = (0.4, 0.02, 0.05, float(rng.uniform(0.01, 0.1)), float(rng.uniform(5.0, 60.0)))
        out.append((x, y, p))
    return out


@needs_lib
def test_scan_equals_oracle_scan():
    for x, y, p in _cases():
        assert native.scan(x, y, p) == O.scan(x, y, p)


@needs_lib
def test_scan_peak_heuristic_closed_form():
    # triangle: rises linearly to its peak at mesh point 6, then falls with slope -1/w; the cut
    # is down to 75 % of the peak from x6 + w/4 on, and zero at the origin (origin rule idle)
    x = _mesh(3.0, 60.0, 30, 0.4)
    n_lin = int(np.ceil(0.4 * (len(x) - 1)))
    pk, w = 6, 5.0
    y = np.where(x <= x[pk], x / x[pk], 1.0 - (x - x[pk]) / w)
    i75 = int(np.argmax(x >= x[pk] + 0.25 * w))
    da = max(x[i75], n_lin * x[pk] / np.ceil(0.5 * n_lin))
    db = n_lin * x[1]
    p = (0.4, 0.02, 0.05, 0.001, 1000.0)
    assert native.scan(x, y, p) == min(da, db)
    # the bounds p3 * num_lin and p4 clamp it
    assert native.scan(x, y, (0.4, 0.02, 0.05, 0.001, 0.5)) == 0.5
    assert native.scan(x, y, (0.4, 0.02, 0.05, 1.0, 1000.0)) == n_lin * 1.0


@needs_lib
def test_scan_origin_heuristic_window():
    p = (0.4, 0.02, 0.05, 1e-6, 1e6)
    n_checked = 0
    for s, lin in ((0.5, 8.0), (2.0, 1.0), (0.2, 3.0), (5.0, 0.4)):
        x = _mesh(lin, 80.0, 24, 0.4)
        y = np.exp(-x / s)
        d = native.scan(x, y, p)
        D = _delta(x, y, p)
        D0 = D(int(np.ceil(0.4 * 24)) * x[1])
        if p[1] < D0 < p[2]:
            assert d == int(np.ceil(0.4 * 24)) * x[1]
            continue
        assert p[1] < D(d) < p[2], (s, lin, d, D(d))
        prev = d / 0.98 if D0 >= p[2] else d / 1.02
        assert not (p[1] < D(prev) < p[2])
        n_checked += 1
    assert n_checked >= 2


@needs_lib
def test_scan_rejects_bad_arguments():
    with pytest.raises(native.PffrgError):
        native.scan(np.zeros(3), np.zeros(4), P_OM)


def _affine_action(prob, rng):
    """Vertex entries affine in each of (Ω, ν, ν') on the old meshes, Σ linear in σ."""
    m, ns = prob.mesh, prob.r.n_sites
    a = Action.empty(ns, m, prob.S)
    a.bare[:] = rng.normal(size=a.bare.shape)
    coef = {}
    for comp in range(2):
        for ch in range(3):
            Om, nu = (m.Ot, m.nt) if ch == 1 else (m.Os, m.ns)
            c = rng.normal(size=(4, ns))
            coef[comp, ch] = c
            a.view(comp, ch, "q1")[:] = c[0][:, None] + c[1][:, None] * Om[None, :]
            for q, j in (("q2_1", 2), ("q2_2", 3)):
                a.view(comp, ch, q)[:] = (c[0][:, None, None] + c[1][:, None, None] * Om[None, :, None] +
                                          c[j][:, None, None] * nu[None, None, :])
            a.view(comp, ch, "q3")[:] = (c[0][:, None, None, None] + c[1][:, None, None, None] * Om[None, :, None, None] +
                                         c[2][:, None, None, None] * nu[None, None, :, None] +
                                         c[3][:, None, None, None] * nu[None, None, None, :])
    a.Sigma[:] = 0.3 * m.sigma
    return a, coef


def test_oracle_resample_reproduces_affine_vertex():
    prob = small_problem("square", 2, 8, 1.0)
    rng = np.random.default_rng(3)
    a, coef = _affine_action(prob, rng)
    m_old = prob.mesh
    # new meshes: finer linear parts and longer tails (points beyond the old range clamp)
    m_new = M.launch_mesh(0.6, m_old.num_sigma - 1, m_old.num_Omega - 1, m_old.num_nu - 1)
    m_new = M.Mesh(m_new.sigma * 1.3, m_new.Os * 1.3, m_new.ns * 1.3, m_new.Ot * 1.3, m_new.nt * 1.3, m_new.chi)
    orc = O.Oracle(prob.r, m_old)
    out = Action.empty(prob.r.n_sites, m_new, prob.S)
    orc.resample(a, m_new, out)
    for comp in range(2):
        for ch in range(3):
            Oo, no = (m_old.Ot, m_old.nt) if ch == 1 else (m_old.Os, m_old.ns)
            On, nnw = (m_new.Ot, m_new.nt) if ch == 1 else (m_new.Os, m_new.ns)
            Oc, nc = np.minimum(On, Oo[-1]), np.minimum(nnw, no[-1])  # get_param clamps
            c = coef[comp, ch]
            exp3 = (c[0][:, None, None, None] + c[1][:, None, None, None] * Oc[None, :, None, None] +
                    c[2][:, None, None, None] * nc[None, None, :, None] + c[3][:, None, None, None] * nc[None, None, None, :])
            got = out.view(comp, ch, "q3")
            assert np.max(np.abs(got - exp3)) <= 1e-12 * np.max(np.abs(exp3))
            exp1 = c[0][:, None] + c[1][:, None] * Oc[None, :]
            assert np.max(np.abs(out.view(comp, ch, "q1") - exp1)) <= 1e-12 * np.max(np.abs(exp1))
            exp21 = c[0][:, None, None] + c[1][:, None, None] * Oc[None, :, None] + c[2][:, None, None] * nc[None, None, :]
            assert np.max(np.abs(out.view(comp, ch, "q2_1") - exp21)) <= 1e-12 * np.max(np.abs(exp21))
    # Σ: linear inside the old mesh, σ_max Σ_max / w beyond it
    so = m_old.sigma
    inside = m_new.sigma <= so[-1]
    np.testing.assert_allclose(out.Sigma[inside], 0.3 * m_new.sigma[inside], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(out.Sigma[~inside], so[-1] * (0.3 * so[-1]) / m_new.sigma[~inside], rtol=1e-15)
    assert np.array_equal(out.bare, a.bare)


def test_oracle_resample_identity_on_same_meshes():
    prob = small_problem("square", 2, 6, 1.0)
    rng = np.random.default_rng(5)
    a = Action.empty(prob.r.n_sites, prob.mesh, prob.S)
    a.gamma[:] = rng.normal(size=a.gamma.shape)
    a.Sigma[:] = rng.normal(size=a.Sigma.shape)
    a.Sigma[0] = 0.0
    out = Action.empty(prob.r.n_sites, prob.mesh, prob.S)
    O.Oracle(prob.r, prob.mesh).resample(a, prob.mesh, out)
    assert np.array_equal(out.gamma, a.gamma) and np.array_equal(out.Sigma, a.Sigma)


def test_oracle_scan_cuts_match_numpy():
    prob = small_problem("kagome", 2, 6, 1.0)
    rng = np.random.default_rng(8)
    a = Action.empty(prob.r.n_sites, prob.mesh, prob.S)
    a.gamma[:] = rng.normal(size=a.gamma.shape)
    # a planted tie: the first occurrence in column-major order must win
    q3 = a.view(1, 2, "q3")
    big = float(np.max(np.abs(a.gamma))) + 1.0
    q3[1, 2, 3, 4] = -big
    q3[0, 2, 3, 5] = big
    idx, cuts = O.Oracle(prob.r, prob.mesh).scan_cuts(a)
    nO, nn = prob.mesh.num_Omega, prob.mesh.num_nu
    for k in range(6):
        q = a.view(k // 3, k % 3, "q3")
        j = int(np.argmax(np.abs(q).ravel(order="F")))
        pos = np.unravel_index(j, q.shape, order="F")
        assert tuple(idx[k]) == tuple(int(t) for t in pos)
        i, w, v, vp = pos
        assert np.array_equal(cuts[k, :nO], q[i, :, v, vp])
        xs = np.arange(nn)
        assert np.array_equal(cuts[k, nO:nO + nn], q[i, w, xs, xs] - q[i, w, -1, -1])
        assert np.array_equal(cuts[k, nO + nn:nO + 2 * nn], q[i, w, :, vp] - q[i, w, -1, vp])
        assert np.array_equal(cuts[k, nO + 2 * nn:], q[i, w, v, :] - q[i, w, v, -1])
    assert tuple(idx[5]) == (1, 2, 3, 4)


@needs_lib
def test_rescanned_mesh_above_scan_scale_is_proportional():
    # above lins[0] = 5 the reference only rescales the meshes with Λ (Action.jl:501-507)
    prob = small_problem("square", 2, 6, 8.0)
    octx = OracleContext(prob)
    m = F.rescanned_mesh(octx, 6.0, F.A)
    ref = M.launch_mesh(6.0, prob.mesh.num_sigma - 1, prob.mesh.num_Omega - 1, prob.mesh.num_nu - 1,
                        prob.mesh.num_chi - 1)
    for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
        assert np.array_equal(getattr(m, f), getattr(ref, f)), f


@needs_lib
def test_oracle_flow_rescans_meshes():
    # a short oracle-driven flow below the scan scale: meshes change every accepted step and stay valid
    prob = small_problem("square", 2, 6, 4.0)
    octx = OracleContext(prob)
    res = F.launch_flow(octx, prob, prob.bare_action(), 4.0, 2.0, max_steps=3)
    assert res.accepted >= 2 and len(res.meshes) == res.accepted
    for m in res.meshes:
        for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
            arr = getattr(m, f)
            assert arr[0] == 0.0 and np.all(np.diff(arr) > 0)
    assert not np.array_equal(res.meshes[0].Os, prob.mesh.Os)
