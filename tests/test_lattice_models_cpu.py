"""The SU(2) lattice models beside plain Heisenberg (model_breathing.jl:14-182): bonds against
their definitions, the symmetry reduction against the couplings, and the special cases that
reduce to the Heisenberg model.  No reference test covers these models, so these checks pin
them: a breathing model with equal couplings is the Heisenberg model; every site folds onto an
irreducible site with the same coupling to the origin; multiplicities add up to the lattice."""
import numpy as np
import pytest

from pffrg import lattice as LT


def _check_reduction(l, r):
    red = LT.get_reduced(l)
    n = len(l.ints)
    assert int(round(r.mult.sum())) == n
    for j in range(n):
        assert np.allclose(l.bonds[0, j], l.bonds[0, red[j]]), j


@pytest.mark.parametrize("name,size", [("kagome", 3), ("pyrochlore", 3)])
def test_breathing_with_equal_couplings_is_heisenberg(name, size):
    l0, r0 = LT.build(name, size, [[1.0], [0.3]])
    l1, r1 = LT.build(name, size, [[1.0, 1.0], [0.3]], "breathing")
    assert r0.n_sites == r1.n_sites
    assert np.array_equal(r0.exchange, r1.exchange) and np.array_equal(r0.mult, r1.mult)
    assert np.allclose(LT.bare_couplings(l0, r0), LT.bare_couplings(l1, r1))
    assert all(np.array_equal(a, b) for a, b in zip(r0.overlap, r1.overlap))


@pytest.mark.parametrize("name,size", [("kagome", 3), ("pyrochlore", 2)])
def test_breathing_bonds_and_reduction(name, size):
    Ja, Jb = 1.0, 0.4
    l, r = LT.build(name, size, [[Ja, Jb]], "breathing")
    seen = set()
    for i in range(len(l.ints)):
        for j in LT.get_nbs(1, i, l.vecs):
            inside = tuple(l.ints[j][:3]) == tuple(l.ints[i][:3])
            expect = Ja if inside else Jb
            assert l.bonds[i, j, 0, 0] == pytest.approx(expect) and l.bonds[i, j, 2, 2] == pytest.approx(expect)
            seen.add(inside)
    assert seen == {True, False}
    _check_reduction(l, r)
    l0, r0 = LT.build(name, size, [[1.0]])
    assert r.n_sites > r0.n_sites  # the breathing lowers the point group


def test_pyrochlore_breathing_c3():
    J = [[1.0, 0.5, 0.1, 0.05]]
    l, r = LT.build("pyrochlore", 2, J, "pyrochlore-breathing-c3")
    for i in range(len(l.ints)):
        for j in LT.get_nbs(1, i, l.vecs):
            inside = tuple(l.ints[j][:3]) == tuple(l.ints[i][:3])
            plus = {l.ints[i][3], l.ints[j][3]} in ({0, 1}, {2, 3})
            Jc, dc = (J[0][0], J[0][2]) if inside else (J[0][1], J[0][3])
            assert l.bonds[i, j, 1, 1] == pytest.approx(Jc + dc if plus else Jc - dc)
    _check_reduction(l, r)
    # without the C3 breaking it is the breathing model
    l0, r0 = LT.build("pyrochlore", 2, [[1.0, 0.5, 0.0, 0.0]], "pyrochlore-breathing-c3")
    l1, r1 = LT.build("pyrochlore", 2, [[1.0, 0.5]], "breathing")
    assert r0.n_sites == r1.n_sites and np.allclose(LT.bare_couplings(l0, r0), LT.bare_couplings(l1, r1))


def test_model_errors():
    with pytest.raises(AssertionError):
        LT.build("square", 2, [[1.0, 0.5]], "breathing")
    with pytest.raises(AssertionError):
        LT.build("kagome", 2, [[1.0]], "breathing")
    with pytest.raises(NotImplementedError):
        LT.build("triangular", 2, [[1.0]], "no-such-model")
    # model_triangular_dm_c3.jl:21-26: triangular lattice only, three couplings per shell, n <= 3
    with pytest.raises(AssertionError):
        LT.build("triangular", 2, [[1.0]], "triangular-dm-c3")
    with pytest.raises(AssertionError):
        LT.build("square", 2, [[1.0, 0.5, 0.2]], "triangular-dm-c3")
