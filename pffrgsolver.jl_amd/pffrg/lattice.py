"""Lattice graphs and their point-group reduction (host side, built once per run).

Produces the tables the flow kernels consume: the irreducible sites, their
multiplicities, the site-exchange images and the overlap rows of the
t-channel site sum.  Follows src/Lattice/{unitcell,site,bond,build,reduced}.jl
and src/Lattice/model_lib/model_{heisenberg,breathing}.jl of PFFRGSolver.jl; tables are
returned 0-based (the reference's are 1-based).

Site coordinates are matched through a hash of coordinates rounded to 1e-6
instead of the reference's linear search with a 1e-8 norm tolerance
(reduced.jl:257-272); lattice points are at least ~0.35 apart, so both
identify the same sites.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

S3 = math.sqrt(3.0)
S2 = math.sqrt(2.0)

# Unit cells: (basis, primitive vectors, bonds per basis site as (n1, n2, n3, Δbasis)).
# Tables restate src/Lattice/unitcell_lib/*.jl (geometry data must match exactly).
_UNITCELLS = {
    "square": (
        [(0.0, 0.0, 0.0)],
        [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.0)],
        [[(1, 0, 0, 0), (-1, 0, 0, 0), (0, 1, 0, 0), (0, -1, 0, 0)]],
    ),
    "honeycomb": (
        [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0)],
        [(1.5, S3 / 2.0, 0.0), (1.5, -S3 / 2.0, 0.0), (0.0, 0.0, 0.0)],
        [[(0, -1, 0, 1), (-1, 0, 0, 1), (0, 0, 0, 1)], [(0, 1, 0, -1), (1, 0, 0, -1), (0, 0, 0, -1)]],
    ),
    "kagome": (
        [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.5, 0.5 * S3, 0.0)],
        [(2.0, 0.0, 0.0), (1.0, S3, 0.0), (0.0, 0.0, 0.0)],
        [
            [(0, 0, 0, 1), (0, 0, 0, 2), (-1, 0, 0, 1), (0, -1, 0, 2)],
            [(0, 0, 0, -1), (1, 0, 0, -1), (0, 0, 0, 1), (1, -1, 0, 1)],
            [(0, 0, 0, -2), (0, 0, 0, -1), (0, 1, 0, -2), (-1, 1, 0, -1)],
        ],
    ),
    "triangular": (
        [(0.0, 0.0, 0.0)],
        [(S3 / 2.0, -0.5, 0.0), (S3 / 2.0, 0.5, 0.0), (0.0, 0.0, 0.0)],
        [[(1, 0, 0, 0), (-1, 0, 0, 0), (0, 1, 0, 0), (0, -1, 0, 0), (1, -1, 0, 0), (-1, 1, 0, 0)]],
    ),
    "cubic": (
        [(0.0, 0.0, 0.0)],
        [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)],
        [[(1, 0, 0, 0), (-1, 0, 0, 0), (0, 1, 0, 0), (0, -1, 0, 0), (0, 0, 1, 0), (0, 0, -1, 0)]],
    ),
    "fcc": (
        [(0.0, 0.0, 0.0)],
        [(0.0, 1.0 / S2, 1.0 / S2), (1.0 / S2, 0.0, 1.0 / S2), (1.0 / S2, 1.0 / S2, 0.0)],
        [[(1, 0, 0, 0), (-1, 0, 0, 0), (0, 1, -1, 0), (0, -1, 1, 0), (0, 1, 0, 0), (0, -1, 0, 0),
          (1, 0, -1, 0), (-1, 0, 1, 0), (0, 0, 1, 0), (0, 0, -1, 0), (1, -1, 0, 0), (-1, 1, 0, 0)]],
    ),
    "bcc": (
        [(0.0, 0.0, 0.0)],
        [(0.5, 0.5, -0.5), (-0.5, 0.5, 0.5), (0.5, -0.5, 0.5)],
        [[(-1, 0, 0, 0), (1, 0, 0, 0), (0, -1, 0, 0), (0, 1, 0, 0), (0, 0, -1, 0), (0, 0, 1, 0),
          (1, 1, 1, 0), (-1, -1, -1, 0)]],
    ),
    "hyperhoneycomb": (
        [(0.0, 0.0, 0.0), (1.0, 1.0, 0.0), (1.0, 2.0, 1.0), (0.0, -1.0, 1.0)],
        [(-1.0, 1.0, -2.0), (-1.0, 1.0, 2.0), (2.0, 4.0, 0.0)],
        [
            [(0, 0, 0, 1), (0, 0, 0, 3), (1, 0, 0, 3)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, -1, 0, 1)],
            [(0, 0, 0, -1), (0, 1, 0, -1), (0, 0, 1, 1)],
            [(0, 0, 0, -3), (-1, 0, 0, -3), (0, 0, -1, -1)],
        ],
    ),
    "hyperkagome": (
        [(0.0, 0.0, 0.0), (-1 / S2, 1 / S2, 0.0), (0.0, 1 / S2, 1 / S2), (0.0, 2 / S2, 2 / S2),
         (-1 / S2, 2 / S2, 3 / S2), (-2 / S2, 1 / S2, 3 / S2), (-2 / S2, 0.0, 2 / S2), (-3 / S2, 0.0, 3 / S2),
         (-1 / S2, 3 / S2, 2 / S2), (-2 / S2, 3 / S2, 1 / S2), (-3 / S2, 3 / S2, 0.0), (-3 / S2, 2 / S2, 1 / S2)],
        [(2.0 * S2, 0.0, 0.0), (0.0, 2.0 * S2, 0.0), (0.0, 0.0, 2.0 * S2)],
        [
            [(0, 0, 0, 1), (0, 0, 0, 2), (1, 0, -1, 7), (1, -1, 0, 10)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, 0, -1, 3), (0, 0, -1, 4)],
            [(0, 0, 0, -1), (0, 0, 0, -2), (0, 0, 0, 1), (1, 0, 0, 9)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, 0, 0, 5), (1, 0, 0, 8)],
            [(0, 0, 0, -1), (0, 0, 0, 4), (0, 0, 0, 1), (0, 0, 1, -3)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, 0, 0, 2), (0, 0, 1, -4)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, -1, 0, 2), (0, -1, 0, 3)],
            [(0, 0, 0, -2), (0, 0, 0, -1), (-1, 0, 1, -7), (0, -1, 1, 3)],
            [(0, 0, 0, -5), (0, 0, 0, -4), (0, 1, 0, -2), (0, 0, 0, 1)],
            [(0, 1, 0, -3), (0, 0, 0, -1), (0, 0, 0, 1), (0, 0, 0, 2)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (-1, 1, 0, -10), (0, 1, -1, -3)],
            [(0, 0, 0, -2), (0, 0, 0, -1), (-1, 0, 0, -9), (-1, 0, 0, -8)],
        ],
    ),
    "pyrochlore": (
        [(0.0, 0.0, 0.0), (0.0, 0.25, 0.25), (0.25, 0.0, 0.25), (0.25, 0.25, 0.0)],
        [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)],
        [
            [(0, 0, 0, 1), (0, 0, 0, 2), (0, 0, 0, 3), (-1, 0, 0, 1), (0, -1, 0, 2), (0, 0, -1, 3)],
            [(0, 0, 0, -1), (0, 0, 0, 1), (0, 0, 0, 2), (1, 0, 0, -1), (1, -1, 0, 1), (1, 0, -1, 2)],
            [(0, 0, 0, -2), (0, 0, 0, -1), (0, 0, 0, 1), (0, 1, 0, -2), (-1, 1, 0, -1), (0, 1, -1, 1)],
            [(0, 0, 0, -3), (0, 0, 0, -2), (0, 0, 0, -1), (0, 0, 1, -3), (-1, 0, 1, -2), (0, -1, 1, -1)],
        ],
    ),
    "diamond": (
        [(0.0, 0.0, 0.0), (1 / S3, 1 / S3, 1 / S3)],
        [(0.0, 2 / S3, 2 / S3), (2 / S3, 0.0, 2 / S3), (2 / S3, 2 / S3, 0.0)],
        [[(0, 0, 0, 1), (-1, 0, 0, 1), (0, -1, 0, 1), (0, 0, -1, 1)],
         [(0, 0, 0, -1), (1, 0, 0, -1), (0, 1, 0, -1), (0, 0, 1, -1)]],
    ),
}


def lattice_avail() -> list[str]:
    """Names of the available unit cells (unitcell.jl:43-61)."""
    return list(_UNITCELLS)


@dataclass
class Unitcell:
    basis: np.ndarray    # (nb, 3)
    vectors: np.ndarray  # (3, 3) rows = primitive vectors
    bonds: list          # per basis site (0-based): list of (n1, n2, n3, Δb)


def get_unitcell(name: str) -> Unitcell:
    if name not in _UNITCELLS:
        raise ValueError(f"Unitcell {name} unknown.")
    b, v, bo = _UNITCELLS[name]
    return Unitcell(np.array(b, dtype=np.float64), np.array(v, dtype=np.float64), [list(x) for x in bo])


def get_vec(it, uc: Unitcell) -> np.ndarray:
    """site.jl:114-125: integer (n1, n2, n3, basis 0-based) -> cartesian."""
    n1, n2, n3, b = it
    return n1 * uc.vectors[0] + n2 * uc.vectors[1] + n3 * uc.vectors[2] + uc.basis[b]


def _step(it, bond):
    return (it[0] + bond[0], it[1] + bond[1], it[2] + bond[2], it[3] + bond[3])


def get_sites(size: int, uc: Unitcell, euclidean: bool = False) -> list[tuple]:
    """BFS truncation of the lattice graph (site.jl:128-248)."""
    ints = [(0, 0, 0, 0)]
    ints_set = {ints[0]}
    current = [ints[0]]
    touched = {ints[0]}
    if euclidean:
        nn_distance = float(np.linalg.norm(get_vec(_step(ints[0], uc.bonds[0][0]), uc)))
        rmax = size * nn_distance
        while True:
            new_ints, out_ints = [], []
            for it in current:
                for bd in uc.bonds[it[3]]:
                    nw = _step(it, bd)
                    if nw not in touched and nw not in ints_set:
                        if np.linalg.norm(get_vec(nw, uc)) <= rmax:
                            new_ints.append(nw); ints.append(nw); ints_set.add(nw)
                        else:
                            out_ints.append(nw)
            if not out_ints:
                touched, current = set(current), new_ints
            else:
                re_ints = []
                cur_set = set(current)
                for it in out_ints:
                    for bd in uc.bonds[it[3]]:
                        nw = _step(it, bd)
                        if nw not in cur_set and nw not in ints_set:
                            if np.linalg.norm(get_vec(nw, uc)) <= rmax:
                                re_ints.append(it)
                                break
                if re_ints:
                    touched, current = set(current), new_ints + re_ints
                elif new_ints:
                    touched, current = set(current), new_ints
                else:
                    break
    else:
        metric = 0
        while metric < size:
            new_ints = []
            for it in current:
                for bd in uc.bonds[it[3]]:
                    nw = _step(it, bd)
                    if nw not in touched and nw not in ints_set:
                        new_ints.append(nw); ints.append(nw); ints_set.add(nw)
            touched, current = set(current), new_ints
            metric += 1
    return ints


def get_metric(it1, it2, uc: Unitcell) -> int:
    """Bond distance between two sites (site.jl:159-205)."""
    if tuple(it1) == tuple(it2):
        return 0
    current = [tuple(it2)]
    touched = {tuple(it2)}
    metric = 0
    target = tuple(it1)
    while True:
        metric += 1
        new = []
        for it in current:
            for bd in uc.bonds[it[3]]:
                nw = _step(it, bd)
                if nw not in touched:
                    new.append(nw)
        if target in new:
            return metric
        touched, current = set(current), new


def get_test_sites(uc: Unitcell) -> tuple[list[tuple], int]:
    """Minimal symmetry test set (site.jl:291-330), bond metric version."""
    test, seen = [], set()
    for i in range(len(uc.basis)):
        b = (0, 0, 0, i)
        for it in [b] + [_step(b, bd) for bd in uc.bonds[i]]:
            key = _key(get_vec(it, uc))
            if key not in seen:
                seen.add(key); test.append(it)
    metric = max(get_metric(test[0], s, uc) for s in test)
    return test, metric


def _key(vec) -> tuple:
    return tuple(int(round(float(x) * 1e6)) for x in vec)


@dataclass
class Lattice:
    name: str
    size: int
    uc: Unitcell
    test_sites: list
    ints: list               # site integer coordinates
    vecs: np.ndarray         # (N, 3)
    bonds: np.ndarray        # (N, N, 3, 3) exchange matrices
    index: dict = field(default_factory=dict)

    def get_site(self, vec) -> int:
        """Index of the site at vec, or -1 (build.jl:257-272)."""
        return self.index.get(_key(vec), -1)


def get_lattice(name: str, size: int, euclidean: bool = False) -> Lattice:
    """build.jl:79-129."""
    uc = get_unitcell(name)
    test_sites, metric = get_test_sites(uc)
    if euclidean:
        raise NotImplementedError("Euclidean truncation is only provided through get_sites(euclidean=True)")
    assert metric <= size, "Lattice is too small to perform symmetry reduction."
    ints = get_sites(size, uc, euclidean)
    vecs = np.array([get_vec(it, uc) for it in ints])
    n = len(ints)
    l = Lattice(name, size, uc, test_sites, ints, vecs, np.zeros((n, n, 3, 3)))
    l.index = {_key(v): i for i, v in enumerate(vecs)}
    return l


def get_nbs(n: int, i: int, vecs: np.ndarray) -> np.ndarray:
    """n-th nearest neighbours of site i by Euclidean distance (site.jl:230-270)."""
    d = np.linalg.norm(vecs - vecs[i], axis=1)
    uniq: list[float] = []
    for x in d:
        if not any(abs(u - x) <= 1e-8 for u in uniq):
            uniq.append(float(x))
    assert len(uniq) >= n + 1, "Could not find neighbors in list."
    dn = sorted(uniq)[n]
    return np.nonzero(np.abs(dn - d) <= 1e-8)[0]


def _add_heisenberg(l: Lattice, i: int, j: int, c: float) -> None:
    for k in range(3):
        l.bonds[i, j, k, k] += c


def _add_shell(l: Lattice, i: int, n: int, Jn, nbs) -> None:
    """n-th neighbour shell of site i: uniform if Jn has one value, else one value per bond
    distance class in ascending order (model_heisenberg.jl:31-52, model_breathing.jl:50-71)."""
    if len(Jn) == 1:
        for j in nbs:
            _add_heisenberg(l, i, j, Jn[0])
    else:
        dist = [get_metric(l.ints[j], l.ints[i], l.uc) for j in nbs]
        kinds = sorted(set(dist))
        assert len(Jn) == len(kinds), (
            f"{l.name} has {len(kinds)} inequivalent {n}-th nearest neighbors, but {len(Jn)} couplings were supplied.")
        for nk, kind in enumerate(kinds):
            for j, dj in zip(nbs, dist):
                if dj == kind:
                    _add_heisenberg(l, i, j, Jn[nk])


def init_model_heisenberg(J: list[list[float]], l: Lattice) -> None:
    """Heisenberg couplings J[n] on n-th neighbour shells (model_heisenberg.jl:13-56)."""
    for i in range(len(l.ints)):
        for n, Jn in enumerate(J, start=1):
            _add_shell(l, i, n, Jn, get_nbs(n, i, l.vecs))


def init_model_breathing(J: list[list[float]], l: Lattice) -> None:
    """Breathing kagome / pyrochlore Heisenberg model (model_breathing.jl:14-76): J[0] =
    [J_a, J_b], J_a on nearest-neighbour bonds inside the unit cell (up triangles/tetrahedra),
    J_b on those that cross it; further shells as in the Heisenberg model."""
    assert l.name in ("kagome", "pyrochlore"), "Breathing model requires Pyrochlore or Kagome lattice."
    assert len(J[0]) == 2, "Breathing model needs two nearest neighbor couplings."
    for i in range(len(l.ints)):
        for n, Jn in enumerate(J, start=1):
            nbs = get_nbs(n, i, l.vecs)
            if n == 1:
                for j in nbs:
                    inside = tuple(l.ints[j][:3]) == tuple(l.ints[i][:3])
                    _add_heisenberg(l, i, j, Jn[0] if inside else Jn[1])
            else:
                _add_shell(l, i, n, Jn, nbs)


def init_model_pyrochlore_breathing_c3(J: list[list[float]], l: Lattice) -> None:
    """Breathing pyrochlore with broken C3 (model_breathing.jl:91-182): J[0] = [J_a, J_b, δ_a,
    δ_b]; a nearest-neighbour bond inside (across) the unit cell carries J_a ± δ_a (J_b ± δ_b),
    + on bonds joining basis sites 1-2 or 3-4, − on the others."""
    assert l.name == "pyrochlore", "Model requires Pyrochlore lattice."
    assert len(J[0]) == 4, "Model requires two nearest neighbor couplings and two C3 breaking perturbations."
    for i in range(len(l.ints)):
        for n, Jn in enumerate(J, start=1):
            nbs = get_nbs(n, i, l.vecs)
            if n == 1:
                for j in nbs:
                    inside = tuple(l.ints[j][:3]) == tuple(l.ints[i][:3])
                    pair = {l.ints[j][3], l.ints[i][3]}
                    plus = pair == {0, 1} or pair == {2, 3}  # basis sites 1-2 or 3-4 (1-based)
                    Jc, dc = (Jn[0], Jn[2]) if inside else (Jn[1], Jn[3])
                    _add_heisenberg(l, i, j, Jc + dc if plus else Jc - dc)
            else:
                _add_shell(l, i, n, Jn, nbs)


def grow_test_sites(l: Lattice, metric: int) -> None:
    """grow_test_sites! (build.jl:87-135), bond metric: add the sites within `metric` bonds of
    the origin to the symmetry test set if it is smaller."""
    current = max(get_metric(l.test_sites[0], s, l.uc) for s in l.test_sites)
    if current < metric:
        keys = {_key(get_vec(s, l.uc)) for s in l.test_sites}
        for s in get_sites(metric, l.uc, False):
            k = _key(get_vec(s, l.uc))
            if k not in keys:
                keys.add(k)
                l.test_sites.append(s)


def init_model_triangular_dm_c3(J: list[list[float]], l: Lattice) -> None:
    """Triangular lattice with out-of-plane DM interaction (model_triangular_dm_c3.jl:14-73):
    J[n] = [Jxx, Jzz, Jxy] on the n-th neighbour shell (n <= 3); the DM coupling Jxy changes
    sign between bonds that are not related by a C3 rotation of the shell's first bond."""
    assert l.name == "triangular", "Model requires triangular lattice."
    assert len(J) <= 3, "Model initialization only works up to third-nearest neighbors."
    for n, Jn in enumerate(J, start=1):
        assert len(Jn) == 3, f"Jxx, Jzz and Jxy all need to be specified for J[{n}]."
    max_nbs = get_nbs(len(J), 0, l.vecs)
    grow_test_sites(l, max(get_metric(l.ints[0], l.ints[j], l.uc) for j in max_nbs))
    ref_vecs: list[np.ndarray] = []
    for i in range(len(l.ints)):
        for n, Jn in enumerate(J, start=1):
            nbs = get_nbs(n, i, l.vecs)
            Jxx, Jzz, Jxy = Jn
            if len(ref_vecs) < n:
                ref_vecs.append(l.vecs[nbs[0]] - l.vecs[i])
            ref = ref_vecs[n - 1]
            for j in nbs:
                vec = l.vecs[j] - l.vecs[i]
                c = round(float(np.dot(ref, vec) / (np.linalg.norm(ref) * np.linalg.norm(vec))), 8)
                sg = 1.0 if math.fmod(math.acos(c), 2.0 * math.pi / 3.0) < 1e-8 else -1.0
                l.bonds[i, j, 0, 1] += sg * Jxy
                l.bonds[i, j, 1, 0] -= sg * Jxy
                l.bonds[i, j, 0, 0] += Jxx
                l.bonds[i, j, 1, 1] += Jxx
                l.bonds[i, j, 2, 2] += Jzz


# models of init_model! (build.jl:183-202): SU(2) models and the U(1)-DM model
MODELS = {
    "heisenberg": init_model_heisenberg,
    "breathing": init_model_breathing,
    "pyrochlore-breathing-c3": init_model_pyrochlore_breathing_c3,
    "triangular-dm-c3": init_model_triangular_dm_c3,
}
# the action each model needs (model_avail, build.jl:147-170)
MODEL_SYMMETRY = {"heisenberg": "su2", "breathing": "su2", "pyrochlore-breathing-c3": "su2",
                  "triangular-dm-c3": "u1-dm"}


def _are_equal(b1, b2) -> bool:
    return float(np.linalg.norm(b1 - b2)) <= 1e-8


def _get_rotation(vec, ref):
    """Rodrigues rotation of vec onto ref (reduced.jl:46-76)."""
    a = vec / np.linalg.norm(vec)
    b = ref / np.linalg.norm(ref)
    n = np.cross(a, b)
    s = np.linalg.norm(n)
    c = float(np.dot(a, b))
    if s < 1e-8:
        return -np.eye(3) if c < 0.0 else np.eye(3)
    n = n / s
    K = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    return np.eye(3) + s * K + (1.0 - c) * (K @ K)


def _get_rotation_for_axis(vec, ref, axis):
    """Rotation of vec onto ref about axis, or zeros (reduced.jl:79-124)."""
    a = vec / np.linalg.norm(vec)
    b = ref / np.linalg.norm(ref)
    ax = axis / np.linalg.norm(axis)
    ovec = vec - np.dot(vec, ax) * ax
    ovec = ovec / np.linalg.norm(ovec)
    oref = ref - np.dot(ref, ax) * ax
    oref = oref / np.linalg.norm(oref)
    s = np.linalg.norm(np.cross(ovec, oref))
    c = float(np.dot(ovec, oref))
    x, y, z = ax
    mat = np.array([
        [x * x + (1.0 - x * x) * c, x * y * (1.0 - c) - z * s, x * z * (1.0 - c) + y * s],
        [x * y * (1.0 - c) + z * s, y * y + (1.0 - y * y) * c, y * z * (1.0 - c) - x * s],
        [x * z * (1.0 - c) - y * s, y * z * (1.0 - c) + x * s, z * z + (1.0 - z * z) * c],
    ])
    if np.linalg.norm(mat @ a - b) > 1e-8:
        if np.linalg.norm(mat.T @ a - b) < 1e-8:
            return mat.T
        return np.zeros((3, 3))
    return mat


def _rotate(v1, v2, r1, r2):
    """reduced.jl:127-141."""
    mat = _get_rotation(v1, r1)
    return _get_rotation_for_axis(mat @ v2, r2, r1) @ mat


def _is_in(mat, lst) -> bool:
    return any(np.max(np.abs(m - mat)) <= 1e-8 for m in lst)


def get_trafos_orig(l: Lattice) -> list[np.ndarray]:
    """Point-group operations fixing the origin (reduced.jl:150-288)."""
    trafos: list[np.ndarray] = []
    ref = (0, 0, 0, 0)
    con = l.uc.bonds[0]
    vec_of = lambda it: get_vec(it, l.uc)

    def valid(mat):
        for t in l.test_sites:
            v = vec_of(t)
            k = l.get_site(mat @ v)
            if k < 0:
                return False
            if not _are_equal(l.bonds[0, l.get_site(v)], l.bonds[0, k]):
                return False
        return True

    for i1 in range(len(con)):
        for i2 in range(len(con)):
            if i1 == i2:
                continue
            vi1, vi2 = vec_of(_step(ref, con[i1])), vec_of(_step(ref, con[i2]))
            if np.linalg.norm(np.cross(vi1, vi2)) <= 1e-8:
                continue
            for j1 in range(len(con)):
                for j2 in range(len(con)):
                    if j1 == j2:
                        continue
                    vj1, vj2 = vec_of(_step(ref, con[j1])), vec_of(_step(ref, con[j2]))
                    if np.linalg.norm(np.cross(vj1, vj2)) <= 1e-8:
                        continue
                    mat = _rotate(vi1, vi2, vj1, vj2)
                    if np.max(np.abs(mat)) > 1e-8:
                        if not _is_in(mat, trafos) and valid(mat):
                            trafos.append(mat)
                        if not _is_in(-mat, trafos) and valid(-mat):
                            trafos.append(-mat)
    return trafos


def get_reduced(l: Lattice) -> np.ndarray:
    """Orbit representatives under the point group (reduced.jl:291-329), 0-based."""
    n = len(l.ints)
    reduced = np.arange(n)
    trafos = get_trafos_orig(l)
    for i in range(1, n):
        if reduced[i] < i:
            continue
        for T in trafos:
            mv = T @ l.vecs[i]
            if np.linalg.norm(mv - l.vecs[i]) > 1e-8:
                k = l.get_site(mv)
                assert k >= 0, "Validity on test set could not be generalized."
                if _are_equal(l.bonds[0, i], l.bonds[0, k]):
                    reduced[k] = i
    return reduced


def get_trafos_uc(l: Lattice) -> list:
    """Maps of each basis site to the origin: (matrix, inversion flag) (reduced.jl:339-502).

    The reference's stacked `break` statements leave only the innermost loop, so
    the last valid transformation in loop order is kept; restated as such."""
    nb = len(l.uc.basis)
    trafos = [None] * (nb - 1)
    ref = (0, 0, 0, 0)
    con = l.uc.bonds[0]
    vec_of = lambda it: get_vec(it, l.uc)
    for b in range(1, nb):
        basis = (0, 0, 0, b)
        bvec = vec_of(basis)
        cb = l.uc.bonds[b]
        ib = l.get_site(bvec)
        in_range = [t for t in l.test_sites if get_metric(t, basis, l.uc) <= l.size]
        for b1 in range(len(cb)):
            for b2 in range(len(cb)):
                if b1 == b2:
                    continue
                vb1, vb2 = vec_of(_step(basis, cb[b1])), vec_of(_step(basis, cb[b2]))
                if np.linalg.norm(np.cross(vb1 - bvec, vb2 - bvec)) <= 1e-8:
                    continue
                bond_b1 = l.bonds[ib, l.get_site(vb1)]
                bond_b2 = l.bonds[ib, l.get_site(vb2)]
                for r1 in range(len(con)):
                    for r2 in range(len(con)):
                        if r1 == r2:
                            continue
                        vr1, vr2 = vec_of(_step(ref, con[r1])), vec_of(_step(ref, con[r2]))
                        if np.linalg.norm(np.cross(vr1, vr2)) <= 1e-8:
                            continue
                        if not _are_equal(bond_b1, l.bonds[0, l.get_site(vr1)]) or not _are_equal(
                                bond_b2, l.bonds[0, l.get_site(vr2)]):
                            continue
                        found = False
                        for inv in (False, True):
                            sg = -1.0 if inv else 1.0
                            mat = _rotate(sg * (vb1 - bvec), sg * (vb2 - bvec), vr1, vr2)
                            if np.max(np.abs(mat)) <= 1e-8:
                                continue
                            ok = True
                            for t in in_range:
                                tv = vec_of(t)
                                k = l.get_site(mat @ (sg * (tv - bvec)))
                                if k < 0 or not _are_equal(l.bonds[ib, l.get_site(tv)], l.bonds[0, k]):
                                    ok = False
                                    break
                            if ok:
                                trafos[b - 1] = (mat, inv)
                                found = True
                                break
                        if found:
                            break
    return trafos


def get_mappings(l: Lattice, reduced: np.ndarray) -> np.ndarray:
    """mappings[i, j] = representative (0-based) of the pair (i, j), -1 if out of range (reduced.jl:528-599)."""
    n = len(l.ints)
    mat = np.full((n, n), -1, dtype=np.int64)
    np.fill_diagonal(mat, reduced[0])
    trafos = get_trafos_uc(l)
    dmax = float(np.max(np.linalg.norm(l.vecs, axis=1)))
    for i in range(n):
        it = l.ints[i]
        b = it[3]
        shift = get_vec((it[0], it[1], it[2], 0), l.uc)
        if b != 0:
            T, inv = trafos[b - 1]
            if inv:
                mapped = (T @ (shift[None, :] - l.vecs + l.uc.basis[b]).T).T
            else:
                mapped = (T @ (l.vecs - shift[None, :] - l.uc.basis[b]).T).T
        else:
            mapped = l.vecs - shift[None, :]
        for j in range(n):
            if i == j:
                continue
            if dmax - np.linalg.norm(mapped[j]) > -1e-8:
                k = l.get_site(mapped[j])
                if k >= 0:
                    mat[i, j] = reduced[k]
    return mat


@dataclass
class ReducedLattice:
    name: str
    size: int
    model: str
    J: list
    irreducible: np.ndarray   # indices of irreducible sites in the full lattice
    ints: list                # integer coordinates of irreducible sites
    vecs: np.ndarray
    overlap: list             # per irreducible site: (rows, 3) int64 of (j1, j2, mult), 0-based
    mult: np.ndarray          # int64 orbit sizes (Reduced_lattice.mult)
    exchange: np.ndarray      # int64, 0-based
    project: np.ndarray       # (N, N) int64, -1 = out of range

    @property
    def n_sites(self) -> int:
        return len(self.irreducible)

    def overlap_csr(self) -> tuple[np.ndarray, np.ndarray]:
        """Overlap rows flattened: offsets (n_sites+1) and rows (R, 3), 0-based."""
        off = np.zeros(self.n_sites + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(o) for o in self.overlap])
        rows = np.concatenate(self.overlap).astype(np.int64) if self.overlap else np.zeros((0, 3), np.int64)
        return off, rows


def get_reduced_lattice(model: str, J: list[list[float]], l: Lattice) -> ReducedLattice:
    """reduced.jl:716-759."""
    if model not in MODELS:
        raise NotImplementedError(f"Model {model} unknown or not provided (models: {', '.join(MODELS)}).")
    MODELS[model](J, l)
    reduced = get_reduced(l)
    irreducible = np.array(list(dict.fromkeys(reduced.tolist())), dtype=np.int64)
    pos = {int(v): k for k, v in enumerate(irreducible)}
    mappings = get_mappings(l, reduced)
    overlap = []
    for irr in irreducible:
        temp = [(int(reduced[j]), int(mappings[j, irr])) for j in range(len(l.ints)) if mappings[irr, j] >= 0]
        pairs = list(dict.fromkeys(temp))
        counts = {p: 0 for p in pairs}
        for p in temp:
            counts[p] += 1
        overlap.append(np.array([[pos[p[0]], pos[p[1]], counts[p]] for p in pairs], dtype=np.int64).reshape(-1, 3))
    mult = np.array([int(np.sum(reduced == irr)) for irr in irreducible], dtype=np.int64)
    exchange = np.array([pos[int(mappings[irr, 0])] for irr in irreducible], dtype=np.int64)
    project = np.full(mappings.shape, -1, dtype=np.int64)
    ok = mappings >= 0
    project[ok] = np.vectorize(lambda v: pos[int(v)])(mappings[ok]) if ok.any() else project[ok]
    return ReducedLattice(l.name, l.size, model, J, irreducible, [l.ints[i] for i in irreducible],
                          l.vecs[irreducible], overlap, mult, exchange, project)


def bare_couplings(l: Lattice, r: ReducedLattice, S: float = 0.5, A: float = 0.0,
                   symmetry: str = "su2") -> np.ndarray:
    """Bare vertex per irreducible site from the bond to the origin.  SU(2): spin J_xx/4/(2S)
    (+ level repulsion A/4/(2S) on site 0) (action_su2.jl:35-66), returns (2, n_sites) [spin,
    density].  U(1)-DM: xx J_xx/4, zz J_zz/4, DM J_xy/4 (+ A/4 on xx and zz of site 0)
    (action_u1_dm.jl:32-75), returns (6, n_sites) [xx, zz, DM, dd, zd, dz]."""
    if symmetry == "u1-dm":
        bare = np.zeros((6, r.n_sites))
        for i, irr in enumerate(r.irreducible):
            bare[0, i] = l.bonds[0, irr, 0, 0] / 4.0
            bare[1, i] = l.bonds[0, irr, 2, 2] / 4.0
            bare[2, i] = l.bonds[0, irr, 0, 1] / 4.0
        bare[0, 0] += A / 4.0
        bare[1, 0] += A / 4.0
        return bare
    bare = np.zeros((2, r.n_sites))
    for i, irr in enumerate(r.irreducible):
        bare[0, i] = l.bonds[0, irr, 0, 0] / 4.0 / (2.0 * S)
    bare[0, 0] += A / 4.0 / (2.0 * S)
    return bare


def normalize_couplings(J: list[list[float]], A: float = 0.0) -> tuple[list[list[float]], float]:
    """normalize([J, [[A]]]) of Launcher.jl:647-650: divide by sqrt(ΣJ² + A²)."""
    nrm = math.sqrt(sum(x * x for Jn in J for x in Jn) + A * A)
    return [[x / nrm for x in Jn] for Jn in J], A / nrm


def build(name: str, size: int, J: list[list[float]], model: str = "heisenberg") -> tuple[Lattice, ReducedLattice]:
    l = get_lattice(name, size)
    r = get_reduced_lattice(model, J, l)
    return l, r
