"""Host-side action: self energy, bare couplings and the packed vertex, for the SU(2) action
(2 components: spin, density; action_lib/action_su2.jl:9-13) and the U(1)-symmetric DM action
(6 components: xx, zz, DM, dd, zd, dz; action_lib/action_u1_dm.jl:1-11).

Packed vertex layout (identical to the reference's column-major Julia arrays and
hence to the HDF5 checkpoint datasets, Action/channel.jl:10-35, vertex.jl:10-15):

    gamma = [comp][channel (s, t, u)][q1 | q2_1 | q2_2 | q3]
    q1[site, w], q2_*[site, w, v], q3[site, w, v, vp]   (site fastest)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh

N_COMP = 2      # SU(2)
N_COMP_U1 = 6   # U(1)-DM
SYMMETRIES = {"su2": N_COMP, "u1-dm": N_COMP_U1}


def channel_sizes(ns: int, nO: int, nn: int) -> tuple[int, int, int, int]:
    return ns * nO, ns * nO * nn, ns * nO * nn, ns * nO * nn * nn


def channel_len(ns: int, nO: int, nn: int) -> int:
    return sum(channel_sizes(ns, nO, nn))


def gamma_len(ns: int, nO: int, nn: int, nc: int = N_COMP) -> int:
    return nc * 3 * channel_len(ns, nO, nn)


@dataclass
class Action:
    S: float
    Sigma: np.ndarray   # (n_sigma,)
    bare: np.ndarray    # (nc, ns)
    gamma: np.ndarray   # flat packed vertex
    ns: int
    nO: int
    nn: int
    nc: int = N_COMP    # vertex components: 2 (SU(2)) or 6 (U(1)-DM)

    @classmethod
    def empty(cls, ns: int, m: Mesh, S: float = 0.5, nc: int = N_COMP) -> "Action":
        """get_action_empty (Action.jl:564-577): zero Σ, zero bare, zero channels."""
        return cls(S, np.zeros(m.num_sigma), np.zeros((nc, ns)),
                   np.zeros(gamma_len(ns, m.num_Omega, m.num_nu, nc)), ns, m.num_Omega, m.num_nu, nc)

    def copy(self) -> "Action":
        return Action(self.S, self.Sigma.copy(), self.bare.copy(), self.gamma.copy(), self.ns, self.nO, self.nn,
                      self.nc)

    def view(self, comp: int, ch: int, q: str) -> np.ndarray:
        """Column-major (Fortran-order) view of one kernel array, e.g. view(0, 0, 'q3')[site, w, v, vp]."""
        ns, nO, nn = self.ns, self.nO, self.nn
        sizes = channel_sizes(ns, nO, nn)
        shapes = {"q1": (ns, nO), "q2_1": (ns, nO, nn), "q2_2": (ns, nO, nn), "q3": (ns, nO, nn, nn)}
        offs = {"q1": 0, "q2_1": sizes[0], "q2_2": sizes[0] + sizes[1], "q3": sizes[0] + sizes[1] + sizes[2]}
        base = (comp * 3 + ch) * channel_len(ns, nO, nn) + offs[q]
        n = int(np.prod(shapes[q]))
        return self.gamma[base:base + n].reshape(shapes[q], order="F")

    def abs_max(self) -> float:
        """get_abs_max over q3 of every channel and component (Action.jl:346-359, channel.jl:406-413)."""
        return max(float(np.max(np.abs(self.view(c, ch, "q3")))) for c in range(self.nc) for ch in range(3))


# ---------------------------------------------------------------- seeded synthetic inputs
MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of the splitmix64 generator started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normals(seed: int, n: int) -> np.ndarray:
    """n N(0,1) draws: Box-Muller on consecutive splitmix64 pairs, u = (x >> 11) * 2^-53."""
    z = splitmix64(seed, 2 * n)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def synthetic_action(ns: int, m: Mesh, bare: np.ndarray, seed: int = 20261203, S: float = 0.5) -> Action:
    """Seeded vertex of the single-RHS benchmark: every channel entry is
    g * 1/(1+|Ω|+|ν|+|ν'|), g ~ N(0,1), filled comp -> channel -> q1,q2_1,q2_2,q3 in
    column-major order; asymptotic slots use ν, ν' = ν_max.  Σ_k = 0.5 σ_k/(1+σ_k²).  The number
    of components is bare's first dimension."""
    a = Action.empty(ns, m, S, nc=np.asarray(bare).shape[0])
    a.bare[:] = bare
    nO, nn = m.num_Omega, m.num_nu
    g = normals(seed, a.gamma.size)
    pos = 0
    for comp in range(a.nc):
        for ch in range(3):
            Om = m.Ot if ch == 1 else m.Os
            nu = m.nt if ch == 1 else m.ns
            vmax = nu[-1]
            W = np.abs(Om)[None, :]
            V = np.abs(nu)
            dec = {
                "q1": np.broadcast_to(1.0 / (1.0 + W + vmax + vmax), (ns, nO)),
                "q2_1": np.broadcast_to(1.0 / (1.0 + W[:, :, None] + V[None, None, :] + vmax), (ns, nO, nn)),
                "q2_2": np.broadcast_to(1.0 / (1.0 + W[:, :, None] + vmax + V[None, None, :]), (ns, nO, nn)),
                "q3": np.broadcast_to(1.0 / (1.0 + W[:, :, None, None] + V[None, None, :, None] + V[None, None, None, :]),
                                      (ns, nO, nn, nn)),
            }
            for q in ("q1", "q2_1", "q2_2", "q3"):
                d = dec[q]
                n = d.size
                a.view(comp, ch, q)[...] = (g[pos:pos + n].reshape(d.shape, order="F") * d)
                pos += n
    a.Sigma[:] = 0.5 * m.sigma / (1.0 + m.sigma ** 2)
    return a
