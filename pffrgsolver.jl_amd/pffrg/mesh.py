"""Frequency meshes (src/Frequency/mesh.jl) and the launch! mesh recipe (Launcher.jl:667-673)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# launch! defaults, Launcher.jl:572-607
P_SIGMA = (0.4, 1.5, 1.0)
P_OMEGA_S = (0.4, 0.02, 0.05, 0.05, 40.0, 1.0)
P_NU_S = (0.4, 0.02, 0.05, 0.03, 30.0, 1.0)
P_OMEGA_T = (0.4, 0.02, 0.05, 0.05, 40.0, 1.0)
P_NU_T = (0.4, 0.02, 0.05, 0.03, 30.0, 1.0)
P_CHI = (0.4, 0.02, 0.05, 0.01, 20.0, 1.0)
LINS = (5.0, 2.0, 6.0, 4.0, 4.0)
BOUNDS = (0.5, 50.0, 200.0, 150.0, 100.0)


def get_mesh(linear: float, upper: float, num: int, p: float) -> np.ndarray:
    """(num+1) frequencies: 0, ceil(p*num) linear points up to `linear`, the rest
    logarithmic up to `upper` (mesh.jl:86-121)."""
    assert linear > 0.0, "Linear bound must be larger than zero."
    assert upper > 0.0, "Upper bound must be larger than zero."
    assert linear < upper, "Linear bound must be smaller than upper bound."
    num_lin = int(math.ceil(p * num))
    num_log = num - num_lin
    assert num_log > 0, "Number of frequencies is too small."
    h = linear / num_lin
    xi = (upper / linear) ** (1.0 / num_log)
    mesh = np.zeros(num + 1)
    for i in range(1, num_lin + 1):
        mesh[i] = i * h
    for i in range(1, num_log + 1):
        mesh[num_lin + i] = xi ** i * linear
    return mesh


@dataclass
class Mesh:
    """Mesh struct of mesh.jl:16-27; arrays have num+1 entries."""
    sigma: np.ndarray
    Os: np.ndarray
    ns: np.ndarray
    Ot: np.ndarray
    nt: np.ndarray
    chi: np.ndarray

    @property
    def num_sigma(self) -> int:
        return len(self.sigma)

    @property
    def num_Omega(self) -> int:
        return len(self.Os)

    @property
    def num_nu(self) -> int:
        return len(self.ns)

    @property
    def num_chi(self) -> int:
        return len(self.chi)


def launch_mesh(initial: float, num_sigma: int = 25, num_Omega: int = 15, num_nu: int = 10, num_chi: int = 10,
                p_sigma=P_SIGMA, p_Os=P_OMEGA_S, p_ns=P_NU_S, p_Ot=P_OMEGA_T, p_nt=P_NU_T, p_chi=P_CHI,
                lins=LINS, bounds=BOUNDS) -> Mesh:
    """Initial meshes built by launch! (Launcher.jl:667-673)."""
    up = max(initial, bounds[0])
    sigma = get_mesh(lins[1] * initial, bounds[1] * up, num_sigma, p_sigma[0]) * p_sigma[2]
    Os = get_mesh(lins[2] * initial, bounds[2] * up, num_Omega, p_Os[0]) * p_Os[5]
    ns = get_mesh(lins[3] * initial, bounds[3] * up, num_nu, p_ns[0]) * p_ns[5]
    Ot = get_mesh(lins[2] * initial, bounds[2] * up, num_Omega, p_Ot[0]) * p_Ot[5]
    nt = get_mesh(lins[3] * initial, bounds[3] * up, num_nu, p_nt[0]) * p_nt[5]
    chi = get_mesh(lins[4] * initial, bounds[4] * up, num_chi, p_chi[0]) * p_chi[5]
    return Mesh(sigma, Os, ns, Ot, nt, chi)


def get_eval(L: float, min_eval: int = 10, max_eval: int = 100) -> int:
    """Sub-interval count of the quadrature, launcher_1l.jl:78."""
    return min(max(int(math.ceil(min_eval / math.sqrt(L))), min_eval), max_eval)
