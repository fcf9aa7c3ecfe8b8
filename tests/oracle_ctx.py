"""An oracle-backed stand-in for pffrg.native.Context: the same slot API (upload, algebra,
compute_*) on host actions, evaluated by the FP64 CPU oracle.  The host flow drivers of
pffrg.flow run unchanged on it, which makes the GPU flow testable against the oracle's."""
from __future__ import annotations

import numpy as np

import oracle as O
from pffrg.action import Action


class OracleContext:
    def __init__(self, prob, n_slots: int = 9):
        self.r = prob.r
        self.nc = getattr(prob, "nc", 2)
        self.orc = O.Oracle(prob.r, prob.mesh, prob.S, self.nc)
        self.m = prob.mesh
        self.ns = prob.r.n_sites
        self.S = prob.S
        self.slots = [Action.empty(self.ns, self.m, self.S, nc=self.nc) for _ in range(n_slots)]

    def set_mesh(self, m):
        self.orc = O.Oracle(self.r, m, self.S, self.nc)
        self.m = m

    def get_sigma(self, s):
        return self.slots[s].Sigma.copy()

    def scan_cuts(self, a):
        return self.orc.scan_cuts(self.slots[a])

    def resample(self, src, dst, m_new):
        out = Action.empty(self.ns, m_new, self.S, nc=self.nc)
        self.orc.resample(self.slots[src], m_new, out)
        self.slots[dst] = out
        self.set_mesh(m_new)

    def upload(self, s, a):
        self.slots[s] = a.copy()

    def download(self, s):
        return self.slots[s].copy()

    def replace_with(self, dst, src, vertex_only=False):
        self.slots[dst].gamma[:] = self.slots[src].gamma
        if not vertex_only:
            self.slots[dst].Sigma[:] = self.slots[src].Sigma

    def mult_with_add_to(self, src, fac, dst, vertex_only=False):
        self.slots[dst].gamma[:] += fac * self.slots[src].gamma
        if not vertex_only:
            self.slots[dst].Sigma[:] += fac * self.slots[src].Sigma

    def abs_max(self, s) -> float:
        a = self.slots[s]
        return max(float(np.max(np.abs(a.view(c, ch, "q3")))) for c in range(self.nc) for ch in range(3))

    def synchronize(self):
        pass

    def compute_dsigma(self, L, a, da, tol=(1e-8, 1e-4)):
        self.orc.compute_dsigma(L, self.slots[a], self.slots[da], tol)

    def compute_dgamma_1l(self, L, a, da, eval_, tol=(1e-8, 1e-4)):
        self.orc.compute_dgamma_1l(L, self.slots[a], self.slots[da], eval_, tol)

    def compute_dgamma_2l(self, L, a, da, dal, eval_, tol=(1e-8, 1e-4)):
        self.orc.compute_dgamma_2l(L, self.slots[a], self.slots[da], self.slots[dal], eval_, tol)

    def compute_dgamma_ml(self, L, loops, a, da, dal, dac, datemp, daS, eval_, tol=(1e-8, 1e-4)):
        sl = self.slots
        self.orc.compute_dgamma_ml(L, loops, sl[a], sl[da], sl[dal], sl[dac], sl[datemp], sl[daS], eval_, tol)

    def compute_dsigma_corr(self, L, a, da, daS, tol=(1e-8, 1e-4)):
        self.orc.compute_dsigma_corr(L, self.slots[a], self.slots[da], self.slots[daS], tol)

    def compute_chi(self, L, a, tol=(1e-8, 1e-4)):
        return self.orc.compute_chi(L, self.slots[a], tol)

    def set_sigma(self, s, sig):
        self.slots[s].Sigma[:] = sig

    def mult_with(self, s, fac, vertex_only=False):
        self.slots[s].gamma[:] *= fac
        if not vertex_only:
            self.slots[s].Sigma[:] *= fac

    def compute_sigma(self, L, a1, a2, eval_, gtol=(1e-8, 1e-4), stol=(1e-8, 1e-4)):
        self.orc.compute_sigma(L, self.slots[a1], self.slots[a2], eval_, gtol, stol)

    def compute_gamma(self, L, a1, a2, eval_, tol=(1e-8, 1e-4)):
        self.orc.compute_gamma(L, self.slots[a1], self.slots[a2], eval_, tol)
