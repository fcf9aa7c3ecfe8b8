"""ctypes wrapper of the FP64 CPU oracle (oracle/pffrg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  It is never used by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# PFFRG_ORACLE_LIB points the checker at another build of this file (the planted-mistake tests)
LIB = Path(os.environ.get("PFFRG_ORACLE_LIB", HERE / "build" / "liboracle.so"))

P_KAT, P_LEFT, P_CENTRAL, P_BSE, P_REDUCED = range(5)


def test_hcubature(kind, p, q, lo, hi, atol, rtol):
    """2-D adaptive cubature of x^p y^q (kind 0) or exp(-(x^2+y^2)) (kind 1); returns (I, regions)."""
    nr = C.c_int32(0)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    I = lib().orc_test_hcubature(kind, p, q, _dp(lo), _dp(hi), atol, rtol, C.byref(nr))
    return I, nr.value


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class System(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("n_sigma", C.c_int32), ("n_Omega", C.c_int32), ("n_nu", C.c_int32),
                ("S", C.c_double),
                ("sigma", C.POINTER(C.c_double)), ("Os", C.POINTER(C.c_double)), ("ns", C.POINTER(C.c_double)),
                ("Ot", C.POINTER(C.c_double)), ("nt", C.POINTER(C.c_double)),
                ("exchange", C.POINTER(C.c_int64)), ("mult", C.POINTER(C.c_double)),
                ("ov_off", C.POINTER(C.c_int64)), ("ov_rows", C.POINTER(C.c_int64)), ("n_comp", C.c_int32)]


class ActionS(C.Structure):
    _fields_ = [("Sigma", C.POINTER(C.c_double)), ("bare", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double))]


_D = C.POINTER(C.c_double)
_S = C.POINTER(System)
_A = C.POINTER(ActionS)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_param.argtypes = [C.c_double, _D, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D, _D]
        L.orc_get_buffer.argtypes = [_S, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32), _D]
        L.orc_get_gamma_avx.argtypes = [_S, _A, _D, C.c_int32, _D]
        L.orc_get_vertex.argtypes = [_S, _A, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_get_vertex.restype = C.c_double
        L.orc_get_sigma.argtypes = [_S, _D, C.c_double]
        L.orc_get_sigma.restype = C.c_double
        L.orc_G_bare.argtypes = [C.c_double, C.c_double]
        L.orc_G_bare.restype = C.c_double
        for n in ("orc_G", "orc_S"):
            getattr(L, n).argtypes = [_S, C.c_double, C.c_double, _D]
            getattr(L, n).restype = C.c_double
        L.orc_prop_kat.argtypes = [_S, C.c_double, C.c_double, C.c_double, _D, _D]
        L.orc_prop_kat.restype = C.c_double
        L.orc_prop.argtypes = [_S, C.c_double, C.c_double, C.c_double, _D]
        L.orc_prop.restype = C.c_double
        L.orc_test_integrate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _D]
        L.orc_test_quadgk.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_test_quadgk.restype = C.c_double
        L.orc_compute_corrs.argtypes = [_S, C.c_double, _A, _A, C.c_int, _D, _D]
        L.orc_compute_element.argtypes = [_S, C.c_int, C.c_double, _A, _A, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, _D, _A]
        L.orc_compute_element.restype = C.c_long
        L.orc_integrand.argtypes = [_S, C.c_int, C.c_double, _A, _A, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, _D]
        L.orc_run_pass.argtypes = [_S, C.c_int, C.c_double, _A, _A, _D, C.c_int, C.c_int, _D, _A]
        L.orc_run_pass.restype = C.c_long
        L.orc_symmetrize.argtypes = [_S, _A]
        L.orc_symmetrize_add_to.argtypes = [_S, _A, _A]
        L.orc_compute_dsigma.argtypes = [_S, C.c_double, _A, _A, _D]
        L.orc_compute_dsigma_corr.argtypes = [_S, C.c_double, _A, _A, _A, _D]
        L.orc_compute_dgamma_1l.argtypes = [_S, C.c_double, _A, _A, C.c_int, _D, _D]
        L.orc_compute_dgamma_1l.restype = C.c_long
        L.orc_compute_dgamma_2l.argtypes = [_S, C.c_double, _A, _A, _A, C.c_int, _D, _D]
        L.orc_compute_dgamma_2l.restype = C.c_long
        L.orc_compute_dgamma_ml.argtypes = [_S, C.c_double, C.c_int, _A, _A, _A, _A, _A, _A, C.c_int, _D, _D]
        L.orc_compute_dgamma_ml.restype = C.c_long
        L.orc_compute_sigma_sde.argtypes = [_S, C.c_double, _A, _A, C.c_int, _D, _D, _D]
        L.orc_compute_sigma_sde.restype = C.c_long
        L.orc_compute_gamma_bse.argtypes = [_S, C.c_double, _A, _A, C.c_int, _D, _D]
        L.orc_compute_gamma_bse.restype = C.c_long
        L.orc_compute_chi.argtypes = [_S, C.c_double, _A, _D, C.c_int, _D, _D, C.POINTER(C.c_int32)]
        L.orc_compute_chi.restype = None
        L.orc_test_hcubature.argtypes = [C.c_int, C.c_int, C.c_int, _D, _D, C.c_double, C.c_double,
                                         C.POINTER(C.c_int32)]
        L.orc_test_hcubature.restype = C.c_double
        L.orc_pf_expm1.argtypes = [C.c_double]
        L.orc_pf_expm1.restype = C.c_double
        L.orc_pf_exp.argtypes = [C.c_double]
        L.orc_pf_exp.restype = C.c_double
        L.orc_scan.argtypes = [_D, _D, C.c_int, _D, _D]
        L.orc_scan.restype = C.c_int
        L.orc_scan_cuts.argtypes = [_S, _A, C.POINTER(C.c_int32), _D]
        L.orc_scan_cuts.restype = None
        L.orc_resample.argtypes = [_S, _D, _D, _D, _D, _D, _A, _A]
        L.orc_resample.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _tol(t):
    return np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(2))


class Oracle:
    """Holds the system tables for one (lattice, mesh) pair and runs the FP64 CPU restatement
    on host Action objects (pffrg.action.Action)."""

    def __init__(self, r, m, S: float = 0.5, nc: int = 2):
        self.L = lib()
        off, rows = r.overlap_csr()
        self._keep = dict(
            sigma=np.ascontiguousarray(m.sigma, dtype=np.float64), Os=np.ascontiguousarray(m.Os, dtype=np.float64),
            ns=np.ascontiguousarray(m.ns, dtype=np.float64), Ot=np.ascontiguousarray(m.Ot, dtype=np.float64),
            nt=np.ascontiguousarray(m.nt, dtype=np.float64),
            exchange=np.ascontiguousarray(r.exchange, dtype=np.int64),
            mult=np.ascontiguousarray(r.mult, dtype=np.float64), off=np.ascontiguousarray(off, dtype=np.int64),
            rows=np.ascontiguousarray(rows.reshape(-1) if rows.size else np.zeros(3, np.int64), dtype=np.int64))
        k = self._keep
        self.sys = System(r.n_sites, m.num_sigma, m.num_Omega, m.num_nu, S, _dp(k["sigma"]), _dp(k["Os"]),
                          _dp(k["ns"]), _dp(k["Ot"]), _dp(k["nt"]), k["exchange"].ctypes.data_as(C.POINTER(C.c_int64)),
                          _dp(k["mult"]), k["off"].ctypes.data_as(C.POINTER(C.c_int64)),
                          k["rows"].ctypes.data_as(C.POINTER(C.c_int64)), nc)
        self.m = m
        self.ns = r.n_sites
        self.nc = nc
        self.corrs = np.zeros(2 * 3 * m.num_Omega)

    def _act(self, a):
        assert a.Sigma.flags.c_contiguous and a.bare.flags.c_contiguous and a.gamma.flags.c_contiguous
        if getattr(a, "nc", 2) != self.nc:
            raise ValueError(f"action with {getattr(a, 'nc', 2)} components on an oracle of {self.nc}")
        return ActionS(_dp(a.Sigma), _dp(a.bare), _dp(a.gamma))

    def set_threads(self, n: int):
        self.L.orc_set_threads(n)

    def compute_dsigma(self, Lam, a, da, tol=(1e-8, 1e-4)):
        self.L.orc_compute_dsigma(C.byref(self.sys), Lam, C.byref(self._act(a)), C.byref(self._act(da)), _dp(_tol(tol)))

    def compute_dsigma_corr(self, Lam, a, da, daS, tol=(1e-8, 1e-4)):
        self.L.orc_compute_dsigma_corr(C.byref(self.sys), Lam, C.byref(self._act(a)), C.byref(self._act(da)),
                                       C.byref(self._act(daS)), _dp(_tol(tol)))

    def compute_dgamma_1l(self, Lam, a, da, eval_, tol=(1e-8, 1e-4)) -> int:
        return self.L.orc_compute_dgamma_1l(C.byref(self.sys), Lam, C.byref(self._act(a)), C.byref(self._act(da)),
                                            eval_, _dp(_tol(tol)), _dp(self.corrs))

    def compute_dgamma_2l(self, Lam, a, da, dal, eval_, tol=(1e-8, 1e-4)) -> int:
        return self.L.orc_compute_dgamma_2l(C.byref(self.sys), Lam, C.byref(self._act(a)), C.byref(self._act(da)),
                                            C.byref(self._act(dal)), eval_, _dp(_tol(tol)), _dp(self.corrs))

    def compute_dgamma_ml(self, Lam, loops, a, da, dal, dac, datemp, daS, eval_, tol=(1e-8, 1e-4)) -> int:
        acts = [C.byref(self._act(x)) for x in (a, da, dal, dac, datemp, daS)]
        return self.L.orc_compute_dgamma_ml(C.byref(self.sys), Lam, loops, *acts, eval_, _dp(_tol(tol)),
                                            _dp(self.corrs))

    def compute_sigma(self, Lam, a1, a2, eval_, gtol=(1e-8, 1e-4), stol=(1e-8, 1e-4)) -> int:
        return self.L.orc_compute_sigma_sde(C.byref(self.sys), Lam, C.byref(self._act(a1)), C.byref(self._act(a2)),
                                            eval_, _dp(_tol(gtol)), _dp(_tol(stol)), _dp(self.corrs))

    def compute_gamma(self, Lam, a1, a2, eval_, tol=(1e-8, 1e-4)) -> int:
        return self.L.orc_compute_gamma_bse(C.byref(self.sys), Lam, C.byref(self._act(a1)), C.byref(self._act(a2)),
                                            eval_, _dp(_tol(tol)), _dp(self.corrs))

    def compute_corrs(self, Lam, a, da, kat: bool, tol=(1e-8, 1e-4)) -> np.ndarray:
        self.L.orc_compute_corrs(C.byref(self.sys), Lam, C.byref(self._act(a)), C.byref(self._act(da)), int(kat),
                                 _dp(self.corrs), _dp(_tol(tol)))
        return self.corrs.copy()

    def run_pass(self, pass_, Lam, a, da, out, full: bool, eval_, tol=(1e-8, 1e-4)) -> int:
        return self.L.orc_run_pass(C.byref(self.sys), pass_, Lam, C.byref(self._act(a)), C.byref(self._act(da)),
                                   _dp(self.corrs), int(full), eval_, _dp(_tol(tol)), C.byref(self._act(out)))

    def compute_element(self, pass_, Lam, a, da, ch, kernel, w1, w2, w3, eval_, out, tol=(1e-8, 1e-4)) -> int:
        return self.L.orc_compute_element(C.byref(self.sys), pass_, Lam, C.byref(self._act(a)), C.byref(self._act(da)),
                                          _dp(self.corrs), ch, kernel, w1, w2, w3, eval_, _dp(_tol(tol)),
                                          C.byref(self._act(out)))

    def integrand(self, pass_, Lam, a, da, ch, kernel, w1, w2, w3, v) -> np.ndarray:
        out = np.zeros((self.nc, self.ns))
        self.L.orc_integrand(C.byref(self.sys), pass_, Lam, C.byref(self._act(a)), C.byref(self._act(da)), ch, kernel,
                             w1, w2, w3, v, _dp(out))
        return out

    def compute_chi(self, Lam, a, tol=(1e-8, 1e-4), chi_mesh=None, with_regions=False):
        """compute_χ! (correlation_su2.jl:58-86): chi[site, w] over the irreducible sites; for the
        U(1)-DM action (correlation_u1_dm.jl:138-190) chi[c, site, w] for χxx, χzz, χxy."""
        cm = np.ascontiguousarray(self.m.chi if chi_mesh is None else chi_mesh, dtype=np.float64)
        ncc = 1 if self.nc == 2 else 3
        out = np.zeros(ncc * self.ns * len(cm))
        nreg = np.zeros(ncc * 9 * self.ns * len(cm), dtype=np.int32)
        self.L.orc_compute_chi(C.byref(self.sys), Lam, C.byref(self._act(a)), _dp(cm), len(cm), _dp(_tol(tol)),
                               _dp(out), nreg.ctypes.data_as(C.POINTER(C.c_int32)))
        chi = out.reshape(ncc, len(cm), self.ns).transpose(0, 2, 1).copy()
        if ncc == 1:
            chi = chi[0]
        return (chi, nreg) if with_regions else chi

    def scan_cuts(self, a):
        """scan_channel's argmax positions idx[3 nc, 4] and cuts[3 nc, nO + 3 nn] (Action.jl:449-476)."""
        k = 3 * self.nc
        idx = np.zeros(4 * k, dtype=np.int32)
        cuts = np.zeros(k * (self.m.num_Omega + 3 * self.m.num_nu))
        self.L.orc_scan_cuts(C.byref(self.sys), C.byref(self._act(a)), idx.ctypes.data_as(C.POINTER(C.c_int32)),
                             _dp(cuts))
        return idx.reshape(k, 4), cuts.reshape(k, -1)

    def resample(self, a_old, m_new, a_new):
        """resample_from_to's interpolation of a_old (on this oracle's meshes) onto m_new."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (m_new.sigma, m_new.Os, m_new.ns, m_new.Ot,
                                                                    m_new.nt)]
        self.L.orc_resample(C.byref(self.sys), *[_dp(x) for x in arrs], C.byref(self._act(a_old)),
                            C.byref(self._act(a_new)))

    def symmetrize(self, a):
        self.L.orc_symmetrize(C.byref(self.sys), C.byref(self._act(a)))

    def symmetrize_add_to(self, al, a):
        self.L.orc_symmetrize_add_to(C.byref(self.sys), C.byref(self._act(al)), C.byref(self._act(a)))

    # low-level probes
    def get_buffer(self, ch, w, v, vp):
        """[exch, map, sgn_mu, sgn_nu, kernel, lo1, hi1, lo2, hi2, lo3, hi3] and the weights; the
        sgn_munu flag (read by the U(1)-DM action only) is get_buffer_munu's."""
        oi = (C.c_int32 * 12)()
        ow = (C.c_double * 6)()
        self.L.orc_get_buffer(C.byref(self.sys), ch, w, v, vp, oi, ow)
        return list(oi)[:11], list(ow)

    def get_buffer_munu(self, ch, w, v, vp) -> bool:
        oi = (C.c_int32 * 12)()
        ow = (C.c_double * 6)()
        self.L.orc_get_buffer(C.byref(self.sys), ch, w, v, vp, oi, ow)
        return bool(oi[11])

    def get_vertex(self, a, comp, ch, w, v, vp, site) -> float:
        return self.L.orc_get_vertex(C.byref(self.sys), C.byref(self._act(a)), comp, ch, w, v, vp, site)

    def get_gamma_avx(self, a, freqs, mask=7) -> np.ndarray:
        out = np.zeros(self.nc * self.ns)
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        self.L.orc_get_gamma_avx(C.byref(self.sys), C.byref(self._act(a)), _dp(f), mask, _dp(out))
        return out.reshape(self.nc, self.ns)

    def get_sigma(self, Sig, w) -> float:
        return self.L.orc_get_sigma(C.byref(self.sys), _dp(np.ascontiguousarray(Sig)), w)

    def G(self, Lam, w, Sig) -> float:
        return self.L.orc_G(C.byref(self.sys), Lam, w, _dp(np.ascontiguousarray(Sig)))

    def S(self, Lam, w, Sig) -> float:
        return self.L.orc_S(C.byref(self.sys), Lam, w, _dp(np.ascontiguousarray(Sig)))


def scan(x, y, p) -> float:
    """scan (Action.jl:369-446): new linear extent for the cut y on mesh x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    d = np.zeros(1)
    if lib().orc_scan(_dp(x), _dp(y), len(x), _dp(pp), _dp(d)) != 0:
        raise ValueError("scan: invalid mesh")
    return float(d[0])


def get_param(val: float, mesh: np.ndarray):
    lo, hi = C.c_int32(), C.c_int32()
    wl, wh = C.c_double(), C.c_double()
    mesh = np.ascontiguousarray(mesh, dtype=np.float64)
    lib().orc_get_param(val, _dp(mesh), len(mesh), C.byref(lo), C.byref(hi), C.byref(wl), C.byref(wh))
    return lo.value, hi.value, wl.value, wh.value


def test_integrate(mode: int, n: int, a: float, b: float, eval_: int, atol: float, rtol: float) -> np.ndarray:
    out = np.zeros(n)
    lib().orc_test_integrate(mode, n, a, b, eval_, atol, rtol, _dp(out))
    return out


def test_quadgk(kind: int, s0: float, atol: float = 1e-10, rtol: float = 1e-10) -> float:
    return lib().orc_test_quadgk(kind, s0, atol, rtol)


def G_bare(Lam: float, w: float) -> float:
    return lib().orc_G_bare(Lam, w)
