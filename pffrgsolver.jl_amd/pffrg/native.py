"""ctypes binding of libpffrg.so (the C ABI declared in include/pffrg.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, every call raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .action import Action

LIB_PATH = Path(os.environ.get("PFFRG_LIB", Path(__file__).resolve().parent / "_lib" / "libpffrg.so"))

FP64 = 0
FP32_MIXED = 1
FP32_ELEMENT = 2
# action symmetries of pffrg_create_symmetry (get_action_empty, Action.jl:564-577)
SU2 = 0
U1_DM = 1

_lib = None


class PffrgError(RuntimeError):
    pass


class Dims(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("n_sigma", C.c_int32), ("n_Omega", C.c_int32), ("n_nu", C.c_int32),
                ("S", C.c_double), ("n_slots", C.c_int32), ("device", C.c_int32)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_i = C.c_int32
_d = C.c_double

_SIGS = {
    "pffrg_create": ([C.POINTER(Dims), C.POINTER(_P)], C.c_int),
    "pffrg_create_symmetry": ([C.POINTER(Dims), _i, C.POINTER(_P)], C.c_int),
    "pffrg_n_components": ([_P], C.c_int32),
    "pffrg_destroy": ([_P], None),
    "pffrg_last_error": ([_P], C.c_char_p),
    "pffrg_set_mesh": ([_P, _D, _D, _D, _D, _D], C.c_int),
    "pffrg_set_lattice": ([_P, _I64, _I64, _I64, _I64, _i], C.c_int),
    "pffrg_upload_action": ([_P, _i, _D, _D, _D], C.c_int),
    "pffrg_download_action": ([_P, _i, _D, _D, _D], C.c_int),
    "pffrg_gamma_length": ([_P], C.c_int64),
    "pffrg_set_precision": ([_P, _i], C.c_int),
    "pffrg_compute_dsigma": ([_P, _d, _i, _i, _D], C.c_int),
    "pffrg_compute_dsigma_corr": ([_P, _d, _i, _i, _i, _D], C.c_int),
    "pffrg_compute_dgamma_1l": ([_P, _d, _i, _i, _i, _D], C.c_int),
    "pffrg_compute_dgamma_2l": ([_P, _d, _i, _i, _i, _i, _D], C.c_int),
    "pffrg_compute_dgamma_ml": ([_P, _d, _i, _i, _i, _i, _i, _i, _i, _i, _D], C.c_int),
    "pffrg_compute_sigma": ([_P, _d, _i, _i, _i, _D, _D], C.c_int),
    "pffrg_compute_gamma": ([_P, _d, _i, _i, _i, _D], C.c_int),
    "pffrg_symmetrize": ([_P, _i], C.c_int),
    "pffrg_symmetrize_add_to": ([_P, _i, _i], C.c_int),
    "pffrg_replace_with": ([_P, _i, _i, _i], C.c_int),
    "pffrg_mult_with": ([_P, _i, _d, _i], C.c_int),
    "pffrg_mult_with_add_to": ([_P, _i, _d, _i, _i], C.c_int),
    "pffrg_abs_max": ([_P, _i, _D], C.c_int),
    "pffrg_get_sigma": ([_P, _i, _D], C.c_int),
    "pffrg_set_sigma": ([_P, _i, _D], C.c_int),
    "pffrg_synchronize": ([_P], C.c_int),
    "pffrg_reset_stats": ([_P], C.c_int),
    "pffrg_get_stats": ([_P, _I64], C.c_int),
    "pffrg_get_corrs": ([_P, _D], C.c_int),
    "pffrg_probe_integrand": ([_P, _i, _d, _i, _i, _i, _i, _i, _i, _i, _i, _D, _D], C.c_int),
    "pffrg_debug_record": ([_P, C.POINTER(C.c_int32), _D, C.c_int64], C.c_int),
    "pffrg_set_timing": ([_P, _i], C.c_int),
    "pffrg_get_timing": ([_P, _D, _I64], C.c_int),
    "pffrg_gather_ceiling": ([_P, _i, _i, _i, C.c_int64, _i, _i, _D, _D], C.c_int),
    "pffrg_compute_chi": ([_P, _d, _i, _D, _i, _D, _D, C.POINTER(C.c_int32)], C.c_int),
    "pffrg_scan_cuts": ([_P, _i, C.POINTER(C.c_int32), _D], C.c_int),
    "pffrg_scan_cuts_noalloc": ([_P, _i, C.POINTER(C.c_int32), _D], C.c_int),
    "pffrg_get_counters": ([_P, _I64, _i], C.c_int),
    "pffrg_scan": ([_D, _D, _i, _D, _D], C.c_int),
    "pffrg_resample": ([_P, _i, _i, _D, _D, _D, _D, _D], C.c_int),
    "pffrg_mlp_create": ([_i, _i, _i, _i, C.c_uint64, _i, C.POINTER(_P)], C.c_int),
    "pffrg_mlp_destroy": ([_P], None),
    "pffrg_mlp_last_error": ([_P], C.c_char_p),
    "pffrg_mlp_num_params": ([_P], C.c_int64),
    "pffrg_mlp_get_params": ([_P, C.POINTER(C.c_float)], C.c_int),
    "pffrg_mlp_set_params": ([_P, C.POINTER(C.c_float)], C.c_int),
    "pffrg_mlp_forward": ([_P, C.POINTER(C.c_float), _i, C.POINTER(C.c_float)], C.c_int),
    "pffrg_mlp_backward": ([_P, C.POINTER(C.c_float), C.POINTER(C.c_float), _i, C.POINTER(C.c_float),
                            C.POINTER(C.c_float)], C.c_int),
    "pffrg_mlp_adam": ([_P, C.c_float, C.c_float, C.c_float, C.c_float, _i], C.c_int),
    "pffrg_decoder_create": ([C.c_int64, _i, C.POINTER(C.c_float), C.POINTER(C.c_float), _i, C.POINTER(_P)], C.c_int),
    "pffrg_decoder_destroy": ([_P], None),
    "pffrg_decoder_apply": ([_P, C.POINTER(C.c_float), _D], C.c_int),
    "pffrg_batch_create": ([C.POINTER(_P), _i, C.POINTER(_P)], C.c_int),
    "pffrg_batch_destroy": ([_P], None),
    "pffrg_batch_last_error": ([_P], C.c_char_p),
    "pffrg_batch_rhs_1l": ([_P, _D, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D, _D],
                           C.c_int),
    "pffrg_batch_compute_chi": ([_P, _D, C.POINTER(C.c_int32), C.POINTER(_D), C.POINTER(C.c_int32), _D,
                                 C.POINTER(_D)], C.c_int),
    "pffrg_set_shared_shape": ([_P, _i, _i, _i], C.c_int),
    "pffrg_get_shared_shape": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "pffrg_tune_shared": ([_P, C.c_double, _i, _i, _i, _D, C.POINTER(C.c_int32), _D, C.POINTER(C.c_int32)], C.c_int),
    "pffrg_batch_replace_with": ([_P, _i, _i, _i], C.c_int),
    "pffrg_batch_mult_with_add_to": ([_P, _i, _D, _i, _i], C.c_int),
    "pffrg_batch_abs_max": ([_P, _i, _D], C.c_int),
}


def load() -> C.CDLL:
    """Load libpffrg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PffrgError(f"{LIB_PATH} not found: build it with __graft_entry__.build()")
        lib = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def scan(x, y, p) -> float:
    """scan (Action.jl:369-446) on the host: the new linear extent of mesh x for the cut y, with
    p = (p0, p1, p2, p3*Λ, p4*Λ)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    if len(x) != len(y) or len(pp) != 5:
        raise PffrgError("scan: x, y must have equal length and p five entries")
    d = np.zeros(1)
    rc = load().pffrg_scan(_dp(x), _dp(y), len(x), _dp(pp), _dp(d))
    if rc != 0:
        raise PffrgError(f"pffrg_scan failed ({rc})")
    return float(d[0])


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ip(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I64)


def _tol(t) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(2))


class Context:
    """One device context: meshes, lattice tables and n_slots action slots on the GPU, for the
    SU(2) action (nc = 2) or the U(1)-DM action (nc = 6)."""

    def __init__(self, n_sites: int, m, n_slots: int, S: float = 0.5, device: int = 0, nc: int = 2):
        self.lib = load()
        self.m = m
        self.ns = n_sites
        self.S = S
        if nc not in (2, 6):
            raise PffrgError(f"vertex components must be 2 (SU(2)) or 6 (U(1)-DM), not {nc}")
        self.nc = nc
        d = Dims(n_sites, m.num_sigma, m.num_Omega, m.num_nu, S, n_slots, device)
        h = _P()
        rc = self.lib.pffrg_create_symmetry(C.byref(d), SU2 if nc == 2 else U1_DM, C.byref(h))
        if rc != 0 or not h:
            raise PffrgError(f"pffrg_create failed with code {rc} (no HIP device, or dims out of range)")
        self.h = h
        self.n_slots = n_slots
        self.set_mesh(m)

    def _ck(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.pffrg_last_error(self.h)
            raise PffrgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pffrg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- setup
    def set_mesh(self, m):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (m.sigma, m.Os, m.ns, m.Ot, m.nt)]
        self._ck(self.lib.pffrg_set_mesh(self.h, *[_dp(a) for a in arrs]), "set_mesh")
        self.m = m

    def set_lattice(self, r):
        off, rows = r.overlap_csr()
        ex = np.ascontiguousarray(r.exchange, dtype=np.int64)
        mult = np.ascontiguousarray(np.rint(r.mult), dtype=np.int64)
        off = np.ascontiguousarray(off, dtype=np.int64)
        rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1)
        if rows.size == 0:
            rows = np.zeros(3, dtype=np.int64)
        self._ck(self.lib.pffrg_set_lattice(self.h, _ip(ex), _ip(mult), _ip(off), _ip(rows), 0), "set_lattice")

    def set_precision(self, mode: int):
        self._ck(self.lib.pffrg_set_precision(self.h, mode), "set_precision")

    # ---------------------------------------------------------------- transfer
    def upload(self, slot: int, a: Action):
        if a.nc != self.nc:
            raise PffrgError(f"upload: action of {a.nc} components into a context of {self.nc}")
        self._ck(self.lib.pffrg_upload_action(self.h, slot, _dp(np.ascontiguousarray(a.Sigma)),
                                              _dp(np.ascontiguousarray(a.bare.reshape(-1))),
                                              _dp(np.ascontiguousarray(a.gamma))), "upload_action")

    def download(self, slot: int) -> Action:
        a = Action.empty(self.ns, self.m, self.S, nc=self.nc)
        bare = np.zeros(self.nc * self.ns)
        self._ck(self.lib.pffrg_download_action(self.h, slot, _dp(a.Sigma), _dp(bare), _dp(a.gamma)), "download_action")
        a.bare[:] = bare.reshape(self.nc, self.ns)
        return a

    def get_sigma(self, slot: int) -> np.ndarray:
        s = np.zeros(self.m.num_sigma)
        self._ck(self.lib.pffrg_get_sigma(self.h, slot, _dp(s)), "get_sigma")
        return s

    def set_sigma(self, slot: int, s: np.ndarray):
        self._ck(self.lib.pffrg_set_sigma(self.h, slot, _dp(np.ascontiguousarray(s, dtype=np.float64))), "set_sigma")

    # ---------------------------------------------------------------- flow right-hand side
    def compute_dsigma(self, L, a, da, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_dsigma(self.h, L, a, da, _dp(_tol(tol))), "compute_dsigma")

    def compute_dsigma_corr(self, L, a, da, da_sigma, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_dsigma_corr(self.h, L, a, da, da_sigma, _dp(_tol(tol))), "compute_dsigma_corr")

    def compute_dgamma_1l(self, L, a, da, eval_, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_dgamma_1l(self.h, L, a, da, eval_, _dp(_tol(tol))), "compute_dgamma_1l")

    def shared_shape(self) -> tuple:
        """(waves per block, batch capacity, most nodes per batch) of the shared-node pass."""
        out = np.zeros(3, dtype=np.int32)
        self._ck(self.lib.pffrg_get_shared_shape(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))), "get_shared_shape")
        return tuple(int(x) for x in out)

    def set_shared_shape(self, waves, batch_nodes, max_nodes):
        self._ck(self.lib.pffrg_set_shared_shape(self.h, waves, batch_nodes, max_nodes), "set_shared_shape")

    def tune_shared(self, L, a, da, eval_, tol=(1e-8, 1e-4)):
        """The autotuner: times every candidate shape on this problem's Katanin pass (slot a ->
        slot da, overwritten) and keeps the fastest; returns (best shape, [ms per candidate])."""
        best = np.zeros(3, dtype=np.int32)
        ms = np.zeros(32)
        n = C.c_int32(len(ms))
        self._ck(self.lib.pffrg_tune_shared(self.h, L, a, da, eval_, _dp(_tol(tol)), best.ctypes.data_as(C.POINTER(C.c_int32)),
                                            _dp(ms), C.byref(n)), "tune_shared")
        return tuple(int(x) for x in best), [float(x) for x in ms[:n.value]]

    def compute_dgamma_2l(self, L, a, da, da_l, eval_, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_dgamma_2l(self.h, L, a, da, da_l, eval_, _dp(_tol(tol))), "compute_dgamma_2l")

    def compute_dgamma_ml(self, L, loops, a, da, da_l, da_c, da_temp, da_sigma, eval_, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_dgamma_ml(self.h, L, loops, a, da, da_l, da_c, da_temp, da_sigma, eval_,
                                                  _dp(_tol(tol))), "compute_dgamma_ml")

    def compute_sigma(self, L, a1, a2, eval_, gtol=(1e-8, 1e-4), stol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_sigma(self.h, L, a1, a2, eval_, _dp(_tol(gtol)), _dp(_tol(stol))),
                 "compute_sigma")

    def compute_gamma(self, L, a1, a2, eval_, tol=(1e-8, 1e-4)):
        self._ck(self.lib.pffrg_compute_gamma(self.h, L, a1, a2, eval_, _dp(_tol(tol))), "compute_gamma")

    # ---------------------------------------------------------------- algebra
    def compute_chi(self, L, a, tol=(1e-8, 1e-4), chi_mesh=None, with_regions=False):
        """compute_χ! (correlation_su2.jl:58-86): chi[site, k] over the irreducible sites and
        the χ frequencies (the mesh's χ by default); for the U(1)-DM action (correlation_u1_dm.jl:
        138-190) chi[c, site, k] with c = χxx, χzz, χxy."""
        cm = np.ascontiguousarray(self.m.chi if chi_mesh is None else chi_mesh, dtype=np.float64)
        ncc = 1 if self.nc == 2 else 3
        out = np.zeros(ncc * self.ns * len(cm))
        nreg = np.zeros(ncc * 9 * self.ns * len(cm), dtype=np.int32)
        self._ck(self.lib.pffrg_compute_chi(self.h, L, a, _dp(cm), len(cm), _dp(_tol(tol)), _dp(out),
                                            nreg.ctypes.data_as(C.POINTER(C.c_int32))), "compute_chi")
        chi = out.reshape(ncc, len(cm), self.ns).transpose(0, 2, 1).copy()
        if ncc == 1:
            chi = chi[0]
        return (chi, nreg) if with_regions else chi

    def scan_cuts(self, a):
        """scan_channel's argmax positions idx[3*comp + ch] = (site, w, v, v') and the cuts through
        them, cuts[3*comp + ch] = (Ω cut, three ν cuts) (Action.jl:449-476)."""
        k = 3 * self.nc
        idx = np.zeros(4 * k, dtype=np.int32)
        cuts = np.zeros(k * (self.m.num_Omega + 3 * self.m.num_nu))
        self._ck(self.lib.pffrg_scan_cuts(self.h, a, idx.ctypes.data_as(C.POINTER(C.c_int32)), _dp(cuts)),
                 "scan_cuts")
        return idx.reshape(k, 4), cuts.reshape(k, -1)

    def resample(self, src, dst, m_new):
        """resample_from_to's interpolation (Action.jl:543-553): slot src onto the meshes m_new into
        slot dst; m_new becomes the context's mesh (its χ mesh included)."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (m_new.sigma, m_new.Os, m_new.ns, m_new.Ot,
                                                                    m_new.nt)]
        if [len(x) for x in arrs] != [self.m.num_sigma, self.m.num_Omega, self.m.num_nu, self.m.num_Omega,
                                      self.m.num_nu]:
            raise PffrgError("resample: new meshes must keep the point counts")
        self._ck(self.lib.pffrg_resample(self.h, src, dst, *[_dp(x) for x in arrs]), "resample")
        self.m = m_new

    def symmetrize(self, s):
        self._ck(self.lib.pffrg_symmetrize(self.h, s), "symmetrize")

    def symmetrize_add_to(self, l, a):
        self._ck(self.lib.pffrg_symmetrize_add_to(self.h, l, a), "symmetrize_add_to")

    def replace_with(self, dst, src, vertex_only=False):
        self._ck(self.lib.pffrg_replace_with(self.h, dst, src, int(vertex_only)), "replace_with")

    def mult_with(self, s, fac, vertex_only=False):
        self._ck(self.lib.pffrg_mult_with(self.h, s, fac, int(vertex_only)), "mult_with")

    def mult_with_add_to(self, src, fac, dst, vertex_only=False):
        self._ck(self.lib.pffrg_mult_with_add_to(self.h, src, fac, dst, int(vertex_only)), "mult_with_add_to")

    def abs_max(self, s) -> float:
        out = C.c_double()
        self._ck(self.lib.pffrg_abs_max(self.h, s, C.byref(out)), "abs_max")
        return out.value

    # ---------------------------------------------------------------- measurement
    def synchronize(self):
        self._ck(self.lib.pffrg_synchronize(self.h), "synchronize")

    def reset_stats(self):
        self._ck(self.lib.pffrg_reset_stats(self.h), "reset_stats")

    def stats(self) -> dict:
        s = np.zeros(4, dtype=np.int64)
        self._ck(self.lib.pffrg_get_stats(self.h, _ip(s)), "get_stats")
        return {"nodes": int(s[0]), "elements": int(s[1]), "launches": int(s[2]), "gathered_bytes": int(s[3])}

    def counters(self, n: int = 32) -> np.ndarray:
        s = np.zeros(n, dtype=np.int64)
        self._ck(self.lib.pffrg_get_counters(self.h, _ip(s), n), "get_counters")
        return s

    def probe_integrand(self, pass_, L, a, da, ch, kernel, w1, w2, w3, v) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros((len(v), self.nc, self.ns))
        self._ck(self.lib.pffrg_probe_integrand(self.h, pass_, L, a, da, ch, kernel, w1, w2, w3, len(v), _dp(v),
                                                out.ctypes.data_as(_D)), "probe_integrand")
        return out

    def corrs(self) -> np.ndarray:
        out = np.zeros(2 * 3 * self.m.num_Omega)
        self._ck(self.lib.pffrg_get_corrs(self.h, _dp(out)), "get_corrs")
        return out

    def set_timing(self, enable: bool):
        self._ck(self.lib.pffrg_set_timing(self.h, int(enable)), "set_timing")

    def timing(self) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_int64()
        self._ck(self.lib.pffrg_get_timing(self.h, C.byref(ms), C.byref(n)), "get_timing")
        return ms.value, n.value

    def gather_ceiling(self, slot: int, precision: int, waves_per_simd: int = 0, footprint: int = 0,
                       n_iter: int = 256, n_rep: int = 5) -> tuple[float, float]:
        """GB/s and ms per launch of random trilinear q3 lookups on the slot's gather copy
        (pffrg_gather_ceiling): the practical ceiling of the vertex pass's gathers, over the
        whole q3 block (footprint 0) or confined to about `footprint` bytes."""
        gbps = C.c_double()
        ms = C.c_double()
        self._ck(self.lib.pffrg_gather_ceiling(self.h, slot, precision, waves_per_simd, footprint, n_iter, n_rep,
                                               C.byref(gbps), C.byref(ms)), "gather_ceiling")
        return gbps.value, ms.value


class Batch:
    """Several contexts of one device whose 1-loop right-hand sides run as one vertex-pass
    launch (pffrg_batch_*).  The contexts must stay open while the batch is used."""

    def __init__(self, ctxs):
        self.lib = load()
        self.ctxs = list(ctxs)
        arr = (_P * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = _P()
        rc = self.lib.pffrg_batch_create(arr, len(self.ctxs), C.byref(h))
        if rc != 0 or not h:
            raise PffrgError(f"pffrg_batch_create failed with code {rc}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.pffrg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def rhs_1l(self, Lambdas, a, da, evals, sigma_tol=(1e-8, 1e-4), gamma_tol=(1e-8, 1e-4)):
        """compute_dΣ! + compute_dΓ_1l! of every problem: slot a[k] -> slot da[k] at Lambdas[k]."""
        n = len(self.ctxs)
        L = np.ascontiguousarray(np.broadcast_to(np.asarray(Lambdas, dtype=np.float64), (n,)))
        ia = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        ida = np.ascontiguousarray(np.broadcast_to(np.asarray(da, dtype=np.int32), (n,)))
        iev = np.ascontiguousarray(np.broadcast_to(np.asarray(evals, dtype=np.int32), (n,)))
        ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
        rc = self.lib.pffrg_batch_rhs_1l(self.h, _dp(L), ip(ia), ip(ida), ip(iev), _dp(_tol(sigma_tol)),
                                         _dp(_tol(gamma_tol)))
        if rc != 0:
            msg = self.lib.pffrg_batch_last_error(self.h)
            raise PffrgError(f"batch_rhs_1l failed ({rc}): {msg.decode() if msg else ''}")

    def _bck(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.pffrg_batch_last_error(self.h)
            raise PffrgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def replace_with(self, dst, src, vertex_only=False):
        """replace_with! of every problem: slot dst <- slot src."""
        self._bck(self.lib.pffrg_batch_replace_with(self.h, dst, src, int(vertex_only)), "batch_replace_with")

    def mult_with_add_to(self, src, facs, dst, vertex_only=False):
        """mult_with_add_to! of every problem: slot dst += facs[k] * slot src."""
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(facs, dtype=np.float64), (len(self.ctxs),)))
        self._bck(self.lib.pffrg_batch_mult_with_add_to(self.h, src, _dp(f), dst, int(vertex_only)),
                  "batch_mult_with_add_to")

    def abs_max(self, slot) -> np.ndarray:
        """get_abs_max of every problem's slot."""
        out = np.zeros(len(self.ctxs))
        self._bck(self.lib.pffrg_batch_abs_max(self.h, slot, _dp(out)), "batch_abs_max")
        return out

    def compute_chi(self, Lambdas, a, tol=(1e-8, 1e-4), chi_meshes=None, active=None):
        """compute_χ! of every problem (slot a[k] at Lambdas[k], the context's χ mesh by default)
        in one set of cubature rounds; returns chi[k][site, w] as Context.compute_chi does, or
        None for a problem that active[k] leaves out."""
        n = len(self.ctxs)
        L = np.ascontiguousarray(np.broadcast_to(np.asarray(Lambdas, dtype=np.float64), (n,)))
        ia = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        on = [True] * n if active is None else [bool(x) for x in active]
        meshes = [np.ascontiguousarray(c.m.chi if chi_meshes is None else chi_meshes[k], dtype=np.float64)
                  for k, c in enumerate(self.ctxs)]
        outs = [np.zeros(c.ns * len(m)) for c, m in zip(self.ctxs, meshes)]
        pm = (_D * n)(*[_dp(m) for m in meshes])
        po = (_D * n)(*[_dp(o) for o in outs])
        nchi = np.ascontiguousarray([len(m) if on[k] else 0 for k, m in enumerate(meshes)], dtype=np.int32)
        rc = self.lib.pffrg_batch_compute_chi(self.h, _dp(L), ia.ctypes.data_as(C.POINTER(C.c_int32)), pm,
                                              nchi.ctypes.data_as(C.POINTER(C.c_int32)), _dp(_tol(tol)), po)
        if rc != 0:
            msg = self.lib.pffrg_batch_last_error(self.h)
            raise PffrgError(f"batch_compute_chi failed ({rc}): {msg.decode() if msg else ''}")
        return [o.reshape(len(m), c.ns).T.copy() if on[k] else None
                for k, (o, m, c) in enumerate(zip(outs, meshes, self.ctxs))]
