"""Observable and checkpoint files in the layout PFFRGSolver.jl writes through HDF5.jl, over the
host library libpffrg_h5.so (include/pffrg_h5.h).

Layout (all names as the reference composes them, Λ printed as Julia prints a Float64):
  observables: "symmetry", "σ/Λ", "Σ/Λ" (save_self!, Action/disk.jl:35-49), "χ/Λ/mesh" and
               "χ/Λ/diag" (save_χ!, Observable/disk.jl:2-29), "reduced_lattice/{name, size,
               model, euclidean, J/i}" (Lattice/disk.jl:2-19);
  checkpoints: "dΛ/Λ", "σ/Λ", "Ωs/Λ", "νs/Λ", "Ωt/Λ", "νt/Λ", "χ/Λ", "S", "symmetry",
               "a/Λ/Σ", "a/Λ/Γ/{spin,dens}/bare" and ".../ch_{s,t,u}/{q1,q2_1,q2_2,q3}"
               (checkpoint!, checkpoint_su2.jl:2-39; save!, Action/disk.jl:2-14, 80-95).
Arrays keep the Julia (column-major) shapes: q3 is (n_sites, n_Ω, n_ν, n_ν), χ is
(n_sites, n_χ).  measure() restates the reference's per-step bookkeeping (Launcher.jl:2-75)."""
from __future__ import annotations

import ctypes as C
import os
import time
from decimal import Decimal
from pathlib import Path

import numpy as np

from .action import Action
from .mesh import Mesh

LIB_PATH = Path(os.environ.get("PFFRG_H5_LIB", Path(__file__).resolve().parent / "_lib" / "libpffrg_h5.so"))

READ, WRITE, CREATE_WRITE = 0, 1, 2
F64, I64, BOOL, STR = 0, 1, 2, 3

_lib = None

_P, _I64P, _DP = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)
_SIGS = {
    "pffrg_h5_open": ([C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "pffrg_h5_close": ([_P], None),
    "pffrg_h5_last_error": ([_P], C.c_char_p),
    "pffrg_h5_exists": ([_P, C.c_char_p], C.c_int),
    "pffrg_h5_write_f64": ([_P, C.c_char_p, C.c_int32, _I64P, _DP], C.c_int),
    "pffrg_h5_write_i64": ([_P, C.c_char_p, C.c_int32, _I64P, _I64P], C.c_int),
    "pffrg_h5_write_bool": ([_P, C.c_char_p, C.c_int32], C.c_int),
    "pffrg_h5_write_str": ([_P, C.c_char_p, C.c_char_p], C.c_int),
    "pffrg_h5_shape": ([_P, C.c_char_p, C.POINTER(C.c_int32), _I64P, C.POINTER(C.c_int32)], C.c_int),
    "pffrg_h5_read_f64": ([_P, C.c_char_p, C.c_int64, _DP], C.c_int),
    "pffrg_h5_read_i64": ([_P, C.c_char_p, C.c_int64, _I64P], C.c_int),
    "pffrg_h5_read_bool": ([_P, C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
    "pffrg_h5_read_str": ([_P, C.c_char_p, C.c_char_p, C.c_int64], C.c_int),
    "pffrg_h5_keys": ([_P, C.c_char_p, C.c_char_p, C.c_int64, _I64P], C.c_int),
}


class H5Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise H5Error(f"{LIB_PATH} not found: build it with __graft_entry__.build() (needs the HDF5 headers)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    """The entry points this binding uses (include/pffrg_h5.h)."""
    return sorted(_SIGS)


def julia_float(x: float) -> str:
    """string(x::Float64) as Julia writes it (Base.Ryu shortest digits): fixed notation when the
    decimal point falls within -4 < pt <= 6, i.e. 1e-4 <= |x| < 1e6 ("0.0001", "50.0",
    "123456.7"), otherwise d.ddd followed by e and the exponent ("1.0e-5", "1.0e6")."""
    x = float(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if np.signbit(x) else "0.0"
    sign, digits, exp = Decimal(repr(abs(x))).as_tuple()
    d = "".join(map(str, digits)).rstrip("0")
    exp += len(digits) - len(d)
    pt = exp + len(d)  # decimal point position relative to the first digit
    s = "-" if x < 0 else ""
    if -4 < pt <= 6:
        if pt <= 0:
            return s + "0." + "0" * (-pt) + d
        if pt >= len(d):
            return s + d + "0" * (pt - len(d)) + ".0"
        return s + d[:pt] + "." + d[pt:]
    return s + d[0] + "." + (d[1:] or "0") + "e" + str(pt - 1)


class H5File:
    """One open HDF5 file (h5open).  mode: "r", "r+" or "cw" (read/write, create if missing)."""

    def __init__(self, path, mode: str = "cw"):
        self.lib = load()
        m = {"r": READ, "r+": WRITE, "cw": CREATE_WRITE}[mode]
        h = C.c_void_p()
        if self.lib.pffrg_h5_open(str(path).encode(), m, C.byref(h)) != 0:
            raise H5Error(f"cannot open {path} ({mode})")
        self.h = h
        self.path = str(path)

    def close(self):
        if self.h:
            self.lib.pffrg_h5_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc: int, what: str):
        if rc < 0:
            raise H5Error(f"{what}: {self.lib.pffrg_h5_last_error(self.h).decode()}")
        return rc

    def haskey(self, path: str) -> bool:
        return self._ck(self.lib.pffrg_h5_exists(self.h, path.encode()), f"haskey {path}") == 1

    def write(self, path: str, value):
        """file[path] = value for Float64/Int64 scalars and arrays, Bool and String."""
        p = path.encode()
        if isinstance(value, str):
            return self._ck(self.lib.pffrg_h5_write_str(self.h, p, value.encode()), f"write {path}")
        if isinstance(value, (bool, np.bool_)):
            return self._ck(self.lib.pffrg_h5_write_bool(self.h, p, int(value)), f"write {path}")
        a = np.asarray(value)
        # the Julia array's bytes: column-major; its dimensions are a.shape
        dims = np.array(a.shape, dtype=np.int64)
        dp = dims.ctypes.data_as(C.POINTER(C.c_int64)) if a.ndim else None
        if np.issubdtype(a.dtype, np.integer):
            buf = np.asfortranarray(a, dtype=np.int64).ravel(order="F")
            return self._ck(self.lib.pffrg_h5_write_i64(self.h, p, a.ndim, dp, buf.ctypes.data_as(C.POINTER(C.c_int64))),
                            f"write {path}")
        buf = np.asfortranarray(a, dtype=np.float64).ravel(order="F")
        return self._ck(self.lib.pffrg_h5_write_f64(self.h, p, a.ndim, dp, buf.ctypes.data_as(C.POINTER(C.c_double))),
                        f"write {path}")

    def shape(self, path: str) -> tuple[tuple[int, ...], int]:
        rank = C.c_int32()
        dims = np.zeros(8, dtype=np.int64)
        kind = C.c_int32()
        self._ck(self.lib.pffrg_h5_shape(self.h, path.encode(), C.byref(rank), dims.ctypes.data_as(C.POINTER(C.c_int64)),
                                         C.byref(kind)), f"shape {path}")
        return tuple(int(d) for d in dims[:rank.value]), kind.value

    def read(self, path: str):
        """read(file, path): Float64/Int64 arrays in their Julia shapes, scalars, Bool, String."""
        shape, kind = self.shape(path)
        p = path.encode()
        if kind == STR:
            buf = C.create_string_buffer(1 << 16)
            self._ck(self.lib.pffrg_h5_read_str(self.h, p, buf, len(buf)), f"read {path}")
            return buf.value.decode()
        if kind == BOOL:
            v = C.c_int32()
            self._ck(self.lib.pffrg_h5_read_bool(self.h, p, C.byref(v)), f"read {path}")
            return bool(v.value)
        n = int(np.prod(shape)) if shape else 1
        if kind == I64:
            out = np.zeros(n, dtype=np.int64)
            self._ck(self.lib.pffrg_h5_read_i64(self.h, p, n, out.ctypes.data_as(C.POINTER(C.c_int64))), f"read {path}")
        else:
            out = np.zeros(n, dtype=np.float64)
            self._ck(self.lib.pffrg_h5_read_f64(self.h, p, n, out.ctypes.data_as(C.POINTER(C.c_double))), f"read {path}")
        return out.reshape(shape, order="F") if shape else out[0].item()

    def keys(self, group: str) -> list[str]:
        need = C.c_int64()
        self._ck(self.lib.pffrg_h5_keys(self.h, group.encode(), None, 0, C.byref(need)), f"keys {group}")
        buf = C.create_string_buffer(int(need.value))
        self._ck(self.lib.pffrg_h5_keys(self.h, group.encode(), buf, len(buf), C.byref(need)), f"keys {group}")
        s = buf.value.decode()
        return s.split("\n") if s else []


_CH = ("ch_s", "ch_t", "ch_u")
_Q = ("q1", "q2_1", "q2_2", "q3")
_COMP = ("spin", "dens")
# U(1)-DM vertex components as checkpoint_u1_dm.jl names them
_COMP_U1 = ("Γxx", "Γzz", "ΓDM", "Γdd", "Γzd", "Γdz")
# χ components of save_χ! (Observable/disk.jl:19-24)
CHI_LABELS = {"su2": ("diag",), "u1-dm": ("xx", "zz", "xy")}


def save_vertex(f: H5File, label: str, a: Action, comp: int):
    """save!(file, label, Γ::Vertex) and save!(file, label, ch::Channel) (Action/disk.jl:2-14, 80-95)."""
    f.write(label + "/bare", a.bare[comp])
    for ch, cn in enumerate(_CH):
        for q in _Q:
            f.write(f"{label}/{cn}/{q}", a.view(comp, ch, q))


def read_vertex(f: H5File, label: str, a: Action, comp: int):
    """read_vertex / read_channel (Action/disk.jl:17-32, 98-115) into component comp of a."""
    a.bare[comp] = f.read(label + "/bare")
    for ch, cn in enumerate(_CH):
        for q in _Q:
            a.view(comp, ch, q)[...] = f.read(f"{label}/{cn}/{q}")


def checkpoint(f: H5File, L: float, dL: float, m: Mesh, a: Action):
    """checkpoint! (checkpoint_su2.jl:2-39; checkpoint_u1_dm.jl:2-37 for the six-component
    action, which stores no spin length)."""
    k = julia_float(L)
    f.write(f"dΛ/{k}", float(dL))
    for name, v in (("σ", m.sigma), ("Ωs", m.Os), ("νs", m.ns), ("Ωt", m.Ot), ("νt", m.nt), ("χ", m.chi)):
        f.write(f"{name}/{k}", v)
    u1 = a.nc == 6
    if not u1 and not f.haskey("S"):
        f.write("S", float(a.S))
    if not f.haskey("symmetry"):
        f.write("symmetry", "u1-dm" if u1 else "su2")
    f.write(f"a/{k}/Σ", a.Sigma)
    names = _COMP_U1 if u1 else _COMP
    for c in range(a.nc):
        save_vertex(f, f"a/{k}/Γ/{names[c]}", a, c)


def read_checkpoint(f: H5File, L: float) -> tuple[float, float, Mesh, Action]:
    """read_checkpoint (Action.jl:589-603) with read_checkpoint_su2 (checkpoint_su2.jl:42-79) or
    read_checkpoint_u1_dm (checkpoint_u1_dm.jl:40-78) by the file's symmetry: the checkpoint at the
    cutoff nearest to L (first of equally near ones in key order, as argmin does)."""
    names = f.keys("σ")
    cut = [float(n) for n in names]
    i = int(np.argmin(np.abs(np.array(cut) - L)))
    k, Lc = names[i], cut[i]
    dL = float(f.read(f"dΛ/{k}"))
    m = Mesh(*(np.asarray(f.read(f"{n}/{k}"), dtype=np.float64) for n in ("σ", "Ωs", "νs", "Ωt", "νt", "χ")))
    u1 = f.haskey("symmetry") and f.read("symmetry") == "u1-dm"
    comps = _COMP_U1 if u1 else _COMP
    S = 0.5 if u1 else float(f.read("S"))
    ns = len(f.read(f"a/{k}/Γ/{comps[0]}/bare"))
    a = Action.empty(ns, m, S, nc=len(comps))
    a.Sigma[:] = f.read(f"a/{k}/Σ")
    for c in range(len(comps)):
        read_vertex(f, f"a/{k}/Γ/{comps[c]}", a, c)
    return Lc, dL, m, a


def save_self(f: H5File, L: float, m: Mesh, Sigma: np.ndarray):
    """save_self! (Action/disk.jl:35-49): the self-energy mesh and Σ of the action."""
    k = julia_float(L)
    f.write(f"σ/{k}", m.sigma)
    f.write(f"Σ/{k}", np.asarray(Sigma, dtype=np.float64))


def save_chi(f: H5File, L: float, symmetry: str, m: Mesh, chi: np.ndarray):
    """save_χ! (Observable/disk.jl:2-29): SU(2) chi[site, ω] as "diag"; U(1)-DM chi[c, site, ω] as
    "xx", "zz", "xy"."""
    if symmetry not in CHI_LABELS:
        raise ValueError(f"unknown symmetry {symmetry}")
    if not f.haskey("symmetry"):
        f.write("symmetry", symmetry)
    k = julia_float(L)
    f.write(f"χ/{k}/mesh", m.chi)
    chi = np.asarray(chi, dtype=np.float64)
    parts = [chi] if symmetry == "su2" else [chi[c] for c in range(3)]
    for label, x in zip(CHI_LABELS[symmetry], parts):
        f.write(f"χ/{k}/{label}", np.ascontiguousarray(x))


def read_chi(f: H5File, L: float, label: str = "diag") -> tuple[float, np.ndarray]:
    """read_χ (Observable/disk.jl:63-87): the nearest stored cutoff and χ[site, ω] of one label
    ("diag" for SU(2); "xx", "zz", "xy" for U(1)-DM)."""
    names = f.keys("χ")
    cut = np.array([float(n) for n in names])
    i = int(np.argmin(np.abs(cut - L)))
    return float(cut[i]), f.read(f"χ/{names[i]}/{label}")


def save_lattice(f: H5File, name: str, size: int, model: str, euclidean: bool, J):
    """save!(file, r::Reduced_lattice, euclidean) (Lattice/disk.jl:2-19)."""
    f.write("reduced_lattice/name", name)
    f.write("reduced_lattice/size", np.int64(size))
    f.write("reduced_lattice/model", model)
    f.write("reduced_lattice/euclidean", bool(euclidean))
    for i, j in enumerate(J, start=1):
        f.write(f"reduced_lattice/J/{i}", np.asarray(j, dtype=np.float64))


def read_lattice(f: H5File) -> tuple[str, int, str, bool, list[np.ndarray]]:
    """The stored description of read_lattice (Lattice/disk.jl:28-48): name, size, model,
    euclidean and the coupling vectors J/1, J/2, ... in index order."""
    n = len(f.keys("reduced_lattice/J"))
    J = [np.atleast_1d(f.read(f"reduced_lattice/J/{i}")) for i in range(1, n + 1)]
    return (f.read("reduced_lattice/name"), int(f.read("reduced_lattice/size")), f.read("reduced_lattice/model"),
            bool(f.read("reduced_lattice/euclidean")), J)


class Measurement:
    """The per-step bookkeeping of measure() (Launcher.jl:2-75) for given observable and
    checkpoint files: Σ and χ of every new cutoff go to the observable file, and a checkpoint is
    written when ct hours have passed since the last one (every cutoff while less than half an
    hour of the wall time wt is left, none in its last 0.1 h).  χ itself is computed by the flow;
    the action is fetched (get_action) only for a checkpoint."""

    def __init__(self, obs_file, cp_file, wt: float = 24.0, ct: float = 1.0, lattice=None):
        """lattice = (name, size, model, euclidean, J): written to both files first, as launch!
        does (Launcher.jl:659-660), unless they already hold one."""
        self.obs_file, self.cp_file, self.wt, self.ct = obs_file, cp_file, wt, ct
        self.t0 = self.t = time.time()
        self.checkpoints: list[float] = []
        if lattice is not None:
            for path in (obs_file, cp_file):
                with H5File(path, "cw") as f:
                    if not f.haskey("reduced_lattice"):
                        save_lattice(f, *lattice)

    def _checkpoint(self, cp: H5File, L, dL, m, get_action):
        if not cp.haskey(f"a/{julia_float(L)}"):
            checkpoint(cp, L, dL, m, get_action())
            self.checkpoints.append(L)

    def __call__(self, L: float, dL: float, m: Mesh, chi: np.ndarray, Sigma: np.ndarray, get_action,
                 force: bool = False):
        """force: the flow's forced checkpoint, measure() called with ct = 0 (at a required
        cutoff, on divergence, on the monotonicity stop; launcher_1l.jl:143-172)."""
        ct = 0.0 if force else self.ct
        with H5File(self.obs_file, "cw") as obs, H5File(self.cp_file, "cw") as cp:
            if not obs.haskey(f"χ/{julia_float(L)}"):
                save_self(obs, L, m, Sigma)
                save_chi(obs, L, "su2" if np.asarray(chi).ndim == 2 else "u1-dm", m, chi)
            h0 = (time.time() - self.t0) / 3600.0
            if self.wt - h0 > 0.5:
                if (time.time() - self.t) / 3600.0 >= ct:
                    self._checkpoint(cp, L, dL, m, get_action)
                    self.t = time.time()
            elif 0.1 < self.wt - h0 <= 0.5:
                self._checkpoint(cp, L, dL, m, get_action)

    def finish(self):
        """The finished flags of both files, set when the flow loop ends (launcher_1l.jl:201-209)."""
        for path in (self.obs_file, self.cp_file):
            with H5File(path, "cw") as f:
                if not f.haskey("finished"):
                    f.write("finished", True)
