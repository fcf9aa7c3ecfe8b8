"""Learned warm start of the parquet iteration (no reference counterpart).

The reference starts launch_parquet! from the bare action (Launcher.jl:676-678, loop input
at parquet.jl:54 and 68).  Here a multilayer perceptron maps (J2/J1, Λ) to coefficients of a
fixed per-channel basis of parquet solutions, and the decoded action starts the same
iteration.  Solutions at different Λ live on meshes proportional to Λ (launch_mesh with
initial = Λ >= 0.5), so array index = scaled frequency and one basis serves every Λ.

The network is trained and evaluated by the library's HIP kernels (pffrg_mlp_*: BF16 MFMA
products with FP32 accumulation, Adam) and decoded by pffrg_decoder (FP32); the basis is the
truncated SVD of the training solutions (numpy, FP64).  Bare initialisation stays the
fallback: WarmStart.initial_action returns the bare action when no model is fitted.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from dataclasses import dataclass

import numpy as np

from .action import Action, splitmix64
from .native import PffrgError, load

_F = C.POINTER(C.c_float)


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


class MLP:
    """n_in -> width (x (n_layers - 1), SiLU) -> n_out on the GPU (pffrg_mlp_*)."""

    def __init__(self, n_in: int, width: int, n_layers: int, n_out: int, seed: int = 1, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.pffrg_mlp_create(n_in, width, n_layers, n_out, seed, device, C.byref(h))
        if rc != 0 or not h:
            raise PffrgError(f"pffrg_mlp_create failed ({rc})")
        self.h = h
        self.n_in, self.n_out = n_in, n_out
        self.n_params = int(self.lib.pffrg_mlp_num_params(h))
        self.step = 0

    def _ck(self, rc, what):
        if rc != 0:
            msg = self.lib.pffrg_mlp_last_error(self.h)
            raise PffrgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pffrg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._ck(self.lib.pffrg_mlp_get_params(self.h, _fp(p)), "mlp_get_params")
        return p

    def set_params(self, p: np.ndarray):
        self._ck(self.lib.pffrg_mlp_set_params(self.h, _fp(np.ascontiguousarray(p, dtype=np.float32))), "mlp_set_params")
        self.step = 0

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self._ck(self.lib.pffrg_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)), "mlp_forward")
        return y

    def backward(self, x: np.ndarray, t: np.ndarray, want_grads=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32) if want_grads else None
        self._ck(self.lib.pffrg_mlp_backward(self.h, _fp(x), _fp(t), x.shape[0], _fp(loss),
                                             _fp(g) if want_grads else None), "mlp_backward")
        return (float(loss[0]), g) if want_grads else float(loss[0])

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        self._ck(self.lib.pffrg_mlp_adam(self.h, lr, b1, b2, eps, self.step), "mlp_adam")


class Decoder:
    """out = mean + coef @ basis (pffrg_decoder, FP32 products, FP64 output)."""

    def __init__(self, mean: np.ndarray, basis: np.ndarray, device: int = 0):
        self.lib = load()
        self.mean = np.ascontiguousarray(mean, dtype=np.float32)
        self.basis = np.ascontiguousarray(basis, dtype=np.float32)
        self.k, self.n = self.basis.shape
        h = C.c_void_p()
        rc = self.lib.pffrg_decoder_create(self.n, self.k, _fp(self.mean), _fp(self.basis), device, C.byref(h))
        if rc != 0 or not h:
            raise PffrgError(f"pffrg_decoder_create failed ({rc})")
        self.h = h

    def __call__(self, coef: np.ndarray) -> np.ndarray:
        out = np.zeros(self.n)
        c = np.ascontiguousarray(coef, dtype=np.float32)
        rc = self.lib.pffrg_decoder_apply(self.h, _fp(c), out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise PffrgError(f"pffrg_decoder_apply failed ({rc})")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.pffrg_decoder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coupling_samples(n: int, seed: int = 5, j2=(-0.5, 0.5), lam=(0.5, 5.0)) -> np.ndarray:
    """n seeded (J2/J1, Λ) pairs, uniform in the given ranges (splitmix64 stream)."""
    s = splitmix64(seed, 2 * n)
    u = (s >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.stack([j2[0] + (j2[1] - j2[0]) * u[0::2], lam[0] + (lam[1] - lam[0]) * u[1::2]], axis=1)


def features(p: np.ndarray) -> np.ndarray:
    """(J2, Λ) -> network input (J2, log Λ)."""
    p = np.atleast_2d(p)
    return np.stack([p[:, 0], np.log(p[:, 1])], axis=1)


def _blocks(a: Action):
    """The compressed blocks of an action: one per channel (both components), then Σ."""
    nch = a.nO + 2 * a.nO * a.nn + a.nO * a.nn * a.nn
    chlen = a.ns * nch
    comp = 3 * chlen
    out = []
    for ch in range(3):
        out.append(np.concatenate([a.gamma[c * comp + ch * chlen: c * comp + (ch + 1) * chlen] for c in range(2)]))
    out.append(a.Sigma.copy())
    return out


def _assemble(template: Action, blocks) -> Action:
    a = template.copy()
    nch = a.nO + 2 * a.nO * a.nn + a.nO * a.nn * a.nn
    chlen = a.ns * nch
    comp = 3 * chlen
    for ch in range(3):
        v = blocks[ch]
        for c in range(2):
            a.gamma[c * comp + ch * chlen: c * comp + (ch + 1) * chlen] = v[c * chlen:(c + 1) * chlen]
    a.Sigma[:] = blocks[3]
    return a


@dataclass
class FitReport:
    losses: list
    n_coef: list
    train_rel_err: float
    train_block_err: list = None  # per block: s, t, u channels, Σ


class WarmStart:
    """Coefficient model of parquet solutions: truncated SVD per block + MLP + decoders."""

    def __init__(self, max_coef: int = 256, width: int = 512, n_layers: int = 4, seed: int = 1, device: int = 0):
        self.max_coef, self.width, self.n_layers, self.seed, self.device = max_coef, width, n_layers, seed, device
        self.mlp = None
        self.decoders = []

    def fit(self, params: np.ndarray, actions: list[Action], steps: int = 3000, lr: float = 1e-3,
            verbose: bool = False) -> FitReport:
        blocks = [_blocks(a) for a in actions]
        n = len(actions)
        self.template = actions[0]
        targets, self.decoders, ncoef = [], [], []
        svds = []
        for b in range(4):
            X = np.stack([bl[b] for bl in blocks])
            mean = X.mean(axis=0)
            U, s, Vt = np.linalg.svd(X - mean, full_matrices=False)
            svds.append((mean, U, s, Vt))
        for mean, U, s, Vt in svds:
            k = int(min(self.max_coef, n, int(np.sum(s > s[0] * 1e-12)) if s[0] > 0 else 1))
            k = max(k, 1)
            # targets c_j = U_j s_j / s_0 sqrt(n) with the block's own largest singular value s_0:
            # the leading coefficient is of order one in every block, and a unit error in a target
            # is a unit relative error of that block's deviation from its mean, so the squared
            # loss weighs the four blocks (three channels and Σ) alike whatever their magnitudes;
            # decoder basis = Vt_k s_0 / sqrt(n)
            sb = float(s[0]) if s[0] > 0 else 1.0
            targets.append(U[:, :k] * (s[:k] / sb) * math.sqrt(n))
            self.decoders.append(Decoder(mean, Vt[:k] * (sb / math.sqrt(n)), self.device))
            ncoef.append(k)
        self.ncoef = ncoef
        T = np.concatenate(targets, axis=1).astype(np.float32)
        F = features(params)
        self.fmean, self.fstd = F.mean(axis=0), F.std(axis=0) + 1e-12
        Xin = ((F - self.fmean) / self.fstd).astype(np.float32)
        self.mlp = MLP(2, self.width, self.n_layers, T.shape[1], self.seed, self.device)
        losses = []
        for it in range(steps):
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))  # cosine decay
            losses.append(self.mlp.backward(Xin, T))
            self.mlp.adam(rate)
            if verbose and (it % 500 == 0 or it == steps - 1):
                print(f"  step {it}: loss {losses[-1]:.3e}", file=sys.stderr, flush=True)
        pred = [self.predict_blocks(p) for p in params]
        errs = self.block_errors(pred, blocks)
        return FitReport(losses, ncoef, max(errs), errs)

    @staticmethod
    def block_errors(pred, blocks) -> list:
        """Worst relative max-norm error over the samples, per block (s, t, u channels, Σ)."""
        return [max(float(np.max(np.abs(pb[b] - bl[b])) / max(np.max(np.abs(bl[b])), 1e-30))
                    for pb, bl in zip(pred, blocks)) for b in range(4)]

    def held_out_errors(self, params, actions: list[Action]) -> list:
        """block_errors of the decoded warm starts against given (converged) solutions."""
        return self.block_errors([self.predict_blocks(p) for p in params], [_blocks(a) for a in actions])

    def predict_blocks(self, p):
        x = ((features(np.asarray(p, dtype=np.float64)) - self.fmean) / self.fstd).astype(np.float32)
        y = self.mlp.forward(x)[0]
        out, o = [], 0
        for b, k in enumerate(self.ncoef):
            out.append(self.decoders[b](y[o:o + k]))
            o += k
        return out

    def initial_action(self, p, bare: Action) -> Action:
        """Decoded warm start for (J2, Λ) with the problem's bare couplings; bare action if unfitted."""
        if self.mlp is None:
            return bare.copy()
        a = _assemble(self.template, self.predict_blocks(p))
        a.bare[:] = bare.bare
        a.S = bare.S
        return a
