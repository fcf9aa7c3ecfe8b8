"""Host-side drivers around the device right-hand side: one RHS evaluation, the
Bogacki-Shampine flow loop (launcher_1l.jl / launcher_2l.jl / launcher_ml.jl) and
the parquet iteration (parquet.jl), all driving the C ABI through native.Context.

Step control, the embedded-error norm and the parquet residual stay in FP64
(pffrg_abs_max reads FP64 device data).  After every accepted step the flow measures χ
and moves the action onto rescanned frequency meshes (resample_from_to, Action.jl:483-557),
as the reference's launchers do.
"""
from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import lattice as LT
from . import mesh as M
from .action import SYMMETRIES, Action
from .native import FP64, FP32_MIXED, Batch, Context, scan

# slot roles
A, DA, A_STAGE, A_INTER, A_ERR, DA_L, DA_C, DA_TEMP, DA_SIGMA = range(9)


@dataclass
class Problem:
    """Lattice + meshes + bare action of a launch! call (Launcher.jl:565-680)."""
    name: str
    size: int
    J: list
    S: float = 0.5
    A_rep: float = 0.0
    mesh: M.Mesh = None
    l: LT.Lattice = None
    r: LT.ReducedLattice = None
    bare: np.ndarray = None
    symmetry: str = "su2"   # "su2" or "u1-dm" (launch!'s symmetry argument)

    @property
    def nc(self) -> int:
        """vertex components of the problem's action (2 or 6)."""
        return SYMMETRIES[self.symmetry]

    @classmethod
    def build(cls, name, size, J, initial, num_sigma=25, num_Omega=15, num_nu=10, num_chi=10, S=0.5, A=0.0,
              normalize=True, model="heisenberg", symmetry=None):
        if symmetry is None:
            symmetry = LT.MODEL_SYMMETRY.get(model, "su2")
        if symmetry != LT.MODEL_SYMMETRY.get(model, symmetry):
            raise ValueError(f"model {model} needs the {LT.MODEL_SYMMETRY[model]} action, not {symmetry}")
        if normalize:
            J, A = LT.normalize_couplings(J, A)
        l, r = LT.build(name, size, J, model)
        m = M.launch_mesh(initial, num_sigma, num_Omega, num_nu, num_chi)
        bare = LT.bare_couplings(l, r, S, A, symmetry)
        return cls(name, size, J, S, A, m, l, r, bare, symmetry)

    def bare_action(self) -> Action:
        a = Action.empty(self.r.n_sites, self.mesh, self.S, nc=self.nc)
        a.bare[:] = self.bare
        return a

    def context(self, n_slots: int = 9, precision: int = FP64, device: int = 0) -> Context:
        c = Context(self.r.n_sites, self.mesh, n_slots, self.S, device, nc=self.nc)
        c.set_lattice(self.r)
        c.set_precision(precision)
        return c


def rhs(ctx: Context, L: float, loops: int, a: int, da: int, eval_: int, Gtol=(1e-8, 1e-4), Stol=(1e-8, 1e-4),
        da_l: int = DA_L, da_c: int = DA_C, da_temp: int = DA_TEMP, da_sigma: int = DA_SIGMA, sigma_corr=True):
    """One complete right-hand side as evaluated per BS3 stage (launcher_1l.jl:85-86,
    launcher_ml.jl:94-95): compute_dΣ! then compute_dΓ_{1l,2l,ml}! (+ compute_dΣ_corr!)."""
    ctx.compute_dsigma(L, a, da, Stol)
    if loops == 1:
        ctx.compute_dgamma_1l(L, a, da, eval_, Gtol)
    elif loops == 2:
        ctx.compute_dgamma_2l(L, a, da, da_l, eval_, Gtol)
    else:
        ctx.compute_dgamma_ml(L, loops, a, da, da_l, da_c, da_temp, da_sigma, eval_, Gtol)
        if sigma_corr:
            ctx.compute_dsigma_corr(L, a, da, da_sigma, Stol)


@dataclass(frozen=True)
class MeshParams:
    """The mesh keywords of launch! (Launcher.jl:580-587)."""
    p_sigma: tuple = M.P_SIGMA
    p_Os: tuple = M.P_OMEGA_S
    p_ns: tuple = M.P_NU_S
    p_Ot: tuple = M.P_OMEGA_T
    p_nt: tuple = M.P_NU_T
    p_chi: tuple = M.P_CHI
    lins: tuple = M.LINS
    bounds: tuple = M.BOUNDS


def chi_components(chi) -> list:
    """The χ matrices [site, ω] of a measurement: one for SU(2) (chi[site, ω]), χxx, χzz, χxy for
    U(1)-DM (chi[c, site, ω]) (get_χ_empty, correlation_lib.jl:37-49)."""
    chi = np.asarray(chi)
    return [chi] if chi.ndim == 2 else [chi[k] for k in range(chi.shape[0])]


def _jl_min(*xs):
    """Julia's min over floats: NaN wins."""
    out = xs[0]
    for x in xs[1:]:
        if x != x or out != out:
            return float("nan")
        out = x if x < out else out
    return out


def rescanned_mesh(ctx, L: float, a: int, chi=None, mp: MeshParams = MeshParams()) -> M.Mesh:
    """The meshes resample_from_to builds at Λ (Action.jl:483-541): linear extents proportional
    to Λ above lins[0], below it from scans of the self-energy of slot a, of the cuts through
    the maxima of its channels (scan_channel, Action.jl:449-480) and of χ[site, :] (the site of
    the largest |χ|).  χ is needed below lins[0]."""
    m_old = ctx.m
    lins, bounds = mp.lins, mp.bounds
    sig_lin, Os_lin, ns_lin = lins[1] * L, lins[2] * L, lins[3] * L
    Ot_lin, nt_lin, chi_lin = lins[2] * L, lins[3] * L, lins[4] * L
    if L < lins[0]:
        sig = ctx.get_sigma(a)
        sig_lin = mp.p_sigma[1] * m_old.sigma[int(np.argmax(np.abs(sig)))]
        _, cuts = ctx.scan_cuts(a)
        nO, nn = m_old.num_Omega, m_old.num_nu

        def scan_channel(k, pO, pn, Om, nu):
            c = cuts[k]
            O_lin = scan(Om, c[:nO], (pO[0], pO[1], pO[2], pO[3] * L, pO[4] * L))
            pv = (pn[0], pn[1], pn[2], pn[3] * L, pn[4] * L)
            v_lin = _jl_min(*[scan(nu, c[nO + j * nn: nO + (j + 1) * nn], pv) for j in range(3)])
            return O_lin, v_lin

        Os_l, ns_l, Ot_l, nt_l = [], [], [], []
        for comp in range(len(cuts) // 3):
            Os_s, ns_s = scan_channel(3 * comp + 0, mp.p_Os, mp.p_ns, m_old.Os, m_old.ns)
            Os_u, ns_u = scan_channel(3 * comp + 2, mp.p_Os, mp.p_ns, m_old.Os, m_old.ns)
            Os_l.append(_jl_min(Os_s, Os_u))
            ns_l.append(_jl_min(ns_s, ns_u))
            Ot_c, nt_c = scan_channel(3 * comp + 1, mp.p_Ot, mp.p_nt, m_old.Ot, m_old.nt)
            Ot_l.append(Ot_c)
            nt_l.append(nt_c)
        Os_lin, ns_lin = _jl_min(*Os_l), _jl_min(*ns_l)
        Ot_lin, nt_lin = _jl_min(*Ot_l), _jl_min(*nt_l)
        if chi is None:
            raise ValueError("rescanning below lins[0] needs the correlations χ")
        pc = mp.p_chi
        chi_lins = []
        for c in chi_components(chi):
            site = int(np.argmax(np.abs(c).T)) % c.shape[0]  # first maximum in column-major order
            chi_lins.append(scan(m_old.chi, c[site, :], (pc[0], pc[1], pc[2], pc[3] * L, pc[4] * L)))
        chi_lin = _jl_min(*chi_lins)
    up = max(L, bounds[0])
    return M.Mesh(
        M.get_mesh(sig_lin, bounds[1] * up, m_old.num_sigma - 1, mp.p_sigma[0]) * mp.p_sigma[2],
        M.get_mesh(Os_lin, bounds[2] * up, m_old.num_Omega - 1, mp.p_Os[0]) * mp.p_Os[5],
        M.get_mesh(ns_lin, bounds[3] * up, m_old.num_nu - 1, mp.p_ns[0]) * mp.p_ns[5],
        M.get_mesh(Ot_lin, bounds[2] * up, m_old.num_Omega - 1, mp.p_Ot[0]) * mp.p_Ot[5],
        M.get_mesh(nt_lin, bounds[3] * up, m_old.num_nu - 1, mp.p_nt[0]) * mp.p_nt[5],
        M.get_mesh(chi_lin, bounds[4] * up, m_old.num_chi - 1, mp.p_chi[0]) * mp.p_chi[5])


def resample_from_to(ctx, L: float, src: int, dst: int, chi=None, mp: MeshParams = MeshParams()) -> M.Mesh:
    """resample_from_to (Action.jl:483-557): new meshes at Λ (rescanned_mesh), then slot src
    interpolated onto them into slot dst; the context continues on the new meshes."""
    m_new = rescanned_mesh(ctx, L, src, chi, mp)
    ctx.resample(src, dst, m_new)
    return m_new


@dataclass
class FlowResult:
    Lambdas: list = field(default_factory=list)
    gamma_max: list = field(default_factory=list)
    accepted: int = 0
    rejected: int = 0
    rhs_evals: int = 0
    wall_s: float = 0.0
    status: str = ""
    chi: list = field(default_factory=list)   # χ[site, ω] after every accepted step (if measured)
    chi_s: float = 0.0                        # wall time spent on χ
    meshes: list = field(default_factory=list)  # meshes after every rescanned step


class BS3Flow:
    """One Bogacki-Shampine flow from Li to Lf (launcher_1l.jl:72-198 and its 2l/ml twins) kept
    as a state machine, so that a caller can evaluate the right-hand sides of many flows
    together.  A step has four stages; before(k) prepares the input of stage k, the caller
    evaluates that right-hand side (stage_input(k) -> slot A_STAGE), after(k) folds it into the
    update and the error estimate, and finish_step() accepts or rejects the step."""

    def __init__(self, ctx: Context, prob: Problem, a0: Action, Li: float, Lf: float, loops: int = 1, sigma_corr=True,
                 bmin=1e-4, bmax=0.1, min_eval=10, max_eval=100, Stol=(1e-8, 1e-4), Gtol=(1e-8, 1e-4),
                 ODE_tol=(1e-8, 1e-2), max_steps: int | None = None, verbose=False, chi_tol=(1e-8, 1e-4),
                 rescan: bool = True, mesh_params: MeshParams = MeshParams(), max_seconds: float | None = None,
                 measurement=None, cps=()):
        self.ctx, self.prob, self.Lf, self.loops, self.sigma_corr = ctx, prob, Lf, loops, sigma_corr
        # measurement(Λ, dΛ, mesh, χ, Σ, get_action) after every accepted step with χ, e.g. an
        # h5io.Measurement writing the observable and checkpoint files (measure, Launcher.jl:2-75)
        self.measurement = measurement
        self.max_seconds = max_seconds
        if rescan and chi_tol is None:
            raise ValueError("mesh rescanning needs the χ measurement (chi_tol)")
        self.chi_tol, self.rescan, self.mesh_params = chi_tol, rescan, mesh_params
        self.bmin, self.bmax, self.min_eval, self.max_eval = bmin, bmax, min_eval, max_eval
        self.Stol, self.Gtol, self.ODE_tol, self.max_steps, self.verbose = Stol, Gtol, ODE_tol, max_steps, verbose
        self.res = FlowResult()
        self.t0 = time.time()
        if ctx.m is not prob.mesh:
            ctx.set_mesh(prob.mesh)  # a0 lives on the problem's initial meshes
        ctx.upload(A, a0)
        # da, a_stage, a_err start empty; a_inter carries the bare couplings (launcher_1l.jl:39-44)
        zero = Action.empty(prob.r.n_sites, prob.mesh, prob.S, nc=prob.nc)
        for s in (DA, A_STAGE, A_ERR, DA_L, DA_C, DA_TEMP, DA_SIGMA):
            ctx.upload(s, zero)
        ctx.upload(A_INTER, prob.bare_action())
        self.L, self.dL = Li, bmax * Li
        # required checkpoints: the user's cutoffs with Λi and Λf (launcher_1l.jl:59-69)
        self.cps = sorted({float(c) for c in cps} | {Li, Lf}, reverse=True)
        for c in self.cps:
            if c < self.L:
                self.dL = min(self.dL, 0.85 * (self.L - c))
                break
        self.steps = 0
        self.done = not (self.L > Lf)
        if self.done:
            self.res.status = "finished"
        self.ev = 0

    def begin_step(self) -> bool:
        """False when the flow has ended (Λ_f reached, step limit, divergence)."""
        if self.done:
            return False
        if self.max_steps is not None and self.steps >= self.max_steps:
            self.res.status = "max_steps"
            self.done = True
            return False
        if self.max_seconds is not None and time.time() - self.t0 > self.max_seconds:
            self.res.status = "time_budget"
            self.done = True
            return False
        self.steps += 1
        self.ev = M.get_eval(self.L, self.min_eval, self.max_eval)
        return True

    # The stage algebra (launcher_1l.jl:84-116) as lists of slot operations, ("copy", dst, src)
    # for replace_with! and ("axpy", src, fac, dst) for mult_with_add_to!: every flow has the same
    # operations on the same slots at a stage, only the factors differ, so that a batch of flows
    # can run each operation for all of them in one call.
    def before_ops(self, k: int) -> list:
        dL = self.dL
        if k == 0:
            return [("copy", DA, A), ("copy", A_ERR, A)]
        if k in (1, 2):
            return [("copy", A_INTER, A), ("axpy", A_STAGE, (-0.5 if k == 1 else -0.75) * dL, A_INTER)]
        return [("copy", A_INTER, DA)]

    def after_ops(self, k: int) -> list:
        dL = self.dL
        if k == 0:
            return [("axpy", A_STAGE, -2.0 * dL / 9.0, DA), ("axpy", A_STAGE, -7.0 * dL / 24.0, A_ERR)]
        if k == 1:
            return [("axpy", A_STAGE, -1.0 * dL / 3.0, DA), ("axpy", A_STAGE, -1.0 * dL / 4.0, A_ERR)]
        if k == 2:
            return [("axpy", A_STAGE, -4.0 * dL / 9.0, DA), ("axpy", A_STAGE, -1.0 * dL / 3.0, A_ERR)]
        return [("axpy", A_STAGE, -1.0 * dL / 8.0, A_ERR)]

    @staticmethod
    def finish_ops() -> list:
        # the embedded error a_err -= a_inter (launcher_1l.jl:113-116)
        return [("axpy", A_INTER, -1.0, A_ERR)]

    def run_ops(self, ops):
        for op in ops:
            if op[0] == "copy":
                self.ctx.replace_with(op[1], op[2])
            else:
                self.ctx.mult_with_add_to(op[1], op[2], op[3])

    def before(self, k: int):
        self.run_ops(self.before_ops(k))

    def stage_input(self, k: int):
        """(Λ, input slot, eval) of stage k; its right-hand side goes to slot A_STAGE."""
        L, dL = self.L, self.dL
        return ((L, A), (L - 0.5 * dL, A_INTER), (L - 0.75 * dL, A_INTER), (L - dL, A_INTER))[k] + (self.ev,)

    def after(self, k: int):
        self.run_ops(self.after_ops(k))

    def rhs(self, k: int):
        L, a, ev = self.stage_input(k)
        rhs(self.ctx, L, self.loops, a, A_STAGE, ev, self.Gtol, self.Stol, sigma_corr=self.sigma_corr)

    def finish_step(self):
        """Accept or reject the step (launcher_1l.jl:113-198): error estimate, step size, χ of an
        accepted step, the stops, new meshes."""
        if self.finish_step_a():
            self.finish_step_b(self.measure_chi())
        else:
            self.finish_step_b(None)

    def finish_step_a(self, maxima=None) -> bool:
        """First half of finish_step: the embedded error, acceptance and the next step size.
        True when an accepted step now needs χ (measure_chi, or a batched compute_chi).  A
        batch passes maxima = (|A_ERR|, |A_INTER|, |A|) after running finish_ops() itself."""
        ctx, res = self.ctx, self.res
        res.rhs_evals += 4
        self._accepted = False
        # embedded error (launcher_1l.jl:113-116)
        if maxima is None:
            self.run_ops(self.finish_ops())
            maxima = (ctx.abs_max(A_ERR), ctx.abs_max(A_INTER), ctx.abs_max(A))
        delta, gmax, amax = maxima
        scale = self.ODE_tol[0] + max(gmax, amax) * self.ODE_tol[1]
        err = delta / scale
        if self.verbose:
            print(f"L={self.L:.6g} dL={self.dL:.3g} eval={self.ev} err={err:.3g} Gmax={gmax:.4g} "
                  f"t={time.time() - self.t0:.1f}s", file=sys.stderr, flush=True)
        if err >= 10.0:
            res.status = "unstable"
            self.done = True
            self._finish_files()
            return False
        accept = err <= 1.0 or self.dL <= self.bmin
        if accept:
            self.L -= self.dL
        # err = 0 gives an unbounded proposal (Julia's 1/0 = Inf), so the step is bmax Λ
        grow = 0.85 * (1.0 / err) ** (1.0 / 3.0) * self.dL if err > 0.0 else math.inf
        self.dL = max(self.bmin, min(self.bmax * self.L, grow))
        lower = [c - self.L for c in self.cps if c - self.L < 0.0]
        L_cp = self.L + lower[0] if lower else 0.0
        self.dL = min(self.dL, self.L - L_cp)
        if accept:
            res.accepted += 1
            res.Lambdas.append(self.L)
            res.gamma_max.append(gmax)
            self._accepted = True
            self._gmax = gmax
        else:
            res.rejected += 1
        return accept and self.chi_tol is not None

    def finish_step_b(self, chi):
        """Second half of finish_step, given χ of the accepted step (None when not measured):
        divergence and monotonicity stops (launcher_1l.jl:140-172), new meshes, the end."""
        ctx, res = self.ctx, self.res
        if self.done:
            return
        if self._accepted:
            # divergence, measurement, monotonicity, then new meshes (launcher_1l.jl:140-178); a
            # checkpoint is forced at a required cutoff (Λ ≈ one of cps) and at either stop
            stop = "diverged" if self._gmax >= 1000.0 else None
            meas = (lambda force: self.measurement(self.L, self.dL, ctx.m, chi, ctx.get_sigma(A_INTER),
                                                   lambda: ctx.download(A_INTER), force=force)) \
                if (chi is not None and self.measurement is not None) else (lambda force: None)
            meas(stop is not None or any(_julia_isapprox(self.L, c) for c in self.cps))
            if chi is not None and self.record_chi(chi) and stop is None:
                stop = "non_monotone"
                meas(True)
            if stop is not None:
                ctx.replace_with(A, A_INTER)
                res.status = stop
                self.done = True
                self._finish_files()
                return
            if self.rescan:
                resample_from_to(ctx, self.L, A_INTER, A, res.chi[-1], self.mesh_params)
                res.meshes.append(ctx.m)
            else:
                ctx.replace_with(A, A_INTER)
        if not (self.L > self.Lf):
            res.status = "finished"
            self.done = True
            self._finish_files()

    def _finish_files(self):
        """The end of the flow loop (Λ_f reached, instability, divergence or the monotonicity
        stop): the finished flags of the observable and checkpoint files (launcher_1l.jl:201-209).
        A flow cut short by max_steps or max_seconds is not finished."""
        if self.measurement is not None and hasattr(self.measurement, "finish"):
            self.measurement.finish()

    def measure_chi(self):
        """compute_χ! of the accepted action (slot A_INTER) on the current χ mesh (measure,
        Launcher.jl:2-40)."""
        t = time.time()
        chi = self.ctx.compute_chi(self.L, A_INTER, self.chi_tol)
        self.res.chi_s += time.time() - t
        return chi

    def record_chi(self, chi) -> bool:
        """Store χ of the step; True when the static on-site correlation of the component with
        the largest |χ| stopped growing (χ/χ_prev < 0.995, the reference's stop rule,
        Launcher.jl:33-35)."""
        prev = self.res.chi[-1] if self.res.chi else None
        self.res.chi.append(chi)
        cs = chi_components(chi)
        k = int(np.argmax([np.max(np.abs(c)) for c in cs]))
        if prev is None or chi_components(prev)[k][0, 0] == 0.0:
            return False
        return not (cs[k][0, 0] / chi_components(prev)[k][0, 0] >= 0.995)

    def result(self) -> FlowResult:
        self.ctx.synchronize()
        self.res.wall_s = time.time() - self.t0
        return self.res


def _julia_isapprox(x: float, y: float) -> bool:
    """Julia's x ≈ y for Float64: |x - y| <= sqrt(eps) max(|x|, |y|)."""
    return x == y or abs(x - y) <= 1.4901161193847656e-08 * max(abs(x), abs(y))


def launch_flow(ctx: Context, prob: Problem, a0: Action, Li: float, Lf: float, loops: int = 1, sigma_corr=True,
                bmin=1e-4, bmax=0.1, min_eval=10, max_eval=100, Stol=(1e-8, 1e-4), Gtol=(1e-8, 1e-4),
                ODE_tol=(1e-8, 1e-2), max_steps: int | None = None, verbose=False, chi_tol=(1e-8, 1e-4),
                rescan: bool = True, mesh_params: MeshParams = MeshParams(),
                max_seconds: float | None = None, measurement=None, cps=()) -> FlowResult:
    """BS3 flow from Li to Lf (launcher_1l.jl:72-198 and its 2l/ml twins).  Slot A holds the
    flowing action on return, on the context's current meshes.  After every accepted step χ is
    measured (chi_tol; None skips it), the flow stops when the static on-site correlation turns
    (launcher_1l.jl:160-172), and the action moves onto rescanned meshes (rescan).  A
    measurement (h5io.Measurement) writes the observable and checkpoint files as launch! does."""
    f = BS3Flow(ctx, prob, a0, Li, Lf, loops, sigma_corr, bmin, bmax, min_eval, max_eval, Stol, Gtol, ODE_tol,
                max_steps, verbose, chi_tol, rescan, mesh_params, max_seconds, measurement, cps)
    while f.begin_step():
        for k in range(4):
            f.before(k)
            f.rhs(k)
            f.after(k)
        f.finish_step()
    return f.result()


def _batch_ops(batch, per_flow):
    """Run the same list of slot operations for every flow of a batch, one batched call per
    operation, with each flow's own factor."""
    for col in zip(*per_flow):
        op = col[0]
        if any(c[0] != op[0] or c[1] != op[1] or c[-1] != op[-1] for c in col):
            raise ValueError("the flows of a batch must run the same slot operations")
        if op[0] == "copy":
            batch.replace_with(op[1], op[2])
        else:
            batch.mult_with_add_to(op[1], [c[2] for c in col], op[3])


def launch_flow_batch(ctxs, probs, a0s, Li, Lf, max_steps: int | None = None, progress: int = 0,
                      **kw) -> list[FlowResult]:
    """1-loop BS3 flows of several problems advanced together: every stage evaluates the
    right-hand sides of all flows still running in one batched launch (pffrg_batch_rhs_1l).
    Each flow keeps its own Λ, step size and acceptance, exactly as launch_flow would.  With
    progress = k > 0, a line goes to stderr every k batched steps."""
    n = len(ctxs)
    Lis = Li if np.ndim(Li) else [Li] * n
    Lfs = Lf if np.ndim(Lf) else [Lf] * n
    flows = [BS3Flow(ctxs[k], probs[k], a0s[k], Lis[k], Lfs[k], 1, max_steps=max_steps, **kw) for k in range(n)]
    batch, members = None, None
    step = 0
    t0 = time.time()
    try:
        while True:
            live = [k for k in range(n) if flows[k].begin_step()]
            if not live:
                break
            step += 1
            if progress and step % progress == 0:
                Ls = [flows[k].L for k in live]
                print(f"batched step {step}: {len(live)} flows running, Lambda {min(Ls):.4g}..{max(Ls):.4g}, "
                      f"{time.time() - t0:.1f} s", file=sys.stderr, flush=True)
            if members != live:
                if batch is not None:
                    batch.close()
                batch, members = Batch([ctxs[k] for k in live]), live
            for k in range(4):
                _batch_ops(batch, [flows[q].before_ops(k) for q in live])
                ins = [flows[q].stage_input(k) for q in live]
                batch.rhs_1l([x[0] for x in ins], [x[1] for x in ins], A_STAGE, [x[2] for x in ins],
                             flows[live[0]].Stol, flows[live[0]].Gtol)
                _batch_ops(batch, [flows[q].after_ops(k) for q in live])
            _batch_ops(batch, [flows[q].finish_ops() for q in live])
            mx = [batch.abs_max(s) for s in (A_ERR, A_INTER, A)]
            need = [flows[q].finish_step_a((mx[0][i], mx[1][i], mx[2][i])) for i, q in enumerate(live)]
            chis = [None] * len(live)
            if any(need):
                t = time.time()
                chis = batch.compute_chi([flows[q].L for q in live], A_INTER, flows[live[0]].chi_tol, active=need)
                dt = time.time() - t
                for q, nd in zip(live, need):
                    if nd:
                        flows[q].res.chi_s += dt / sum(need)
            for q, chi in zip(live, chis):
                flows[q].finish_step_b(chi)
    finally:
        if batch is not None:
            batch.close()
    return [f.result() for f in flows]


@dataclass
class ParquetResult:
    iterations: int = 0
    sde_iterations: int = 0
    abs_err: float = float("inf")
    rel_err: float = float("inf")
    history: list = field(default_factory=list)
    status: str = ""


def launch_parquet(ctx: Context, L: float, a0: Action, max_iter=50, min_eval=10, max_eval=100, Stol=(1e-8, 1e-4),
                   Gtol=(1e-8, 1e-4), parquet_tol=(1e-8, 1e-4), slot_a=A, slot_ap=DA, slot_err=A_ERR,
                   max_seconds: float | None = None, verbose=False) -> ParquetResult:
    """Damped parquet fixed point (parquet.jl:1-100); slot_a holds the solution on return.
    Besides the reference's stops (tolerance, max_iter), the iteration ends when the vertex
    maximum reaches 1000 (the flow's divergence bound, launcher_1l.jl:145) and, optionally,
    after max_seconds of wall time (checked between iterations)."""
    t0 = time.time()
    res = ParquetResult()
    ctx.upload(slot_a, a0)
    zero = Action.empty(a0.ns, ctx.m, a0.S, nc=a0.nc)
    ctx.upload(slot_ap, zero)
    ctx.upload(slot_err, zero)
    beta = min(L, 1.0)
    ev = M.get_eval(L, min_eval, max_eval)
    abs_err = rel_err = np.inf
    count = 1
    while abs_err >= parquet_tol[0] and rel_err >= parquet_tol[1] and count <= max_iter:
        s_abs = s_rel = np.inf
        s_count = 1
        while s_abs >= parquet_tol[0] and s_rel >= parquet_tol[1] and s_count <= max_iter:
            ctx.compute_sigma(L, slot_a, slot_ap, ev, Gtol, Stol)
            sa, sp = ctx.get_sigma(slot_a), ctx.get_sigma(slot_ap)
            s_abs = float(np.max(np.abs(sp - sa)))
            s_rel = s_abs / max(float(np.max(np.abs(sa))), float(np.max(np.abs(sp))))
            ctx.set_sigma(slot_a, (1.0 - beta) * sa + beta * sp)
            s_count += 1
            res.sde_iterations += 1
            if verbose and s_count % 10 == 0:
                print(f"      SDE iteration {s_count - 1}: abs err {s_abs:.2e} ({time.time() - t0:.1f} s)", file=sys.stderr,
                      flush=True)
            if max_seconds is not None and time.time() - t0 > max_seconds:
                break
        ctx.compute_gamma(L, slot_a, slot_ap, ev, Gtol)
        ctx.replace_with(slot_err, slot_ap, vertex_only=True)
        ctx.mult_with_add_to(slot_a, -1.0, slot_err, vertex_only=True)
        abs_err = ctx.abs_max(slot_err)
        rel_err = abs_err / max(ctx.abs_max(slot_a), ctx.abs_max(slot_ap))
        res.history.append((abs_err, rel_err))
        ctx.mult_with(slot_a, 1.0 - beta, vertex_only=True)
        ctx.mult_with_add_to(slot_ap, beta, slot_a, vertex_only=True)
        if verbose:
            print(f"    parquet iteration {count}: abs err {abs_err:.2e} rel err {rel_err:.2e} "
                  f"(SDE total {res.sde_iterations}, {time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
        count += 1
        if ctx.abs_max(slot_a) >= 1000.0:
            res.status = "diverged"
            break
        if max_seconds is not None and time.time() - t0 > max_seconds:
            res.status = "time"
            break
    else:
        res.status = "converged" if (abs_err < parquet_tol[0] or rel_err < parquet_tol[1]) else "max_iter"
    res.iterations = count - 1
    res.abs_err, res.rel_err = abs_err, rel_err
    return res
