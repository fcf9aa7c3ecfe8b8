"""Error behaviour of the C ABI (include/pffrg.h): every argument and state check returns a
negative status with a message (pffrg_last_error) instead of launching anything, in the places
where the reference's functions stop at an @assert or a Julia method error (quadrature.jl:94,
125-126; compute.jl:57-286 take distinct actions).  After the failed calls the context still
computes exactly what a fresh one does."""
import numpy as np
import pytest

from pffrg import action as A
from pffrg import native
from pffrg.native import PffrgError

from helpers import small_problem

pytestmark = pytest.mark.gpu


def _rhs(ctx, a):
    ctx.upload(0, a)
    ctx.compute_dsigma(1.0, 0, 1)
    ctx.compute_dgamma_1l(1.0, 0, 1, 4)
    return ctx.download(1)


def test_argument_and_state_errors_leave_the_context_usable():
    prob = small_problem("square", 2, 6)
    a = A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=11)
    with prob.context(n_slots=3) as fresh:
        ref = _rhs(fresh, a)
    with prob.context(n_slots=3) as ctx:
        ctx.upload(0, a)
        cases = [
            ("slot index out of range", lambda: ctx.upload(3, a)),
            ("slot index out of range", lambda: ctx.download(-1)),
            ("eval must be >= 1", lambda: ctx.compute_dgamma_1l(1.0, 0, 1, 0)),
            ("invalid", lambda: ctx.compute_dgamma_1l(0.0, 0, 1, 4)),
            ("invalid", lambda: ctx.compute_dgamma_1l(-1.0, 0, 1, 4)),
            ("invalid", lambda: ctx.compute_dgamma_1l(1.0, 0, 0, 4)),
            ("invalid", lambda: ctx.compute_dgamma_2l(1.0, 0, 1, 1, 4)),
            ("distinct", lambda: ctx.compute_dgamma_ml(1.0, 3, 0, 1, 2, 2, 1, 0, 4)),
            ("unknown precision", lambda: ctx.set_precision(7)),
        ]
        for msg, call in cases:
            with pytest.raises(PffrgError, match=msg):
                call()
        # a rejected mesh or lattice keeps the previous one
        m, r = prob.mesh, prob.r
        Os = np.array(m.Os, dtype=np.float64)
        Os[3] = Os[2]
        rc = ctx.lib.pffrg_set_mesh(ctx.h, *[native._dp(np.ascontiguousarray(x, dtype=np.float64))
                                             for x in (m.sigma, Os, m.ns, m.Ot, m.nt)])
        assert rc == -1 and b"strictly increasing" in ctx.lib.pffrg_last_error(ctx.h)
        off, rows = r.overlap_csr()
        off = np.ascontiguousarray(off, dtype=np.int64)
        rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1)
        mult = np.ascontiguousarray(np.rint(r.mult), dtype=np.int64)
        ex = np.array(r.exchange, dtype=np.int64)
        ex[0] = r.n_sites
        rc = ctx.lib.pffrg_set_lattice(ctx.h, native._ip(ex), native._ip(mult), native._ip(off), native._ip(rows), 0)
        assert rc == -1 and b"exchange index out of range" in ctx.lib.pffrg_last_error(ctx.h)
        bad_rows = rows.copy()
        bad_rows[0] = r.n_sites
        ex_ok = np.ascontiguousarray(r.exchange, dtype=np.int64)
        rc = ctx.lib.pffrg_set_lattice(ctx.h, native._ip(ex_ok), native._ip(mult), native._ip(off),
                                       native._ip(bad_rows), 0)
        assert rc == -1 and b"overlap row out of range" in ctx.lib.pffrg_last_error(ctx.h)
        bad_mult = mult.copy()
        bad_mult[0] = 0
        rc = ctx.lib.pffrg_set_lattice(ctx.h, native._ip(ex_ok), native._ip(bad_mult), native._ip(off),
                                       native._ip(rows), 0)
        assert rc == -1 and b"multiplicity out of range" in ctx.lib.pffrg_last_error(ctx.h)
        got = _rhs(ctx, a)
        print(f"{len(cases) + 4} rejected calls; RHS afterwards equals a fresh context's bitwise", flush=True)
        np.testing.assert_array_equal(got.gamma, ref.gamma)
        np.testing.assert_array_equal(got.Sigma, ref.Sigma)


def test_compute_before_lattice_is_a_state_error():
    prob = small_problem("square", 2, 6)
    with native.Context(prob.r.n_sites, prob.mesh, 2) as ctx:  # meshes set, lattice not
        with pytest.raises(PffrgError, match="lattice not set"):
            ctx.compute_dsigma(1.0, 0, 1)


def test_batch_rejects_bad_problem_arguments():
    # pffrg_batch_rhs_1l checks every problem before it launches anything
    prob = small_problem("square", 2, 6)
    a = A.synthetic_action(prob.r.n_sites, prob.mesh, prob.bare, seed=12)
    ctxs = [prob.context(n_slots=2), prob.context(n_slots=2)]
    try:
        for c in ctxs:
            c.upload(0, a)
        with native.Batch(ctxs) as b:
            with pytest.raises(PffrgError, match="invalid Lambda or eval of problem 1"):
                b.rhs_1l([1.0, 1.0], 0, 1, [4, 0])
            with pytest.raises(PffrgError, match="invalid slot of problem 0"):
                b.rhs_1l([1.0, 1.0], 0, 2, [4, 4])
            b.rhs_1l([1.0, 1.0], 0, 1, [4, 4])
        with prob.context(n_slots=3) as fresh:
            ref = _rhs(fresh, a)
        for c in ctxs:
            np.testing.assert_array_equal(c.download(1).gamma, ref.gamma)
    finally:
        for c in ctxs:
            c.close()
