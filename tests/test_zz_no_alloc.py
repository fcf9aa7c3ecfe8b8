"""The allocation-free scan_cuts entry (pffrg_scan_cuts_noalloc) against pffrg_scan_cuts.

Both run the same kernels (scan_channel's argmax and cuts, Action.jl:449-480); the no-alloc entry
keeps its partials and cuts in one static scratch per GPU.  Two calls in a row through the C ABI,
the second on a smaller problem (fewer sites and mesh points, so the scratch still holds the
first call's larger contents), must each equal pffrg_scan_cuts bitwise: the outputs are positions
and plain copies or single subtractions of vertex entries."""
import ctypes as C

import numpy as np
import pytest

from pffrg import native
from pffrg.action import Action

from helpers import small_problem

pytestmark = pytest.mark.gpu


def _random_action(prob, seed):
    rng = np.random.default_rng(seed)
    a = Action.empty(prob.r.n_sites, prob.mesh, prob.S)
    a.gamma[:] = rng.normal(size=a.gamma.shape)
    a.bare[:] = prob.bare
    return a


def _noalloc(ctx, a):
    idx = np.full(24, -7, dtype=np.int32)
    cuts = np.full(6 * (ctx.m.num_Omega + 3 * ctx.m.num_nu), np.nan)
    rc = ctx.lib.pffrg_scan_cuts_noalloc(ctx.h, a, idx.ctypes.data_as(C.POINTER(C.c_int32)), native._dp(cuts))
    assert rc == 0, ctx.lib.pffrg_last_error(ctx.h)
    return idx.reshape(6, 4), cuts.reshape(6, -1)


def test_scan_cuts_noalloc_equals_scan_cuts():
    # larger problem first, then a smaller one on the same GPU's scratch
    for k, (name, size, N) in enumerate([("square", 4, 12), ("square", 2, 6)]):
        prob = small_problem(name, size, N, 1.0)
        a = _random_action(prob, 30 + k)
        print(f"scan_cuts_noalloc call {k + 1}: {name} L={size} N={N} sites {prob.r.n_sites}", flush=True)
        with prob.context(n_slots=2) as ctx:
            ctx.upload(1, a)
            got_idx, got_cuts = _noalloc(ctx, 1)
            ref_idx, ref_cuts = ctx.scan_cuts(1)
        assert np.array_equal(got_idx, ref_idx)
        assert np.array_equal(got_cuts, ref_cuts)
        assert np.all(np.isfinite(got_cuts))
