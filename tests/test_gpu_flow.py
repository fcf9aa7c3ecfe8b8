"""Host flow drivers on the GPU: a BS3 flow (launch_flow, launcher_1l.jl:72-198) with χ
measurement and mesh rescanning after every accepted step against the same driver run on the
FP64 oracle, and batched flows (launch_flow_batch) against single ones.

Tolerance: the GPU right-hand sides agree with the oracle's to ~1e-15 and the stage algebra
differs only by FMA rounding, so the accepted Λ sequence is identical and the final actions
agree to <= 1e-10 relative (max-norm)."""
import numpy as np
import pytest

from pffrg import action as A
from pffrg import flow as F
from pffrg import native

from helpers import rel_err, small_problem
from oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu


def _flow_setup(name, size, N, J=((1.0,),)):
    prob = small_problem(name, size, N, 5.0, J)
    return prob, prob.bare_action()


def test_flow_matches_oracle_flow():
    prob, a0 = _flow_setup("square", 2, 6)
    with prob.context(n_slots=9) as ctx:
        res = F.launch_flow(ctx, prob, a0, 5.0, 1.0, max_steps=6, chi_tol=(1e-8, 1e-4))
        got = ctx.download(F.A)
    octx = OracleContext(prob)
    ref_res = F.launch_flow(octx, prob, a0, 5.0, 1.0, max_steps=6, chi_tol=(1e-8, 1e-4))
    ref = octx.download(F.A)
    print(f"GPU Λ {res.Lambdas} ({res.accepted}/{res.rejected}); oracle Λ {ref_res.Lambdas}", flush=True)
    assert res.accepted == ref_res.accepted and res.rejected == ref_res.rejected
    np.testing.assert_allclose(res.Lambdas, ref_res.Lambdas, rtol=1e-12)
    e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
    ec = max(rel_err(c, r) for c, r in zip(res.chi, ref_res.chi))
    print(f"final action rel err {e:.3e}; chi along the flow {ec:.3e}", flush=True)
    assert e < 1e-10 and ec < 1e-10 and len(res.chi) == res.accepted
    # rescanned meshes after every accepted step (resample_from_to, Action.jl:483-557)
    assert len(res.meshes) == len(ref_res.meshes) == res.accepted
    for m, rm in zip(res.meshes, ref_res.meshes):
        for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
            np.testing.assert_allclose(getattr(m, f), getattr(rm, f), rtol=1e-12)
    assert not np.array_equal(res.meshes[-1].Os, prob.mesh.Os)


def test_multiloop_flow_matches_oracle_flow():
    # 3-loop flow with self-energy corrections (launcher_ml.jl:78-198) for a few steps
    prob, a0 = _flow_setup("square", 2, 6)
    with prob.context(n_slots=9) as ctx:
        res = F.launch_flow(ctx, prob, a0, 5.0, 1.0, loops=3, max_steps=3)
        got = ctx.download(F.A)
    octx = OracleContext(prob)
    ref_res = F.launch_flow(octx, prob, a0, 5.0, 1.0, loops=3, max_steps=3)
    ref = octx.download(F.A)
    print(f"3-loop: GPU Λ {res.Lambdas}; oracle Λ {ref_res.Lambdas}", flush=True)
    assert res.accepted == ref_res.accepted and res.rejected == ref_res.rejected
    np.testing.assert_allclose(res.Lambdas, ref_res.Lambdas, rtol=1e-12)
    e = max(rel_err(got.gamma, ref.gamma), rel_err(got.Sigma, ref.Sigma))
    print(f"3-loop final action rel err {e:.3e}", flush=True)
    assert e < 1e-10


def test_fp32_mixed_flow_chi_within_tolerance():
    # the headline tolerance: χ of the FP32-mixed flow against the FP64 flow, rel err <= 1e-4
    prob, a0 = _flow_setup("square", 3, 8)
    runs = {}
    for mode in (native.FP64, native.FP32_MIXED):
        with prob.context(n_slots=9, precision=mode) as ctx:
            runs[mode] = F.launch_flow(ctx, prob, a0, 5.0, 0.5, max_steps=12)
    r64, r32 = runs[native.FP64], runs[native.FP32_MIXED]
    print(f"FP64 Λ {r64.Lambdas}\nFP32-mixed Λ {r32.Lambdas}", flush=True)
    common = [(i, j) for i, L in enumerate(r64.Lambdas) for j, L2 in enumerate(r32.Lambdas)
              if abs(L - L2) <= 1e-12 * L]
    assert len(common) >= min(len(r64.Lambdas), len(r32.Lambdas)) - 1 >= 3
    err = max(rel_err(r32.chi[j], r64.chi[i]) for i, j in common)
    print(f"χ FP32-mixed vs FP64 over {len(common)} steps: {err:.2e}", flush=True)
    assert err <= 1e-4


def test_batched_flows_equal_single_flows():
    setups = [_flow_setup("square", 2, 6), _flow_setup("kagome", 3, 6), _flow_setup("square", 3, 6, ((1.0,), (0.4,)))]
    single = []
    for prob, a0 in setups:
        with prob.context(n_slots=9) as ctx:
            res = F.launch_flow(ctx, prob, a0, 5.0, 2.0, max_steps=4)
            single.append((res, ctx.download(F.A)))
    ctxs = [prob.context(n_slots=9) for prob, _ in setups]
    try:
        results = F.launch_flow_batch(ctxs, [p for p, _ in setups], [a for _, a in setups], 5.0, 2.0, max_steps=4)
        for k, ctx in enumerate(ctxs):
            got = ctx.download(F.A)
            print(f"problem {k}: Λ {results[k].Lambdas} status {results[k].status}", flush=True)
            assert results[k].Lambdas == single[k][0].Lambdas
            assert results[k].accepted == single[k][0].accepted and results[k].rejected == single[k][0].rejected
            np.testing.assert_array_equal(got.gamma, single[k][1].gamma)
            # χ measured in one batched call per step equals the single flow's
            assert len(results[k].chi) == len(single[k][0].chi)
            for x, y in zip(results[k].chi, single[k][0].chi):
                np.testing.assert_array_equal(x, y)
    finally:
        for ctx in ctxs:
            ctx.close()


@pytest.mark.skipif(not __import__("pffrg.h5io", fromlist=["LIB_PATH"]).LIB_PATH.exists(),
                    reason="libpffrg_h5.so not built (no HDF5 headers)")
def test_flow_writes_observable_and_checkpoint_files(tmp_path):
    # measure() of launch! (Launcher.jl:2-75): Σ and χ of every accepted step in the observable
    # file, a checkpoint per step with ct = 0; a checkpoint read back reproduces its χ exactly
    from pffrg import h5io as H
    prob, a0 = _flow_setup("square", 2, 6)
    obs, cp = tmp_path / "obs.h5", tmp_path / "cp.h5"
    meas = H.Measurement(obs, cp, ct=0.0, lattice=(prob.name, prob.size, "heisenberg", False, prob.J))
    with prob.context(n_slots=9) as ctx:
        res = F.launch_flow(ctx, prob, a0, 5.0, 1.0, max_steps=5, measurement=meas)
        with H.H5File(obs, "r") as f:
            assert len(f.keys("χ")) == len(res.chi) == res.accepted
            for L, c in zip(res.Lambdas, res.chi):
                np.testing.assert_array_equal(f.read(f"χ/{H.julia_float(L)}/diag"), c)
            assert H.read_lattice(f)[0] == prob.name
        assert meas.checkpoints == res.Lambdas
        with H.H5File(cp, "r") as f:
            Lc, dL, m, a = H.read_checkpoint(f, res.Lambdas[-2])
        assert Lc == res.Lambdas[-2] and dL > 0
        ctx.set_mesh(m)
        ctx.upload(F.A_INTER, a)
        chi = ctx.compute_chi(Lc, F.A_INTER, (1e-8, 1e-4))
    print(f"{res.accepted} steps written; checkpoint at {Lc}: chi max {np.max(np.abs(chi)):.4g}", flush=True)
    np.testing.assert_array_equal(chi, res.chi[-2])
