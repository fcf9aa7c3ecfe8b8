"""Learned parquet warm start on the GPU: the MLP kernels against the numpy restatement
(tests/mlp_ref.py, same BF16 roundings, FP64 sums), the decoder against numpy, training, and
the warm start's effect on the parquet iteration count.

Tolerances: forward outputs and gradients agree to <= 2e-3 relative (norm per array): FP32
MFMA accumulation against FP64 sums, plus the occasional BF16 rounding that lands on the other
side of a tie."""
import numpy as np
import pytest

from pffrg import flow as F
from pffrg import learn

import mlp_ref as R

pytestmark = pytest.mark.gpu

DIMS = (2, 64, 3, 20)


def _net(seed=4):
    return learn.MLP(*DIMS, seed=seed)


def _relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_mlp_forward_matches_reference():
    net = _net()
    p = net.get_params().astype(np.float64)
    x = np.random.default_rng(1).standard_normal((37, 2)).astype(np.float32)
    y = net.forward(x)
    ref, _, _ = R.forward(p, R.dims_of(*DIMS), x)
    e = _relnorm(y, ref)
    print(f"forward rel err {e:.2e}", flush=True)
    assert e <= 2e-3


def test_mlp_gradients_match_reference():
    net = _net()
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((37, 2)).astype(np.float32)
    t = rng.standard_normal((37, 20)).astype(np.float32)
    loss, g = net.backward(x, t, want_grads=True)
    rl, rg = R.loss_grad(p, R.dims_of(*DIMS), x, t)
    dims = R.dims_of(*DIMS)
    o = 0
    for l in range(len(dims) - 1):
        for n in (dims[l] * dims[l + 1], dims[l + 1]):
            e = _relnorm(g[o:o + n], rg[o:o + n])
            print(f"layer {l} block of {n}: grad rel err {e:.2e}", flush=True)
            assert e <= 2e-3
            o += n
    assert abs(loss - rl) <= 1e-3 * rl


def test_decoder_matches_numpy():
    rng = np.random.default_rng(3)
    mean = rng.standard_normal(10007).astype(np.float32)
    basis = rng.standard_normal((13, 10007)).astype(np.float32)
    coef = rng.standard_normal(13).astype(np.float32)
    got = learn.Decoder(mean, basis)(coef)
    ref = mean.astype(np.float64) + coef.astype(np.float64) @ basis.astype(np.float64)
    assert _relnorm(got, ref) <= 1e-6


def test_mlp_training_reduces_loss():
    net = learn.MLP(2, 128, 4, 8, seed=3)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (64, 2)).astype(np.float32)
    t = np.stack([np.sin(k * x[:, 0]) * np.cos(x[:, 1]) for k in range(1, 9)], axis=1).astype(np.float32)
    l0 = net.backward(x, t)
    for _ in range(600):
        net.backward(x, t)
        net.adam(2e-3)
    l1 = net.backward(x, t)
    print(f"loss {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 1e-2 * l0


def _parquet_problem(j2, lam, N=6):
    return F.Problem.build("square", 2, [[1.0], [j2]], initial=lam, num_sigma=N + 2, num_Omega=N, num_nu=N)


def _solve(prob, lam, a0, max_iter=200):
    with prob.context(n_slots=9) as ctx:
        res = F.launch_parquet(ctx, lam, a0, max_iter=max_iter, parquet_tol=(1e-8, 0.0))
        print(f"  parquet at L={lam:.3f}: {res.iterations} iterations, residual {res.abs_err:.1e}", flush=True)
        return res, ctx.download(F.A)


def test_warm_start_reduces_parquet_iterations():
    """Parquet solutions of the square J1-J2 model (L=2, N=6) at 12 seeded (J2, Λ in [1, 5])
    train the model; at 3 held-out couplings the decoded warm start converges to the same
    absolute residual (1e-8) in fewer iterations than the bare start."""
    pts = learn.coupling_samples(15, seed=5, lam=(1.0, 5.0))
    train, test = pts[:12], pts[12:]
    acts = []
    for j2, lam in train:
        prob = _parquet_problem(j2, lam)
        res, a = _solve(prob, lam, prob.bare_action())
        assert res.abs_err < 1e-8
        acts.append(a)
    ws = learn.WarmStart(max_coef=256, width=256, n_layers=4, seed=1)
    rep = ws.fit(train, acts, steps=1500, lr=1e-3)
    print(f"coefficients per block {rep.n_coef}, final loss {rep.losses[-1]:.2e}, train rel err {rep.train_rel_err:.2e}",
          flush=True)
    gains = []
    for j2, lam in test:
        prob = _parquet_problem(j2, lam)
        bare = prob.bare_action()
        rb, ab = _solve(prob, lam, bare)
        rw, aw = _solve(prob, lam, ws.initial_action((j2, lam), bare))
        d = float(np.max(np.abs(aw.gamma - ab.gamma)) / np.max(np.abs(ab.gamma)))
        print(f"J2={j2:+.3f} L={lam:.3f}: bare {rb.iterations} it, warm {rw.iterations} it, "
              f"residuals {rb.abs_err:.1e}/{rw.abs_err:.1e}, solutions differ by {d:.1e}", flush=True)
        assert rb.abs_err < 1e-8 and rw.abs_err < 1e-8
        gains.append(rb.iterations / rw.iterations)
    assert np.mean(gains) > 1.0
