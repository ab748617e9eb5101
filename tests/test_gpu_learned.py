"""GPU: the learned schedule's kernels against the FP64 oracle, the instance
features, and what the trained network gains over the heuristic."""
import numpy as np
import pytest

import multistageqw_amd as M
from multistageqw_amd import learned as Lr
from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_mlp_forward_grad_adam_vs_fp64(ctx):
    _say("mlp kernels vs fp64")
    mlp = Lr.MLP(ctx, seed=3)
    p = mlp.params().astype(np.float64)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (37, 32)).astype(np.float32)
    T = rng.normal(0, 0.5, (37, 10)).astype(np.float32)
    Mk = Lr.loss_mask(rng.integers(1, 6, 37))
    Y = mlp.forward(X)
    Yr, _ = R.forward(p, X)
    assert np.max(np.abs(Y - Yr)) < 1e-5 * max(1.0, np.max(np.abs(Yr)))
    loss, g = mlp.loss_grad(X, T, Mk, want_grads=True)
    lr_, gr = R.loss_grad(p, X, T, Mk)
    assert abs(loss - lr_) < 1e-5 * max(1.0, lr_)
    assert np.max(np.abs(g - gr)) < 1e-4 * max(1e-3, np.max(np.abs(gr)))
    mlp.adam(lr=1e-3)
    p1 = mlp.params().astype(np.float64)
    pr, _, _ = R.adam(p, g.astype(np.float64), np.zeros(R.P), np.zeros(R.P), 1)
    assert np.max(np.abs(p1 - pr)) < 1e-6
    mlp.set_params(p.astype(np.float32))
    assert np.array_equal(mlp.params(), p.astype(np.float32))
    mlp.close()


def test_mlp_training_reduces_loss(ctx):
    _say("mlp training")
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (512, 32)).astype(np.float32)
    W = rng.normal(0, 0.3, (32, 10))
    T = np.tanh(X @ W).astype(np.float32)
    Mk = np.ones((512, 10), np.float32)
    mlp = Lr.MLP(ctx, seed=1)
    l0 = mlp.loss_grad(X, T, Mk)
    for _ in range(300):
        mlp.loss_grad(X, T, Mk)
        mlp.adam(lr=3e-3)
    assert mlp.loss_grad(X, T, Mk) < 0.1 * l0
    mlp.close()


def test_features_vs_oracle(ctx, oracle):
    _say("features")
    inst = [(n, m, M.generate_sk(n, 1, 60 + n)[0]) for n, m in [(8, 2), (12, 3), (15, 1)]]
    F = Lr.features(ctx, inst)
    for k, (n, m, rec) in enumerate(inst):
        J, h = oracle.unpack(n, rec)
        hp, _, info = oracle.energy_compat(n, J, h)
        t = np.array([2.0 * J[i, j] for i in range(1, n) for j in range(i)], np.float64)
        hh = h.astype(np.float64)
        ref = [n / 20, m / 5, t.mean(), t.var(), np.abs(t).max(), hh.mean(), hh.var(), np.abs(hh).max()]
        assert np.allclose(F[k, :8], ref, rtol=1e-5, atol=1e-6)
        r = np.float32(info["e_abs"])
        b = np.floor((hp + r) / (np.float32(2) * r) * np.float32(24)).astype(int).clip(0, 23)
        counts = np.bincount(b, minlength=24) / float(1 << n)
        assert np.max(np.abs(F[k, 8:] - counts)) <= 1.0 / (1 << n) + 1e-7
        assert abs(F[k, 8:].sum() - 1) < 1e-6


def test_learned_schedule_pipeline(ctx):
    """The training pipeline end to end on a small problem: simulator-scored candidates
    (n = 12..14), training by the HIP kernels, evaluation on fresh instances at equal total
    time.  What the shipped network gains is tested in test_gpu_learned_shipped.py."""
    _say("learned pipeline (small)")
    data = Lr.make_dataset(ctx, 240, seed=11, n_range=(12, 14), candidates=8, samples=16)
    assert data.scored == 240 * 8
    # candidate 0 of every instance is the heuristic, so the best is at least as good
    assert np.all(data.p_best >= data.p_heur)
    assert np.mean(data.p_best > data.p_heur) > 0.5
    mlp = Lr.MLP(ctx, seed=5)
    info = Lr.train(mlp, data, epochs=150, batch=128, lr=1e-3)
    assert info["history"][-1][1] < info["history"][0][1]
    ev = Lr.evaluate(ctx, mlp, 32, seed=99, n=14, samples=32)
    print("[learned]", ev, flush=True)
    assert ev["equal_time"]
    mlp.close()
