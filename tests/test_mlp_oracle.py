"""CPU pins of the FP64 MLP restatement (oracle/mlp_ref.py) and of the schedule
arithmetic of the learned path (multistageqw_amd/learned.py)."""
import numpy as np
from scipy.special import erf

from oracle import mlp_ref as R


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, R.P)
    X = rng.normal(0, 1, (7, R.I))
    T = rng.normal(0, 1, (7, R.O))
    M = (rng.uniform(size=(7, R.O)) > 0.3).astype(float)
    loss, g = R.loss_grad(p, X, T, M)
    for k in rng.choice(R.P, 60, replace=False):
        e = np.zeros(R.P)
        e[k] = 1e-6
        fd = (R.loss_grad(p + e, X, T, M)[0] - R.loss_grad(p - e, X, T, M)[0]) / 2e-6
        assert abs(fd - g[k]) < 1e-6 + 1e-5 * abs(g[k]), k


def test_gelu_and_its_derivative():
    z = np.linspace(-6, 6, 101)
    assert np.allclose(R.gelu(z), z * 0.5 * (1 + erf(z / np.sqrt(2))))
    fd = (R.gelu(z + 1e-6) - R.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(R.gelu_grad(z), fd, atol=1e-7)


def test_masked_loss_ignores_masked_outputs():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, R.P)
    X = rng.normal(0, 1, (4, R.I))
    T = rng.normal(0, 1, (4, R.O))
    M = np.zeros((4, R.O))
    M[:, :6] = 1
    T2 = T.copy()
    T2[:, 6:] += 100.0
    assert R.loss_grad(p, X, T, M)[0] == R.loss_grad(p, X, T2, M)[0]


def test_apply_offsets_keeps_total_time():
    from multistageqw_amd.learned import apply_offsets, loss_mask
    g, t = np.array([3.0, 1.2, 0.4], np.float32), np.array([0.1, 0.3, 0.5], np.float32)
    g0, t0 = apply_offsets(g, t, np.zeros(10))
    assert np.allclose(g0, g) and np.allclose(t0, t)
    g1, t1 = apply_offsets(g, t, np.array([0.3, -0.5, 0.0, 0.8, -0.2, 0.1, 9, 9, 9, 9]))
    assert abs(t1.sum() - t.sum()) < 1e-6
    assert np.allclose(g1, g * np.exp([0.3, 0.0, -0.2]), rtol=1e-6)
    assert loss_mask([3, 1]).tolist()[1] == [1, 1] + [0] * 8
