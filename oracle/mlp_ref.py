"""FP64 numpy restatement of the learned-schedule MLP (32 -> 128 -> 128 -> 10,
GELU, masked MSE) with analytic gradients.

TEST INFRASTRUCTURE: checks libmqw's MLP kernels; pinned itself by central
finite differences in the tests.  Parameter layout as include/mqw.h states:
W1[32x128] b1 W2[128x128] b2 W3[128x10] b3, row-major W[in][out].
"""
import numpy as np
from scipy.special import erf

I, H, O = 32, 128, 10
SIZES = [("W1", I * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * O), ("b3", O)]
P = sum(s for _, s in SIZES)


def unflatten(p):
    p = np.asarray(p, np.float64)
    out, k = {}, 0
    for name, s in SIZES:
        out[name] = p[k:k + s]
        k += s
    out["W1"] = out["W1"].reshape(I, H)
    out["W2"] = out["W2"].reshape(H, H)
    out["W3"] = out["W3"].reshape(H, O)
    return out


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_grad(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(p, X):
    w = unflatten(p)
    X = np.asarray(X, np.float64)
    Z1 = X @ w["W1"] + w["b1"]
    A1 = gelu(Z1)
    Z2 = A1 @ w["W2"] + w["b2"]
    A2 = gelu(Z2)
    Y = A2 @ w["W3"] + w["b3"]
    return Y, (X, Z1, A1, Z2, A2)


def loss_grad(p, X, T, M):
    w = unflatten(p)
    Y, (X, Z1, A1, Z2, A2) = forward(p, X)
    M = np.asarray(M, np.float64)
    msum = M.sum()
    D = M * (Y - np.asarray(T, np.float64))
    loss = float(np.sum(D * (Y - T)) / msum)
    dY = 2 * D / msum
    gW3, gb3 = A2.T @ dY, dY.sum(0)
    dZ2 = (dY @ w["W3"].T) * gelu_grad(Z2)
    gW2, gb2 = A1.T @ dZ2, dZ2.sum(0)
    dZ1 = (dZ2 @ w["W2"].T) * gelu_grad(Z1)
    gW1, gb1 = X.T @ dZ1, dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam(p, g, m1, m2, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    p = p - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return p, m1, m2
