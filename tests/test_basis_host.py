"""CPU: the compression of a shared basis (mqw_basis.cpp: compress_first_stage and its
eigensolver), built with g++ into a small harness and checked against numpy's SVD.

For a basis of kc vectors in two parity classes and samples' coefficient rows, the
compression must keep every sample's state (sum_k c_jk i^(k&1) V_k) to within tol of its
norm, with the smallest rank that does so, from the Gram pair sums alone."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("basis") / "libbasis_harness.so")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT}/include", f"-I{ROOT}/multistageqw_amd/csrc", f"-I{ROCM}/include",
           os.path.join(ROOT, "tests", "native", "basis_harness.cpp"),
           os.path.join(ROOT, "multistageqw_amd", "csrc", "mqw_basis.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    L = C.CDLL(out)
    d = C.POINTER(C.c_double)
    L.harness_compress.argtypes = [C.c_int, d, C.c_int, d, C.c_double, C.POINTER(C.c_int), d, d]
    L.harness_compress.restype = C.c_int
    return L


def _pairs(V):
    """Gram pair sums in the layout of gram_pair_start: (k1 <= k2, same parity) by k2, then k1."""
    kc = V.shape[0]
    out = []
    for k2 in range(kc):
        for k1 in range(k2 & 1, k2 + 1, 2):
            out.append(float(V[k1] @ V[k2]))
    return np.array(out)


def _compress(L, V, coefs, tol):
    kc = V.shape[0]
    S = coefs.shape[0]
    g = _pairs(V)
    R = C.c_int(0)
    B = np.zeros(kc * kc)
    u = np.zeros(S * kc)
    d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    coefs = np.ascontiguousarray(coefs, np.float64)
    ok = L.harness_compress(kc, d(g), S, d(coefs), tol, C.byref(R), d(B), d(u))
    if not ok:
        return None
    R = R.value
    return R, B[: 2 * R * kc].reshape(2 * R, kc), u[: S * 2 * R].reshape(S, 2 * R)


def _class_rank(V, coefs, tol):
    """numpy: the smallest R (the same for both classes) keeping every sample's discarded part,
    both classes together, within tol of its whole norm."""
    kc = V.shape[0]
    states, tails = [], []
    for cls in (0, 1):
        ks = np.arange(cls, kc, 2)
        X = coefs[:, ks] @ V[ks]                       # samples x N, this class's parts
        U, s, Wt = np.linalg.svd(X, full_matrices=False)
        m = U * s                                      # coordinates on the right singular vectors
        states.append(X)
        tails.append(m)
    tot = (states[0] ** 2).sum(1) + (states[1] ** 2).sum(1)
    for R in range(0, kc):
        tail = sum(((t[:, R:]) ** 2).sum(1) for t in tails)
        if np.all(tail <= tol * tol * tot):
            return R
    return kc


@pytest.mark.parametrize("kc,S,rank,seed", [(16, 30, 5, 1), (37, 100, 10, 2), (64, 100, 22, 3), (200, 100, 38, 4)])
def test_compression_keeps_every_sample(harness, kc, S, rank, seed):
    rng = np.random.default_rng(seed)
    N = 2048
    # vectors with a decaying spectrum: numerically rank-deficient, like Chebyshev bases
    base = rng.standard_normal((rank + 8, N))
    mix = rng.standard_normal((kc, rank + 8)) * np.concatenate([np.ones(rank), 10.0 ** -np.arange(5, 13)])
    V = mix @ base / np.sqrt(N)
    coefs = rng.standard_normal((S, kc)) * np.exp(-0.05 * np.arange(kc))
    tol = 2e-7
    res = _compress(harness, V, coefs, tol)
    assert res is not None
    R, B, u = res
    assert R == _class_rank(V, coefs, tol)
    for cls in (0, 1):
        ks = np.arange(cls, kc, 2)
        Q = B[cls::2] @ V                              # the class's kept directions, R x N
        # unit directions, orthogonal within the class to the accuracy the Gram sums give them
        # (a direction of singular value sigma is Q = sum_j m_j X_j / sigma^2)
        G = Q @ Q.T
        assert np.allclose(np.diag(G), 1.0, atol=1e-2) and np.max(np.abs(G - np.diag(np.diag(G)))) < 1e-3
        X = coefs[:, ks] @ V[ks]
        err = np.linalg.norm(X - u[:, cls::2] @ Q, axis=1)
        tot = np.sqrt(np.linalg.norm(coefs[:, 0::2] @ V[0::2], axis=1) ** 2
                      + np.linalg.norm(coefs[:, 1::2] @ V[1::2], axis=1) ** 2)
        assert np.all(err <= 1.01 * tol * tot + 1e-13), (cls, (err / tot).max())


def test_compression_identity_rows_of_states(harness):
    """The third stage's form: the vectors are the samples' own real and imaginary parts
    (vector 2j = Re psi_j, 2j + 1 = Im psi_j, coefficient ones at 2j and 2j + 1)."""
    rng = np.random.default_rng(7)
    S, N, r = 60, 4096, 20
    basis = rng.standard_normal((r, N)) / np.sqrt(N)
    re = rng.standard_normal((S, r)) @ basis
    im = rng.standard_normal((S, r)) @ basis
    V = np.empty((2 * S, N))
    V[0::2], V[1::2] = re, im
    coefs = np.zeros((S, 2 * S))
    for j in range(S):
        coefs[j, 2 * j] = coefs[j, 2 * j + 1] = 1.0
    res = _compress(harness, V, coefs, 2e-7)
    assert res is not None
    R, B, u = res
    assert R == r  # each part spans exactly the r directions
    for cls, X in ((0, re), (1, im)):
        Q = B[cls::2] @ V
        assert np.max(np.linalg.norm(X - u[:, cls::2] @ Q, axis=1)) < 1e-9


def test_no_compression_when_full_rank(harness):
    """States that span the whole basis: compression refused (2R would not be below kc)."""
    rng = np.random.default_rng(9)
    kc, S, N = 12, 40, 256
    V = rng.standard_normal((kc, N))
    coefs = rng.standard_normal((S, kc))
    assert _compress(harness, V, coefs, 2e-7) is None
