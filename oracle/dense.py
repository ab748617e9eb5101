"""Dense-matrix FP64 reference for small n (numpy ``eigh``).

TEST INFRASTRUCTURE (same rules as ``oracle.py``).  It fixes what the
mathematics fixes, independent of the Chebyshev/Clenshaw machinery:

* ``hypercube(n)``: H_G = sum_i X_i as a dense matrix, (H_G b)[k] = sum_i b[k ^ 2^i].
* ``evolve(E, gamma, t, psi)``: exp(+i (diag(E) - gamma H_G) t) psi, the
  propagator MultiQW.cpp:175-236 approximates (the sign convention of
  MultiQW.cpp:92 and the +i of the odd terms at :206-217).
* ``product_probability``: the J = 0 closed form.  H = sum_i (h_i z_i - gamma X_i)
  factorises into 2x2 blocks, so the ground-state probability of any
  multi-stage schedule is a product of single-spin probabilities at any n.
"""
from __future__ import annotations

import numpy as np


def hypercube(n: int, masks=None, xy: bool = False) -> np.ndarray:
    """sum_i X_i; with masks, the X-string mixer sum_k X^{m_k} (X^m flips the bits of m).

    xy=True: every mask is a pair 2^a + 2^b and the mixer is sum_k (X_a X_b + Y_a Y_b):
    <x ^ m| . |x> = 2 where exactly one of bits a, b is set in x, else 0 (pinned by
    ``xy_pauli``, the same operator from Pauli matrices)."""
    N = 1 << n
    H = np.zeros((N, N))
    k = np.arange(N)
    for m in (masks if masks is not None else [1 << i for i in range(n)]):
        m = int(m)
        if xy:
            w = np.where(np.array([bin(int(x) & m).count("1") == 1 for x in k]), 2.0, 0.0)
            H[k ^ m, k] += w
        else:
            H[k, k ^ m] += 1.0
    return H


def xy_pauli(n: int, pairs) -> np.ndarray:
    """sum over pairs (a, b) of X_a X_b + Y_a Y_b from Kronecker products of Pauli
    matrices (qubit i = bit i of the index, qubit 0 the least significant)."""
    X = np.array([[0, 1], [1, 0]], np.complex128)
    Y = np.array([[0, -1j], [1j, 0]], np.complex128)
    I = np.eye(2, dtype=np.complex128)

    def op(ops):
        out = np.ones((1, 1), np.complex128)
        for q in range(n - 1, -1, -1):
            out = np.kron(out, ops.get(q, I))
        return out

    H = np.zeros((1 << n, 1 << n), np.complex128)
    for a, b in pairs:
        H += op({a: X, b: X}) + op({a: Y, b: Y})
    return H


def hamiltonian(E, gamma: float, masks=None, xy: bool = False) -> np.ndarray:
    E = np.asarray(E, np.float64)
    n = int(np.log2(E.size))
    return np.diag(E) - gamma * hypercube(n, masks, xy)


def evolve(E, gamma: float, t: float, psi) -> np.ndarray:
    """exp(+i H t) psi with H = diag(E) - gamma*H_G, complex128."""
    w, V = np.linalg.eigh(hamiltonian(E, gamma))
    return V @ (np.exp(1j * w * t) * (V.T @ np.asarray(psi, np.complex128)))


def walk_probability(E, gammas, times, eloc: int, reset: bool = True, masks=None, xy: bool = False):
    """Multi-stage walk from |+>, renormalised after each stage (MultiQW.cpp:432-451).

    times: [m, samples].  reset=True resets to |+> per sample (m != 1);
    reset=False chains the samples (the m == 1 path of MultiQW.cpp:426-430).
    Returns per-sample probabilities of state ``eloc``.  masks: an X-string
    mixer in place of sum_i X_i (xy: pairs with X_a X_b + Y_a Y_b).
    """
    E = np.asarray(E, np.float64)
    N = E.size
    n = int(np.log2(N))
    m, samples = np.asarray(times).shape
    eig = [np.linalg.eigh(np.diag(E) - g * hypercube(n, masks, xy)) for g in gammas]
    plus = np.full(N, 1 / np.sqrt(N), np.complex128)
    psi = plus.copy()
    out = np.zeros(samples)
    for j in range(samples):
        for i in range(m):
            w, V = eig[i]
            psi = V @ (np.exp(1j * w * float(times[i][j])) * (V.T @ psi))
            psi /= np.linalg.norm(psi)
        out[j] = abs(psi[eloc]) ** 2
        if reset:
            psi = plus.copy()
    return out


def product_probability(h, gammas, times_col, target_bits) -> float:
    """Closed-form ground-state probability for J = 0 (independent spins).

    Spin i evolves under h_i z_i - gamma X (z = +1 on bit 1) from |+>; each
    stage of ``times_col`` (one sample, m stages) is applied in turn.
    target_bits[i] in {0, 1} is the bit of the target state.
    """
    p = 1.0
    for i, hi in enumerate(h):
        psi = np.array([1.0, 1.0], np.complex128) / np.sqrt(2.0)
        for g, t in zip(gammas, times_col):
            H = np.array([[-hi, -g], [-g, hi]], np.float64)   # basis |0> (s=-1), |1> (s=+1)
            w, V = np.linalg.eigh(H)
            psi = V @ (np.exp(1j * w * t) * (V.T @ psi))
        p *= abs(psi[target_bits[i]]) ** 2
    return p
