"""The dense XY mixer of the oracle (the bit rule the GPU kernel implements:
2 psi[x ^ m] where exactly one bit of the pair m is set in x) against the same
operator built from Pauli matrices, and the symmetry that defines it."""
import numpy as np
import pytest

from oracle import dense as D

PAIRS = [(0, 1), (1, 2), (0, 3), (2, 4), (1, 4)]


def test_bit_rule_equals_pauli_products():
    n = 5
    masks = [(1 << a) | (1 << b) for a, b in PAIRS]
    H = D.hypercube(n, masks, xy=True)
    P = D.xy_pauli(n, PAIRS)
    assert np.max(np.abs(P.imag)) < 1e-12
    assert np.array_equal(H, P.real)
    assert np.array_equal(H, H.T)


def test_xy_keeps_hamming_weight_and_norm_bound():
    n = 5
    masks = [(1 << a) | (1 << b) for a, b in PAIRS]
    H = D.hypercube(n, masks, xy=True)
    w = np.array([bin(x).count("1") for x in range(1 << n)])
    rows, cols = np.nonzero(H)
    assert np.all(w[rows] == w[cols])
    # the engine's spectral bound: 2 per pair
    assert np.max(np.abs(np.linalg.eigvalsh(H))) <= 2 * len(PAIRS) + 1e-9


@pytest.mark.parametrize("a,b", [(0, 1), (2, 4)])
def test_single_pair_spectrum(a, b):
    """X_a X_b + Y_a Y_b has eigenvalues +-2 on the one-excitation subspace and 0 elsewhere."""
    n = 5
    ev = np.sort(np.linalg.eigvalsh(D.hypercube(n, [(1 << a) | (1 << b)], xy=True)))
    N = 1 << n
    assert np.allclose(ev[: N // 4], -2) and np.allclose(ev[-N // 4:], 2) and np.allclose(ev[N // 4: -N // 4], 0)
