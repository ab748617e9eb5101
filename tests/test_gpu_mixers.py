"""X-string mixers: H = H_P - gamma * sum_k X^{m_k} in place of the reference's
sum_i X_i (the reference's README.md:41 lists multi-X mixers as a to-do; there
is no reference implementation, so the dense FP64 propagator is the oracle)."""
import numpy as np
import pytest

import multistageqw_amd as M
from oracle import dense as D

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_single_bit_masks_are_the_hypercube(ctx, oracle):
    n, m = 12, 2
    _say("single-bit masks n=12")
    recs = M.generate_sk(n, 2, 71)
    _, tt, a = ctx.solve(n, m, recs, samples=8, with_samples=True, mixer_masks=[1 << i for i in range(n)])
    _, _, b = ctx.solve(n, m, recs, samples=8, with_samples=True, small_n_max=0, tile_bits=12)
    assert np.max(np.abs(a - b)) < 1e-6
    _, tto, c = oracle.solve(n, m, recs, samples=8, threads=2, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(a - c)) < 1e-5


RING_AND_STRINGS = [0b11 << i for i in range(7)] + [0b10000001, 0b10101, 0b11100000]


@pytest.mark.parametrize("precision,tol", [(M.FP32, 1e-5), (M.FP64, 1e-6)])
def test_xstring_mixer_vs_dense(ctx, oracle, precision, tol):
    n, m, samples = 8, 2, 5
    _say("X-string mixer vs dense, precision", precision)
    rec = M.generate_sk(n, 1, 72)[0]
    gamma = np.array([0.8, 0.45], np.float32)
    short_t = np.array([0.7, 1.1], np.float32)
    _, _, sp = ctx.solve_batch([(n, m, rec, gamma, short_t)], samples=samples, with_samples=True,
                               precision=precision, mixer_masks=RING_AND_STRINGS)
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    times = oracle.sample_times(m, samples, short_t)
    ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gamma.astype(np.float64), times, info["eloc"],
                             masks=RING_AND_STRINGS)
    assert np.max(np.abs(sp[0] - ref)) < tol
    # the mixer matters: the hypercube gives a different walk
    hyp = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gamma.astype(np.float64), times, info["eloc"])
    assert np.max(np.abs(hyp - ref)) > 1e-3


def test_xstring_mixer_at_n20_is_unit_norm_and_bounded(ctx):
    """A large instance on the mixer path: probabilities stay in [0, 1] (renormalised state)."""
    n = 20
    _say("X-string mixer n=20")
    rec = M.generate_sk(n, 1, 73)[0]
    masks = [(0b11 << i) & ((1 << n) - 1) | (1 if i == n - 1 else 0) for i in range(n)]
    masks = [mk for mk in masks if mk]
    p, _ = ctx.solve_batch([(n, 2, rec, np.array([0.6, 0.3], np.float32), np.array([0.5, 0.8], np.float32))],
                           samples=2, mixer_masks=masks)
    assert 0.0 <= p[0] <= 1.0


@pytest.mark.parametrize("masks", [[0], [1 << 12], list(range(1, 70))])
def test_bad_masks_are_errors(ctx, masks):
    rec = M.generate_sk(10, 1, 74)
    with pytest.raises(M.MQWError):
        ctx.solve(10, 2, rec, samples=2, mixer_masks=masks)


XY_PAIRS = [(1 << i) | (1 << ((i + 1) % 8)) for i in range(8)] + [0b10001, 0b1000100]   # ring + two chords


@pytest.mark.parametrize("precision,tol", [(M.FP32, 1e-5), (M.FP64, 1e-6)])
def test_xy_mixer_vs_dense(ctx, oracle, precision, tol):
    """XY mixer sum_k (X_a X_b + Y_a Y_b) over pairs (the reference README.md:41 to-do)
    against the dense FP64 propagator, from |+> (which the XY terms do not leave invariant)."""
    n, m, samples = 8, 2, 5
    _say("XY mixer vs dense, precision", precision)
    rec = M.generate_sk(n, 1, 75)[0]
    gamma = np.array([0.5, 0.3], np.float32)
    short_t = np.array([0.8, 1.2], np.float32)
    _, _, sp = ctx.solve_batch([(n, m, rec, gamma, short_t)], samples=samples, with_samples=True,
                               precision=precision, mixer_masks=XY_PAIRS, mixer="xy")
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    times = oracle.sample_times(m, samples, short_t)
    ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gamma.astype(np.float64), times, info["eloc"],
                             masks=XY_PAIRS, xy=True)
    assert np.max(np.abs(sp[0] - ref)) < tol
    # the same masks as X-strings are a different walk
    xs = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gamma.astype(np.float64), times, info["eloc"],
                            masks=XY_PAIRS)
    assert np.max(np.abs(xs - ref)) > 1e-3


def test_xy_mixer_chain_m1_vs_dense(ctx, oracle):
    """m = 1: the samples chain their states (MultiQW.cpp:426-430) under the XY mixer."""
    n, m, samples = 8, 1, 6
    _say("XY mixer m=1 chain")
    rec = M.generate_sk(n, 1, 76)[0]
    gamma = np.array([0.6], np.float32)
    short_t = np.array([0.9], np.float32)
    _, _, sp = ctx.solve_batch([(n, m, rec, gamma, short_t)], samples=samples, with_samples=True,
                               mixer_masks=XY_PAIRS, mixer="xy")
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    times = oracle.sample_times(m, samples, short_t)
    ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gamma.astype(np.float64), times, info["eloc"],
                             reset=False, masks=XY_PAIRS, xy=True)
    assert np.max(np.abs(sp[0] - ref)) < 1e-5


def test_xy_mixer_at_n20_is_bounded(ctx):
    n = 20
    _say("XY mixer n=20")
    rec = M.generate_sk(n, 1, 77)[0]
    pairs = [(1 << i) | (1 << ((i + 1) % n)) for i in range(n)]
    p, _ = ctx.solve_batch([(n, 2, rec, np.array([0.4, 0.2], np.float32), np.array([0.5, 0.8], np.float32))],
                           samples=2, mixer_masks=pairs, mixer="xy")
    assert 0.0 <= p[0] <= 1.0


@pytest.mark.parametrize("masks", [[0b1], [0b111], [0b11, 0b1000]])
def test_xy_masks_need_two_bits(ctx, masks):
    rec = M.generate_sk(10, 1, 78)
    with pytest.raises(M.MQWError):
        ctx.solve(10, 2, rec, samples=2, mixer_masks=masks, mixer="xy")
