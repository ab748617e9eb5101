"""Edge cases of the solve path: an empty batch, and degenerate ground states.

The reference keeps the first minimum of its Gray-code energy walk (strict <,
MultiQW.cpp:349-352), so with several minima E_loc is the one that comes first
in Gray order (index grey(i) = i ^ (i >> 1) at walk step i).  With h = 0 every
minimum comes with its complement, and with couplings of +-1 every energy is
an integer, exact in FP32 and FP64, so the minima are massively degenerate.
With Gaussian couplings the walk's FP32-rounded steps separate a minimum from
its complement by rounding (about 1e-4 at n = 16), and the first in Gray order
of the values the walk computed is still the one kept.
"""
import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def _record(n, seed, pm_one):
    rng = np.random.default_rng(seed)
    rec = np.zeros(n * (n + 1) // 2)
    npair = n * (n - 1) // 2
    rec[:npair] = rng.choice([-1.0, 1.0], npair) if pm_one else rng.normal(size=npair)
    return rec  # h = 0


def test_empty_batch(ctx):
    _say("empty batch")
    p, tt = ctx.solve_batch([])
    assert len(p) == 0 and len(tt) == 0


@pytest.mark.parametrize("n,pm_one", [(10, True), (14, True), (16, False), (20, True)])
def test_degenerate_minima_first_in_gray_order(ctx, oracle, n, pm_one):
    _say("degenerate minima n", n, "+-1" if pm_one else "normal", "h = 0")
    rec = _record(n, 7000 + n, pm_one)
    hp, e, info = ctx.energy(n, rec, M.ENERGY_COMPAT, want_e64=True)
    J, h = oracle.unpack(n, rec)
    hpo, eo, infoo = oracle.energy_compat(n, J, h)
    assert info["eloc"] == infoo["eloc"]
    assert np.array_equal(hp, hpo)
    N = 1 << n
    gray = np.arange(N, dtype=np.int64)
    gray ^= gray >> 1
    walk = e[gray]  # energies in the order the walk visits them
    first = int(gray[int(np.argmax(walk == walk.min()))])
    assert info["eloc"] == first
    comp = (N - 1) ^ first  # h = 0: the complement has the same energy
    if pm_one:
        # integer energies: the walk is exact, so the tie is exact
        assert np.all(e == np.round(e)) and np.count_nonzero(e == e.min()) >= 2
        assert e[first] == e[comp]
    else:
        # the walk adds FP32-rounded steps in FP64: the complement drifts by rounding only
        assert abs(e[first] - e[comp]) < 1e-3


@pytest.mark.parametrize("n,m", [(10, 2), (14, 3)])
def test_degenerate_instance_solve_vs_oracle(ctx, oracle, n, m):
    _say("degenerate instance solve n", n, "m", m)
    recs = np.stack([_record(n, 7100 + n + k, k % 2 == 0) for k in range(2)])
    p, tt, sp = ctx.solve(n, m, recs, samples=8, with_samples=True)
    po, tto, spo = oracle.solve(n, m, recs, samples=8, threads=2, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


@pytest.mark.parametrize("samples", [0, -1])
def test_bad_sample_count_is_an_error(ctx, samples):
    """samples < 1 is rejected with MQW_EINVAL before any problem is prepared (a negative
    count once reached a std::vector size and aborted the caller through the C ABI)."""
    _say("samples", samples)
    recs = M.generate_sk(10, 2, 1)
    with pytest.raises(M.MQWError) as e:
        ctx.solve(10, 2, recs, samples=samples)
    assert e.value.status == 1  # MQW_EINVAL
    with pytest.raises(M.MQWError) as e:
        ctx.solve_batch([(10, 2, recs[0]), (12, 3, M.generate_sk(12, 1, 2)[0])], samples=samples)
    assert e.value.status == 1
    # the context still works afterwards
    p, _ = ctx.solve(10, 2, recs, samples=4)
    assert np.all((p >= 0) & (p <= 1))
