"""GPU: uneven batches, schedule overrides, the autotuner, FP64 precision study.
Reference: MultiQW.cpp (per-problem loop :296-455, heuristic :386-423)."""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
PROB_TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def test_uneven_batch_vs_oracle(ctx, oracle):
    """One call with n in 4..17 and m in 1..5 (the uneven workload, scaled to oracle size)."""
    rng = np.random.default_rng(5)
    inst = []
    for k in range(14):
        n = int(rng.integers(4, 18))
        m = int(rng.integers(1, 6))
        inst.append((n, m, M.generate_sk(n, 1, 4000 + k)[0]))
    _say("uneven batch", [(n, m) for n, m, _ in inst])
    p, tt, sp = ctx.solve_batch(inst, samples=8, with_samples=True)
    for k, (n, m, rec) in enumerate(inst):
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=8, with_samples=True)
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < PROB_TOL, (n, m)


def test_schedule_override_vs_oracle(ctx, oracle):
    """Caller-supplied per-stage (gamma, t): the learned-schedule entry of mqw_solve_batch."""
    n, m = 11, 3
    recs = M.generate_sk(n, 3, 77)
    rng = np.random.default_rng(1)
    gam = rng.uniform(0.3, 3.0, (3, m)).astype(np.float32)
    st = rng.uniform(0.05, 0.6, (3, m)).astype(np.float32)
    _say("override n=11 m=3")
    p, tt = ctx.solve_batch([(n, m, recs[k], gam[k], st[k]) for k in range(3)], samples=20)
    po, tto = oracle.solve(n, m, recs, samples=20, gamma=gam, short_t=st)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(p - po)) < PROB_TOL
    ph, _ = ctx.solve(n, m, recs, samples=20)
    assert np.max(np.abs(ph - p)) > 1e-4  # the override really changed the walk


def test_autotuner_cache_and_results(ctx, oracle, tmp_path):
    n = 16
    cache = str(tmp_path / "tune.txt")
    _say("tune n=16")
    cfg = ctx.tune(n, cache)
    assert cfg["tile_bits"] in (11, 12, 13, 14) and cfg["high_pass"] in (0, 1, 2, 3)
    lines = open(cache).read().split("\n")
    assert lines[0].split()[1] == "16" and len(lines[0].split()) == 7
    again = ctx.tune(n, cache)  # reuses the recorded choice
    assert again == cfg and len(open(cache).read().strip().split("\n")) == 1
    recs = M.generate_sk(n, 1, 9)
    p, _, sp = ctx.solve(n, 2, recs, samples=4, with_samples=True)
    po, _, spo = oracle.solve(n, 2, recs, samples=4, with_samples=True, threads=4)
    assert np.max(np.abs(sp - spo)) < PROB_TOL


def test_fp64_path_agrees_with_fp32_and_oracle_fp64(ctx, oracle):
    """The precision study at small size: FP64 GPU vs FP64 oracle and vs FP32."""
    n, m = 14, 3
    recs = M.generate_sk(n, 1, 3)
    _say("fp64 n=14")
    p64, _, s64 = ctx.solve(n, m, recs, samples=6, precision=M.FP64, with_samples=True)
    p32, _, s32 = ctx.solve(n, m, recs, samples=6, with_samples=True)
    po, _, spo = oracle.solve(n, m, recs, samples=6, precision=1, with_samples=True)
    assert np.max(np.abs(s64 - spo)) < 1e-9 + 1e-7  # float32 output rounding of the probability
    assert np.max(np.abs(s64 - s32)) < PROB_TOL


def test_exact_energy_mode(ctx, oracle):
    n = 12
    recs = M.generate_sk(n, 2, 8)
    p_c, _ = ctx.solve(n, 2, recs, samples=10)
    p_e, _ = ctx.solve(n, 2, recs, samples=10, energy_mode=M.ENERGY_EXACT)
    # the diagonal differs only by the reference's float-delta drift
    assert np.max(np.abs(p_c - p_e)) < 1e-4


def test_autotuner_weighs_the_high_bit_pass(ctx, tmp_path):
    """From n = 22 the real-solve stage also times the sweep with the high-bit pass."""
    n = 23
    _say("tune n=23")
    cache = str(tmp_path / "tune23.txt")
    cfg = ctx.tune(n, cache)
    assert cfg["high_pass"] in (0, 2, 4)
    f = open(cache).read().split()
    assert int(f[1]) == n and int(f[5]) == cfg["high_pass"] and float(f[6]) > 0
