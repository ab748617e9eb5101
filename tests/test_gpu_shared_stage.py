"""Shared first stage: the m > 1 samples of an instance start their first stage
from |+> under the same H (MultiQW.cpp:432-451 resets psi per sample), so one
real Chebyshev basis per instance and a combination per sample stand in for
every sample's own Clenshaw terms of that stage.  Checked per sample against
the oracle and against the per-sample Clenshaw path (shared_first_stage=False)."""
import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5
# forward (basis) and backward (Clenshaw) recurrences round differently in FP32
PATH_TOL = 2e-6


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,count,samples,kw", [
    (14, 2, 3, 5, {}),
    (14, 3, 2, 7, dict(tile_bits=13, variant=1)),
    (15, 5, 2, 4, dict(tile_bits=13, variant=16)),   # both halves per workgroup in the later stages
    (16, 3, 2, 3, dict(tile_bits=13, variant=10)),   # tile pairs in the later stages
    (17, 4, 1, 3, {}),
])
def test_shared_first_stage_vs_oracle_and_clenshaw(ctx, oracle, n, m, count, samples, kw):
    _say("shared first stage n", n, "m", m, "samples", samples, kw)
    recs = M.generate_sk(n, count, 4000 + 10 * n + m)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    st = ctx.last_stats()
    assert st["aux_launches"] > 0 and st["aux_bytes"] > 0
    assert st["psi_const_amp_terms"] == 0          # no first-stage Clenshaw term ran
    _, tt_off, sp_off = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0,
                                  shared_first_stage=False, **kw)
    st_off = ctx.last_stats()
    assert st_off["psi_const_amp_terms"] > 0
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto) and np.array_equal(tt_off, tto)
    assert np.max(np.abs(sp - sp_off)) < PATH_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_shared_first_stage_batches_and_groups(ctx, oracle):
    """150 samples (three combine jobs per instance) in small trajectory groups, without graphs."""
    n, m, samples = 14, 3, 150
    _say("shared first stage: 150 samples, small groups")
    recs = M.generate_sk(n, 2, 4101)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0,
                          group_bytes=40 << 20, use_graph=False)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert ctx.last_stats()["aux_launches"] > 3


def test_shared_first_stage_in_uneven_batch(ctx):
    """m = 1 chains (no shared stage) beside m > 1 instances, overrides and fixed times."""
    _say("shared first stage: uneven batch")
    inst = [
        (14, 1, M.generate_sk(14, 1, 4201)[0]),
        (14, 3, M.generate_sk(14, 1, 4202)[0]),
        (15, 2, M.generate_sk(15, 1, 4203)[0], np.array([0.9, 0.4], np.float32), np.array([0.6, 0.9], np.float32)),
        (16, 4, M.generate_sk(16, 1, 4204)[0]),
    ]
    for fixed in (False, True):
        _, _, a = ctx.solve_batch(inst, samples=6, with_samples=True, small_n_max=0, fixed_times=fixed)
        _, _, b = ctx.solve_batch(inst, samples=6, with_samples=True, small_n_max=0, fixed_times=fixed,
                                  shared_first_stage=False)
        assert np.max(np.abs(a - b)) < PATH_TOL


def test_shared_first_stage_with_high_bit_pass(ctx):
    """n = 25: the later stages take the high-bit pass; the basis recursion loads every partner."""
    n, m, samples = 25, 3, 2
    _say("shared first stage n=25 vs per-sample Clenshaw")
    rec = M.generate_sk(n, 1, 4301)
    _, _, a = ctx.solve(n, m, rec, samples=samples, with_samples=True, final_readout=False)
    assert ctx.last_stats()["aux_launches"] > 0
    _, _, b = ctx.solve(n, m, rec, samples=samples, with_samples=True, shared_first_stage=False, final_readout=False)
    assert ctx.last_stats()["aux_launches"] == 0
    assert np.max(np.abs(a - b)) < PATH_TOL


def test_long_first_stage_falls_back(ctx):
    """A first stage of more than 64 Chebyshev terms runs its own Clenshaw terms per sample,
    with the shared stage on or off alike."""
    n, m, samples = 14, 2, 3
    _say("first stage beyond the basis size")
    rec = M.generate_sk(n, 1, 4401)[0]
    g = np.array([0.8, 0.3], np.float32)
    t = np.array([4.0, 0.5], np.float32)    # s = onenorm * t > 4 * 0.8 * 14: far above 64 terms
    _, _, a = ctx.solve_batch([(n, m, rec, g, t)], samples=samples, with_samples=True, small_n_max=0,
                              final_readout=False)
    st = ctx.last_stats()
    assert st["aux_launches"] == 0 and st["psi_const_amp_terms"] > 0
    _, _, b = ctx.solve_batch([(n, m, rec, g, t)], samples=samples, with_samples=True, small_n_max=0,
                              shared_first_stage=False, final_readout=False)
    assert np.array_equal(a, b)
