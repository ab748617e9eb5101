"""Shared second stage: for m >= 3 the second-stage outputs of an instance's
samples are combinations of one basis W_lk = T_l(H_2/onenorm_2) V_k (V_k the
shared first-stage basis), formed on the matrix cores.  Checked per sample
against the per-sample Clenshaw path (shared_second_stage=False), the oracle
and the J = 0 closed form.  Reference: MultiQW.cpp:432-451."""
import numpy as np
import pytest

import multistageqw_amd as M
from oracle import dense as D

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5
PATH_TOL = 2e-6   # FP32 rounding of two evaluation orders


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,count,samples,kw", [
    (14, 3, 2, 64, {}),                                   # every stage from a basis: no Clenshaw terms
    (14, 4, 2, 100, dict(tile_bits=13, variant=1)),       # 4 row blocks; the third stage as terms
    (15, 5, 1, 150, {}),                                  # tiles of 128 and 22 samples
    (16, 3, 3, 64, dict(max_device_bytes=2 << 30)),       # one instance per batch of bases (W 336-394 MiB each)
    (14, 3, 2, 70, dict(group_bytes=16 << 20, use_graph=False)),  # an instance's samples over several groups
    (15, 3, 2, 64, dict(final_readout=False)),            # the last stage as terms from the GEMM's output
    (16, 5, 2, 48, dict(max_device_bytes=3 << 29)),       # one instance per batch of bases (W 203 MiB each), two middle stages as terms
])
def test_second_stage_vs_per_sample_and_oracle(ctx, oracle, n, m, count, samples, kw):
    _say("second stage n", n, "m", m, "samples", samples, kw)
    recs = M.generate_sk(n, count, 6000 + 10 * n + m)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    st = ctx.last_stats()
    _, tt_f, sp_f = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0,
                              shared_second_stage=False, **kw)
    st_f = ctx.last_stats()
    assert st["amp_terms"] < st_f["amp_terms"]            # the second stage ran no terms
    if m == 3 and kw.get("final_readout", True):
        assert st["sweeps"] == 0
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto) and np.array_equal(tt_f, tto)
    assert np.max(np.abs(sp - sp_f)) < PATH_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_second_stage_few_samples_runs_terms(ctx):
    """With few samples the shared basis would cost more than the samples' own terms."""
    _say("second stage: 4 samples")
    recs = M.generate_sk(16, 1, 6101)
    _, _, a = ctx.solve(16, 3, recs, samples=4, with_samples=True, small_n_max=0)
    st = ctx.last_stats()
    _, _, b = ctx.solve(16, 3, recs, samples=4, with_samples=True, small_n_max=0, shared_second_stage=False)
    assert st["amp_terms"] == ctx.last_stats()["amp_terms"]
    assert np.array_equal(a, b)


def test_second_stage_j0_closed_form(ctx, oracle):
    """J = 0 instance with a schedule whose first two stages fit the bases: the product formula
    per sample (the stage times are the reference's MT19937 draws around the given ones)."""
    n, m, samples = 16, 3, 64
    _say("second stage J=0 n=16")
    rng = np.random.default_rng(6201)
    rec = np.zeros(n * (n + 1) // 2)
    rec[n * (n - 1) // 2:] = rng.uniform(0.4, 1.6, n) * rng.choice([-1.0, 1.0], n)
    J, h = oracle.unpack(n, rec)
    gam = np.array([0.9, 0.5, 0.2], np.float32)
    st0 = np.array([0.6, 0.8, 0.7], np.float32)
    times = oracle.sample_times(m, samples, st0)
    bits = [1 if hi < 0 else 0 for hi in h]
    inst = [(n, m, rec, gam, st0)]
    _, _, sp = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0)
    st = ctx.last_stats()
    _, _, sp_f = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0,
                                 shared_second_stage=False)
    assert st["amp_terms"] < ctx.last_stats()["amp_terms"]
    for j in range(samples):
        ref = D.product_probability(h.astype(np.float64), gam.astype(np.float64), times[:, j], bits)
        assert abs(sp[0, j] - ref) < PROB_TOL, (j, sp[0, j], ref)
    assert np.max(np.abs(sp - sp_f)) < PATH_TOL


def test_second_stage_in_uneven_batch(ctx):
    """Mixed n and m in one call: bases where they pay, terms elsewhere."""
    _say("second stage: uneven batch")
    inst = [(14, 1, M.generate_sk(14, 1, 6301)[0]), (14, 3, M.generate_sk(14, 1, 6302)[0]),
            (15, 4, M.generate_sk(15, 1, 6303)[0]), (16, 2, M.generate_sk(16, 1, 6304)[0]),
            (15, 3, M.generate_sk(15, 1, 6305)[0])]
    _, _, a = ctx.solve_batch(inst, samples=48, with_samples=True, small_n_max=0)
    _, _, b = ctx.solve_batch(inst, samples=48, with_samples=True, small_n_max=0, shared_second_stage=False)
    assert np.max(np.abs(a - b)) < PATH_TOL


@pytest.mark.parametrize("n,m,count,samples,kw", [
    (14, 3, 2, 150, {}),                                  # tiles of 128 and 22 samples
    (16, 3, 2, 100, {}),                                  # the headline's 100 samples: one tile
    (15, 4, 2, 40, dict(tile_bits=13, variant=1)),        # one 64-sample tile, the third stage as terms
    (14, 3, 3, 32, dict(final_readout=False)),            # one 32-sample tile, the last stage as terms
    (14, 3, 1, 129, {}),                                  # tiles of 128 and 1 sample
])
def test_split_fp16_gemm_matches_fp32_gemm(ctx, oracle, n, m, count, samples, kw):
    """The split-FP16 GEMM (FP16 matrix ops on high/low operand parts, FP32 accumulation) against
    the FP32 matrix-op GEMM, the per-sample Clenshaw path and the oracle."""
    _say("split fp16 vs fp32 gemm n", n, "m", m, "samples", samples, kw)
    recs = M.generate_sk(n, count, 6500 + 10 * n + m)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    st = ctx.last_stats()
    assert st["gemm_launches"] > 0
    p32, _, sp32 = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, gemm="fp32", **kw)
    assert ctx.last_stats()["gemm_launches"] == st["gemm_launches"]
    _, _, spc = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0,
                          shared_second_stage=False, **kw)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - sp32)) < PATH_TOL
    assert np.max(np.abs(sp - spc)) < PATH_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_gemm_option_rejects_unknown():
    with pytest.raises(ValueError):
        M.api.make_opts(gemm="bf16")


def test_split_gemm_near_the_basis_limit(ctx, oracle):
    """First and second stages of 55-64 Chebyshev terms (fixed times, so every sample has the same
    s): the largest bases the shared second stage takes (kc1, kc2 <= 64, rows up to 64 x 64)."""
    n, m, samples = 14, 3, 64
    _say("split gemm near kc = 64")
    rec = M.generate_sk(n, 1, 6601)[0]
    J, h = oracle.unpack(n, rec)
    _, _, info = oracle.energy_compat(n, J, h)
    e_abs = (info["emax"] - info["e0"]) / 2
    gam = np.array([0.8, 0.5, 0.2], np.float32)
    onenorm = e_abs + np.abs(gam.astype(np.float64)) * n
    # stage times giving s = onenorm t = 44 for the first two stages (about 58 terms), a short last stage
    st = np.array([44.0 / onenorm[0], 44.0 / onenorm[1], 0.5], np.float32)
    k1 = len(oracle.cheb_coeffs(float(np.float32(onenorm[0]) * st[0])))
    k2 = len(oracle.cheb_coeffs(float(np.float32(onenorm[1]) * st[1])))
    assert 50 <= k1 <= 64 and 50 <= k2 <= 64
    inst = [(n, m, rec, gam, st)]
    _, _, a = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True)
    assert ctx.last_stats()["gemm_launches"] > 0
    _, _, b = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              gemm="fp32")
    _, _, c = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              shared_second_stage=False)
    assert ctx.last_stats()["gemm_launches"] == 0
    assert np.max(np.abs(a - b)) < PATH_TOL
    assert np.max(np.abs(a - c)) < PATH_TOL
    # fixed times: every sample is the same walk
    assert np.max(np.abs(a - a[0, 0])) < 1e-7


def test_split_gemm_smallest_bases(ctx, oracle):
    """Second-stage bases of 5 x 5, 6 x 6 and 8 x 10 terms in one batch: one, two and three 16-row
    blocks per class, the shortest runs of the split GEMM's copy pipeline (its first copies and
    waits differ below four blocks)."""
    n, m, samples = 14, 3, 64
    _say("split gemm, 1-3 blocks per class")
    gam = np.array([0.8, 0.5, 0.2], np.float32)
    inst, blocks = [], []
    for i, (s1, s2) in enumerate([(0.2, 0.2), (0.5, 0.5), (1.0, 2.0)]):
        rec = M.generate_sk(n, 1, 6700 + i)[0]
        J, h = oracle.unpack(n, rec)
        _, _, info = oracle.energy_compat(n, J, h)
        onenorm = (info["emax"] - info["e0"]) / 2 + np.abs(gam.astype(np.float64)) * n
        st = np.array([s1 / onenorm[0], s2 / onenorm[1], 0.5], np.float32)
        k1 = len(oracle.cheb_coeffs(float(np.float32(onenorm[0]) * st[0])))
        k2 = len(oracle.cheb_coeffs(float(np.float32(onenorm[1]) * st[1])))
        real_rows = sum(1 for l in range(k2) for k in range(k1) if (l & 1) == (k & 1))
        blocks.append(((real_rows + 15) // 16, (k1 * k2 - real_rows + 15) // 16))
        inst.append((n, m, rec, gam, st))
    assert blocks == [(1, 1), (2, 2), (3, 3)], blocks
    _, _, a = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True)
    assert ctx.last_stats()["gemm_launches"] > 0
    _, _, b = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              gemm="fp32")
    _, _, c = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              shared_second_stage=False)
    assert ctx.last_stats()["gemm_launches"] == 0
    assert np.max(np.abs(a - b)) < PATH_TOL
    assert np.max(np.abs(a - c)) < PATH_TOL
