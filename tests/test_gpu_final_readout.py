"""Adjoint readout of the last stage.  The reference keeps only
|psi_m[E_loc]|^2 / ||psi_m||^2 of the last stage's output (MultiQW.cpp:441-446).
With X_k = T_k(H/onenorm) e_Eloc (real symmetric T_k), psi_m[E_loc] is
sum_k a_k <X_k, psi> over the last stage's input psi, and ||psi_m|| equals
||psi|| to within the truncated series' tail (<= tol = 1e-6), so per-sample
probabilities move by at most 2 tol relative, plus FP32 rounding.  Checked
against the propagated last stage (final_readout=False), the oracle and
closed forms."""
import numpy as np
import pytest

import multistageqw_amd as M
from oracle import dense as D

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5
# 2 tol relative (probabilities <= 1) plus the FP32 rounding of either path
READOUT_TOL = 3e-6


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,count,samples,kw", [
    (14, 2, 3, 5, {}),                                 # shared first stage + readout: no Clenshaw terms
    (14, 3, 2, 20, dict(tile_bits=13, variant=1)),     # two readout tiles per instance
    (15, 5, 2, 4, dict(tile_bits=13, variant=16)),
    (16, 3, 2, 40, dict(group_bytes=24 << 20, use_graph=False)),  # several groups
    (17, 4, 1, 3, dict(shared_first_stage=False)),     # readout after Clenshaw first stages
])
def test_final_readout_vs_propagated_and_oracle(ctx, oracle, n, m, count, samples, kw):
    _say("final readout n", n, "m", m, "samples", samples, kw)
    recs = M.generate_sk(n, count, 5000 + 10 * n + m)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    st = ctx.last_stats()
    assert st["aux_launches"] > 0
    if m == 2 and kw.get("shared_first_stage", True):
        assert st["sweeps"] == 0                     # every stage came from a basis
    _, tt_f, sp_f = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, final_readout=False,
                              **kw)
    assert ctx.last_stats()["amp_terms"] > st["amp_terms"]
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto) and np.array_equal(tt_f, tto)
    assert np.max(np.abs(sp - sp_f)) < READOUT_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_final_readout_headline_size(ctx, oracle):
    """n = 20, m = 3, the bench instances (seed 2): 2 instances x 18 samples against the oracle."""
    _say("final readout n=20 m=3")
    recs = M.generate_sk(20, 2, 2)
    _, tt, sp = ctx.solve(20, 3, recs, samples=18, with_samples=True)
    _, _, sp_f = ctx.solve(20, 3, recs, samples=18, with_samples=True, final_readout=False)
    po, tto, spo = oracle.solve(20, 3, recs, samples=18, threads=8, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - sp_f)) < READOUT_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL


def _j0_record(n, seed, scale=0.25):
    rng = np.random.default_rng(seed)
    rec = np.zeros(n * (n + 1) // 2)
    rec[n * (n - 1) // 2:] = rng.uniform(0.4, 1.6, n) * scale * rng.choice([-1.0, 1.0], n)
    return rec


@pytest.mark.parametrize("n,m,samples", [(16, 2, 5), (22, 3, 3), (25, 3, 2)])
def test_final_readout_j0_closed_form(ctx, oracle, n, m, samples):
    """J = 0: the walk factorises into spins, so every sample's probability has a closed form
    (n = 25 takes the high-bit pass in its middle stage)."""
    _say("final readout J=0 n", n, "m", m)
    rec = _j0_record(n, 60 + n)
    J, h = oracle.unpack(n, rec)
    sch = oracle.schedule(n, m, J, h)
    times = oracle.sample_times(m, samples, sch["short_t"])
    bits = [1 if hi < 0 else 0 for hi in h]
    _, tt, sp = ctx.solve(n, m, rec[None, :], samples=samples, with_samples=True, small_n_max=0)
    assert ctx.last_stats()["aux_launches"] > 0
    for j in range(samples):
        ref = D.product_probability(h.astype(np.float64), sch["gamma"].astype(np.float64), times[:, j], bits)
        assert abs(sp[0, j] - ref) < PROB_TOL, (j, sp[0, j], ref)


def test_m1_chains_are_untouched(ctx):
    """m = 1 samples chain their states (MultiQW.cpp:426-430): every stage is propagated."""
    _say("final readout: m = 1")
    recs = M.generate_sk(15, 2, 5101)
    _, _, a = ctx.solve(15, 1, recs, samples=6, with_samples=True, small_n_max=0)
    _, _, b = ctx.solve(15, 1, recs, samples=6, with_samples=True, small_n_max=0, final_readout=False)
    assert np.array_equal(a, b)


def test_long_last_stage_falls_back(ctx):
    """A last stage of more than 128 Chebyshev terms is propagated, with the option on or off alike."""
    n, m, samples = 14, 2, 3
    _say("last stage beyond the readout basis size")
    rec = M.generate_sk(n, 1, 5201)[0]
    g = np.array([0.8, 0.3], np.float32)
    t = np.array([0.5, 8.0], np.float32)     # s = onenorm * t well above 128 terms
    inst = [(n, m, rec, g, t)]
    _, _, a = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0)
    _, _, b = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, final_readout=False)
    assert np.array_equal(a, b)


def test_final_readout_in_uneven_batch(ctx):
    """m = 1 chains, m = 2..5 instances, overrides and fixed times in one call."""
    _say("final readout: uneven batch")
    inst = [
        (14, 1, M.generate_sk(14, 1, 5301)[0]),
        (14, 2, M.generate_sk(14, 1, 5302)[0]),
        (15, 3, M.generate_sk(15, 1, 5303)[0], np.array([0.9, 0.5, 0.2], np.float32),
         np.array([0.6, 0.9, 0.7], np.float32)),
        (16, 5, M.generate_sk(16, 1, 5304)[0]),
    ]
    for fixed in (False, True):
        _, _, a = ctx.solve_batch(inst, samples=7, with_samples=True, small_n_max=0, fixed_times=fixed)
        _, _, b = ctx.solve_batch(inst, samples=7, with_samples=True, small_n_max=0, fixed_times=fixed,
                                  final_readout=False)
        assert np.max(np.abs(a - b)) < READOUT_TOL


@pytest.mark.parametrize("n,m,count,samples,kw", [
    (14, 2, 2, 100, {}),                               # tiles of 64 and 36 samples
    (16, 3, 2, 20, dict(shared_second_stage=False)),   # one partial tile per instance
    (15, 4, 1, 64, dict(tile_bits=13, variant=1)),     # one full tile
])
def test_matrix_core_readout_matches_vector_readout(ctx, oracle, n, m, count, samples, kw):
    """Readout dots on the FP16 matrix cores (split operands, the default) against the FP32
    vector-unit dots (gemm="fp32"), the propagated last stage and the oracle."""
    _say("matrix-core readout n", n, "m", m, "samples", samples, kw)
    recs = M.generate_sk(n, count, 5300 + 10 * n + m)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    assert ctx.last_stats()["aux_launches"] > 0
    p32, tt32, sp32 = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, gemm="fp32", **kw)
    _, _, spp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, final_readout=False, **kw)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto) and np.array_equal(tt32, tto)
    assert np.max(np.abs(sp - sp32)) < 1e-6
    assert np.max(np.abs(sp - spp)) < READOUT_TOL
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_readout_near_the_basis_limit(ctx, oracle):
    """A last stage of 100-128 Chebyshev terms (fixed times): the largest adjoint basis the readout
    takes (kx <= 128, the blocks of both classes and a padding block), matrix-core and vector dots
    against the propagated last stage."""
    n, m, samples = 14, 2, 40
    _say("readout near kx = 128")
    rec = M.generate_sk(n, 1, 5701)[0]
    J, h = oracle.unpack(n, rec)
    _, _, info = oracle.energy_compat(n, J, h)
    e_abs = (info["emax"] - info["e0"]) / 2
    gam = np.array([0.7, 0.3], np.float32)
    onenorm = e_abs + np.abs(gam.astype(np.float64)) * n
    st = np.array([0.6, 95.0 / onenorm[1]], np.float32)
    kx = len(oracle.cheb_coeffs(float(np.float32(onenorm[1]) * st[1])))
    assert 100 <= kx <= 128
    inst = [(n, m, rec, gam, st)]
    _, _, a = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True)
    assert ctx.last_stats()["aux_launches"] > 0
    _, _, b = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              gemm="fp32")
    _, _, c = ctx.solve_batch(inst, samples=samples, with_samples=True, small_n_max=0, fixed_times=True,
                              final_readout=False)
    assert np.max(np.abs(a - b)) < 1e-6
    assert np.max(np.abs(a - c)) < READOUT_TOL
