"""GPU: the last two stages of m >= 4 samples read out together (adjoint_two_stages).

With e = e_Eloc and P_i = sum_k a_ik T_k(Ht_i) the stage propagators (T_k
real symmetric), psi_m[E_loc] = (P_{m-1}^T P_m^T e)^T psi_{m-2}.  The samples'
adjoint states P_m^T e lie in the span of the last stage's basis T_k(Ht_m) e,
which is compressed per sample to 1e-7 like the first stage's; P_{m-1}^T is
then a second-stage GEMM over T_l(Ht_{m-1}) Z_r, and the probability the
squared bilinear dot over the norm of psi_{m-2} (the reference renormalises
after each stage, MultiQW.cpp:441-446; the two truncated series are unitary to
within their coefficient tails, so the norms agree to 4 tol relative).
Checked per sample against the same solves with stage m - 1 as per-sample
Clenshaw terms, and against the oracle with FP64 norm sums.
"""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,samples,gemm", [(16, 4, 64, "split_fp16"), (15, 5, 48, "split_fp16"),
                                              (16, 4, 100, "fp32")])
def test_two_stage_adjoint_vs_stage_terms_and_oracle(ctx, oracle, n, m, samples, gemm):
    recs = M.generate_sk(n, 2, 300 + n + m)
    _say("two-stage adjoint n", n, "m", m, "samples", samples, gemm)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, gemm=gemm)
    st = ctx.last_stats()
    pt, ttt, spt = ctx.solve(n, m, recs, samples=samples, with_samples=True, gemm=gemm, adjoint_two_stages=False)
    stt = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d_terms = np.max(np.abs(sp - spt))
    d_oracle = np.max(np.abs(sp - spo))
    print(f"[adjoint2] n={n} m={m}: {st['adjoint2_instances']} instances, adjoint vectors "
          f"{st['adjoint2_uncompressed']} -> {st['adjoint2_vectors']}, term sweeps {stt['sweeps']} -> {st['sweeps']}, "
          f"max|dp| vs terms {d_terms:.3e}, vs oracle {d_oracle:.3e}", flush=True)
    assert st["adjoint2_instances"] == len(recs) and stt["adjoint2_instances"] == 0
    assert st["adjoint2_vectors"] < 0.6 * st["adjoint2_uncompressed"]
    assert st["sweeps"] < stt["sweeps"]
    if m == 4:
        assert st["sweeps"] == 0  # first two stages shared, last two read out: no term runs
    assert np.array_equal(tt, ttt) and np.array_equal(tt, tto)
    assert d_terms < 2e-6
    assert d_oracle < 1e-6


def test_two_stage_adjoint_headline_size(ctx, oracle):
    """n = 20, m = 4: every stage from shared bases, no term sweep, against the oracle."""
    n, m, samples = 20, 4, 64  # at 32 samples the stage-3 terms cost about as much: not taken
    recs = M.generate_sk(n, 1, 2)
    _say("two-stage adjoint n=20 m=4,", samples, "samples")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d = np.max(np.abs(sp - spo))
    print(f"[adjoint2] n=20 m=4: adjoint vectors {st['adjoint2_uncompressed']} -> {st['adjoint2_vectors']}, "
          f"gemm launches {st['gemm_launches']}, sweeps {st['sweeps']}, max|dp| vs oracle {d:.3e}", flush=True)
    assert st["adjoint2_instances"] == 1 and st["sweeps"] == 0 and st["gemm_launches"] >= 2
    assert np.array_equal(tt, tto)
    assert d < 1e-6


def test_two_stage_adjoint_in_uneven_batch_with_overrides(ctx, oracle):
    """Two-stage adjoint instances beside one-stage readouts and m = 1 chains, with schedule overrides."""
    inst = []
    rng = np.random.default_rng(17)
    for k, (n, m) in enumerate([(15, 4), (16, 3), (14, 5), (15, 1), (16, 4), (15, 2)]):
        rec = M.generate_sk(n, 1, 1700 + k)[0]
        s = M.schedule_heuristic(n, m, rec)
        g = (s.gamma * rng.uniform(0.7, 1.3, m)).astype(np.float32)
        inst.append((n, m, rec, g, s.short_t))
    _say("two-stage adjoint in an uneven batch")
    p, tt, sp = ctx.solve_batch(inst, samples=48, with_samples=True)
    assert ctx.last_stats()["adjoint2_instances"] == 3
    for k, (n, m, rec, g, t) in enumerate(inst):
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=48, threads=THREADS, with_samples=True,
                                    gamma=g[None, :], short_t=t[None, :], norm="f64")
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < 1e-6, (n, m)


def test_two_stage_adjoint_off_where_it_does_not_pay(ctx):
    """With few samples the per-sample stage m - 1 terms cost less: the path is not taken, and the
    results are bit for bit those with the option off."""
    n, m, samples = 15, 4, 6
    recs = M.generate_sk(n, 2, 404)
    _say("two-stage adjoint with", samples, "samples")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    assert ctx.last_stats()["adjoint2_instances"] == 0
    pt, ttt, spt = ctx.solve(n, m, recs, samples=samples, with_samples=True, adjoint_two_stages=False)
    assert np.array_equal(tt, ttt) and np.array_equal(sp, spt)


def test_two_stage_adjoint_overlapped_batches_bit_identical(ctx):
    """Several batches of bases (a device cap) overlapped on two streams against one after the other."""
    n, m, samples = 16, 4, 64
    recs = M.generate_sk(n, 6, 515)
    cap = 4 << 30  # a quarter of it per batch of bases: one instance (W and Y about 0.6 GiB)
    _say("two-stage adjoint, overlapped batches")
    # (one launch per recurrence step: the on-chip recurrences do not overlap batches)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, max_device_bytes=cap,
                          chain_recurrences=False)
    st = ctx.last_stats()
    p2, tt2, sp2 = ctx.solve(n, m, recs, samples=samples, with_samples=True, max_device_bytes=cap,
                             overlap_bases=False, chain_recurrences=False)
    st2 = ctx.last_stats()
    print(f"[adjoint2] batches {st['basis_batches']} (overlapped {st['overlap_batches']}), "
          f"{st2['basis_batches']} without overlap", flush=True)
    assert st["adjoint2_instances"] == 6 and st["basis_batches"] >= 2 and st["overlap_batches"] >= 1
    assert np.array_equal(tt, tt2) and np.array_equal(sp, sp2)


def _wide_instances(n, ms, seed):
    """Schedules whose last two stages run twice the heuristic's time: last-stage bases beyond 128
    vectors and stage m - 1 series beyond 64 terms."""
    inst = []
    for k, m in enumerate(ms):
        rec = M.generate_sk(n, 1, seed + k)[0]
        s = M.schedule_heuristic(n, m, rec)
        t = s.short_t.copy()
        t[-1] *= 2.0
        if m >= 3:
            t[-2] *= 2.0
        inst.append((n, m, rec, s.gamma.copy(), t.astype(np.float32)))
    return inst


@pytest.mark.parametrize("gemm", ["split_fp16", "fp32"])
def test_wide_bases_vs_oracle(ctx, oracle, gemm):
    """Last-stage bases of 129..254 vectors (readout and compressed adjoint, two combine jobs) and
    second-stage series of 65..128 terms on the split-FP16 path; the FP32 GEMM keeps its limits
    (64 and 128) and runs those stages as terms.  Every sample against the oracle."""
    inst = _wide_instances(16, [2, 3, 4, 5, 4], 2600)
    _say("wide bases,", gemm)
    p, tt, sp = ctx.solve_batch(inst, samples=64, with_samples=True, gemm=gemm)
    st = ctx.last_stats()
    print(f"[wide] {gemm}: adjoint2 {st['adjoint2_instances']} instances, kx sum {st['adjoint2_uncompressed']} -> "
          f"{st['adjoint2_vectors']}, sweeps {st['sweeps']}, gemm launches {st['gemm_launches']}", flush=True)
    if gemm == "split_fp16":
        assert st["adjoint2_instances"] == 3
        assert st["adjoint2_uncompressed"] > 128 * 3
    else:
        assert st["adjoint2_instances"] == 0
    for k, (n, m, rec, g, t) in enumerate(inst):
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=64, threads=THREADS, with_samples=True,
                                    gamma=g[None, :], short_t=t[None, :], norm="f64")
        d = np.max(np.abs(sp[k] - spo[0]))
        print(f"[wide] {gemm} m={m}: max|dp| vs oracle {d:.3e}", flush=True)
        assert tt[k] == tto[0]
        assert d < 1e-6, (m, d)
