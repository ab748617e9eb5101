"""GPU: the compressed first-stage basis of the shared second stage.

Where the second stage comes from the shared basis, the samples' first-stage
states sum_k c_j[k] i^(k&1) T_k(Ht_1)|+> are first reduced to the smallest set
of orthonormal combinations Q_r (per parity class) that holds every sample's
state to within 1e-7 relative, and the second-stage basis is T_l(Ht_2) Q_r.
The compressed path is compared per sample with the same solve built from
every V_k (compress_bases = False), with the per-sample Clenshaw path and with
the oracle (MultiQW.cpp:432-454).
"""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,samples,gemm", [(14, 3, 100, "split_fp16"), (16, 4, 64, "split_fp16"),
                                              (16, 3, 150, "fp32"), (20, 3, 100, "split_fp16")])
def test_compressed_matches_full_basis(ctx, n, m, samples, gemm):
    recs = M.generate_sk(n, 3, 70 + n)
    _say("compressed vs full basis n", n, "m", m, "samples", samples, gemm)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, gemm=gemm)
    st = ctx.last_stats()
    pf, ttf, spf = ctx.solve(n, m, recs, samples=samples, with_samples=True, gemm=gemm, compress_bases=False)
    stf = ctx.last_stats()
    print(f"[compress] n={n} m={m}: {st['compressed_instances']} instances, vectors "
          f"{st['uncompressed_vectors']} -> {st['compressed_vectors']}, max|dp|={np.max(np.abs(sp - spf)):.3e}",
          flush=True)
    assert st["compressed_instances"] == len(recs) and stf["compressed_instances"] == 0
    assert st["compressed_vectors"] < 0.8 * st["uncompressed_vectors"]
    assert st["gemm_launches"] > 0 and stf["gemm_launches"] > 0
    assert np.array_equal(tt, ttf)
    assert np.max(np.abs(sp - spf)) < 2e-7


def test_compressed_vs_oracle_and_terms(ctx, oracle):
    n, m, samples = 15, 5, 48
    recs = M.generate_sk(n, 2, 91)
    _say("compressed n=15 m=5 vs oracle and per-sample terms")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    assert ctx.last_stats()["compressed_instances"] == 2
    pt, _, spt = ctx.solve(n, m, recs, samples=samples, with_samples=True, shared_first_stage=False,
                           shared_second_stage=False, final_readout=False)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spt)) < 3e-6
    assert np.max(np.abs(sp - spo)) < 1e-6


def test_compressed_in_uneven_batch_with_overrides(ctx, oracle):
    """Compressed and uncompressed instances side by side in one batch, with schedule overrides."""
    inst = []
    rng = np.random.default_rng(4)
    for k, (n, m) in enumerate([(15, 3), (16, 2), (15, 4), (14, 1), (16, 3)]):
        rec = M.generate_sk(n, 1, 800 + k)[0]
        s = M.schedule_heuristic(n, m, rec)
        g = (s.gamma * rng.uniform(0.7, 1.3, m)).astype(np.float32)
        inst.append((n, m, rec, g, s.short_t))
    _say("compressed in an uneven batch")
    p, tt, sp = ctx.solve_batch(inst, samples=40, with_samples=True)
    assert ctx.last_stats()["compressed_instances"] >= 2
    for k, (n, m, rec, g, t) in enumerate(inst):
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=40, threads=THREADS, with_samples=True,
                                    gamma=g[None, :], short_t=t[None, :], norm="f64")
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < 1e-6, (n, m)


def test_compression_in_chunks_is_bit_identical(ctx):
    """A device cap under which the first-stage bases of the compressed instances (computed before
    the batches) take several chunks gives the same results bit for bit as one chunk."""
    n, m, samples = 14, 3, 40
    ni = 64  # their V together exceed the cap's V share while each instance's W fits its share
    recs = M.generate_sk(n, ni, 990)
    _say("compression in chunks,", ni, "instances n=14")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    assert st["compressed_instances"] == ni and st["first_stage_chunks"] == 1
    tried = []
    for cap_mib in (640, 384, 320, 256, 224):
        pc, ttc, spc = ctx.solve(n, m, recs, samples=samples, with_samples=True, max_device_bytes=cap_mib << 20)
        stc = ctx.last_stats()
        tried.append((cap_mib, stc["compressed_instances"], stc["first_stage_chunks"]))
        if stc["compressed_instances"] != ni:
            continue
        assert np.array_equal(tt, ttc)
        assert np.array_equal(sp, spc), cap_mib
        if stc["first_stage_chunks"] >= 2:
            break
    print("[chunks] (cap MiB, compressed, chunks):", tried, flush=True)
    assert any(c == ni and k >= 2 for _, c, k in tried), tried


def test_direct_m1_in_chunks_is_bit_identical(ctx):
    """Direct m = 1 chains whose first-stage bases do not fit together run one chunk at a time
    with the same results bit for bit."""
    n, samples = 16, 40
    recs = M.generate_sk(n, 6, 1990)
    _say("direct m=1 in chunks, 6 instances n=16")
    p, tt, sp = ctx.solve(n, 1, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    pc, ttc, spc = ctx.solve(n, 1, recs, samples=samples, with_samples=True, max_device_bytes=64 << 20)
    stc = ctx.last_stats()
    print(f"[chunks] direct m=1: {st['first_stage_chunks']} -> {stc['first_stage_chunks']} chunks", flush=True)
    assert st["direct_chains"] == stc["direct_chains"] == 6
    assert st["first_stage_chunks"] == 1 and stc["first_stage_chunks"] >= 2
    assert np.array_equal(tt, ttc) and np.array_equal(sp, spc)


def test_compressed_with_few_samples(ctx, oracle):
    """Just above the sample count where the shared second stage pays: few rows per class."""
    n, m, samples = 16, 3, 30
    recs = M.generate_sk(n, 2, 1201)
    _say("compressed, 30 samples")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    pf, _, spf = ctx.solve(n, m, recs, samples=samples, with_samples=True, compress_bases=False)
    print(f"[compress] 30 samples: {st['compressed_instances']} compressed, vectors "
          f"{st['uncompressed_vectors']} -> {st['compressed_vectors']}", flush=True)
    if st["gemm_launches"] == 0:
        pytest.skip("the shared second stage does not pay at this sample count for these instances")
    # an instance whose samples need 2R >= kc directions keeps its V_k
    assert st["compressed_instances"] >= 1
    assert np.max(np.abs(sp - spf)) < 2e-7
    po, _, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    assert np.max(np.abs(sp - spo)) < 1e-6
