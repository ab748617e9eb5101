"""Overlapped batches of shared bases: when a solve's shared bases take several
batches, the next batch's basis recurrences run on a second stream while the
current batch's GEMM, readout and sweeps run on the main one.  The kernels and
their inputs are the same as one batch after the other, so the results must be
bit-identical with overlap_bases off; also checked against the oracle.  The
on-chip recurrences (n = 15..20) take the whole device, so batches do not
overlap with them: these tests run the recurrences one launch per step."""
import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


# The memory caps give one instance per batch of second-stage bases built from every V_k
# (compress_bases=False); the last case compresses them, with a cap that still splits.
@pytest.mark.parametrize("n,m,count,samples,kw", [
    (16, 3, 3, 64, dict(max_device_bytes=2 << 30, compress_bases=False)),  # three batches (W 336-394 MiB each)
    (16, 3, 2, 64, dict(max_device_bytes=2 << 30, use_graph=False, compress_bases=False)),  # no graphs
    (16, 3, 3, 70, dict(max_device_bytes=2 << 30, group_bytes=24 << 20, compress_bases=False)),  # groups
    # middle stages as terms (no two-stage adjoint, whose bases would not fit the cap twice); the
    # seed of the second-stage test, whose two instances both share their second stage (W 203 MiB each)
    (16, 5, 2, 48, dict(max_device_bytes=3 << 29, seed=6165, compress_bases=False, adjoint_two_stages=False)),
    # compressed bases (W about 190-250 MiB each, 336-456 MiB uncompressed): count = 0 means
    # at least two batches of 512 MiB
    (16, 3, 0, 64, dict(max_device_bytes=2 << 30)),
])
def test_overlapped_bases_identical(ctx, oracle, n, m, count, samples, kw):
    _say("overlapped bases n", n, "m", m, "count", count, "samples", samples, kw)
    kw = dict(kw)
    recs = M.generate_sk(n, count or 3, kw.pop("seed", 7000 + 10 * n + m))
    kw["chain_recurrences"] = False
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, **kw)
    st = ctx.last_stats()
    batches = count or st["basis_batches"]
    assert st["basis_batches"] == batches >= 2 and st["overlap_batches"] == batches
    assert st["compressed_instances"] == (0 if kw.get("compress_bases") is False else len(recs))
    p_off, tt_off, sp_off = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0,
                                      overlap_bases=False, **kw)
    st_off = ctx.last_stats()
    assert st_off["basis_batches"] == batches and st_off["overlap_batches"] == 0
    assert st["gemm_launches"] == st_off["gemm_launches"] > 0
    assert np.array_equal(sp, sp_off) and np.array_equal(p, p_off) and np.array_equal(tt, tt_off)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_overlap_repeated_solves(ctx):
    """Back-to-back overlapped solves on one context reuse both basis buffers."""
    n, m, samples = 16, 3, 64
    recs = M.generate_sk(n, 3, 7101)
    _say("overlapped bases: three solves in a row")
    kw = dict(samples=samples, with_samples=True, small_n_max=0, max_device_bytes=2 << 30, compress_bases=False,
              chain_recurrences=False)
    first = ctx.solve(n, m, recs, **kw)[2]
    for _ in range(2):
        again = ctx.solve(n, m, recs, **kw)[2]
        assert ctx.last_stats()["overlap_batches"] == 3
        assert np.array_equal(first, again)


def test_single_batch_does_not_overlap(ctx):
    """One batch of bases has nothing to overlap with: it runs on the main stream."""
    _say("overlapped bases: one batch")
    recs = M.generate_sk(14, 2, 7201)
    ctx.solve(14, 3, recs, samples=64, small_n_max=0)
    st = ctx.last_stats()
    assert st["basis_batches"] == 1 and st["overlap_batches"] == 0


def test_overlap_in_uneven_batch(ctx):
    """Several batches of bases in one n-group of an uneven batch, with m = 1 chains (no bases),
    an instance whose second stage stays as terms (few samples would not pay) and override
    schedules: overlapped and one-after-the-other results are bit-identical."""
    _say("overlapped bases: uneven batch")
    inst = [
        (16, 3, M.generate_sk(16, 1, 7301)[0]),
        (16, 1, M.generate_sk(16, 1, 7302)[0]),
        (16, 3, M.generate_sk(16, 1, 7303)[0]),
        (16, 4, M.generate_sk(16, 1, 7304)[0], np.array([0.9, 0.6, 0.4, 0.2], np.float32),
         np.array([0.5, 0.7, 0.6, 0.4], np.float32)),
        (16, 3, M.generate_sk(16, 1, 7305)[0]),
        (14, 3, M.generate_sk(14, 1, 7306)[0]),
    ]
    # (the m = 4 instance's two-stage adjoint bases would not fit the cap twice: off here; its
    # overlapped batches are tested with the two-stage adjoint)
    kw = dict(samples=64, with_samples=True, small_n_max=0, max_device_bytes=2 << 30, adjoint_two_stages=False,
              chain_recurrences=False)
    p, tt, sp = ctx.solve_batch(inst, **kw)
    st = ctx.last_stats()
    assert st["overlap_batches"] >= 2
    p0, tt0, sp0 = ctx.solve_batch(inst, overlap_bases=False, **kw)
    assert ctx.last_stats()["overlap_batches"] == 0
    assert np.array_equal(sp, sp0) and np.array_equal(p, p0) and np.array_equal(tt, tt0)
