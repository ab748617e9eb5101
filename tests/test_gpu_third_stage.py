"""GPU: the third stage from compressed second-stage states (shared_third_stage).

For m >= 4, where the second stage comes from the shared basis W and the third
stage is neither the last nor one of the two read out together, the samples'
second-stage states psi2_j (MultiQW.cpp:441-443 propagates them per sample)
are compressed to directions U_r (real parts and imaginary parts, each to
2e-7 of the state's norm) and the third stage is a GEMM over T_l(Ht_3) U_r.
Checked per sample against the same solves with the third stage as terms and
against the oracle with FP64 norm sums; where the compression does not pay the
samples continue from their stored second-stage states.
"""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print("[case]", *a, flush=True)


# (the states of 100 samples span about 2 x 38 of 200 directions; with fewer samples the
# compression gains less than the GEMM costs)
@pytest.mark.parametrize("n,m,samples,kw", [
    (16, 5, 100, {}),                                  # stages 4, 5 read out together
    (15, 4, 100, dict(adjoint_two_stages=False)),      # the fourth stage read out alone
    (15, 6, 100, {}),                                  # the fourth stage as terms between
])
def test_third_stage_vs_terms_and_oracle(ctx, oracle, n, m, samples, kw):
    recs = M.generate_sk(n, 2, 500 + n + m)
    _say("third stage n", n, "m", m, "samples", samples, kw)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, **kw)
    st = ctx.last_stats()
    pt, ttt, spt = ctx.solve(n, m, recs, samples=samples, with_samples=True, shared_third_stage=False, **kw)
    stt = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d_terms = np.max(np.abs(sp - spt))
    d_oracle = np.max(np.abs(sp - spo))
    print(f"[mid] n={n} m={m}: {st['third_stage_instances']} instances, states {st['third_stage_states']} -> "
          f"{st['third_stage_vectors']} directions, term sweeps {stt['sweeps']} -> {st['sweeps']}, "
          f"max|dp| vs terms {d_terms:.3e}, vs oracle {d_oracle:.3e}", flush=True)
    assert st["third_stage_instances"] == len(recs) and stt["third_stage_instances"] == 0
    assert st["third_stage_vectors"] < 0.8 * st["third_stage_states"]
    assert st["sweeps"] < stt["sweeps"]
    if m == 5:
        assert st["sweeps"] == 0  # every stage from a shared basis
    assert np.array_equal(tt, ttt) and np.array_equal(tt, tto)
    assert d_terms < 2e-6
    assert d_oracle < 1e-6


def test_third_stage_headline_size(ctx, oracle):
    """n = 20, m = 5, 100 samples: no term sweep at all, against the oracle."""
    n, m, samples = 20, 5, 100
    recs = M.generate_sk(n, 1, 3)
    _say("third stage n=20 m=5,", samples, "samples")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d = np.max(np.abs(sp - spo))
    print(f"[mid] n=20 m=5: states {st['third_stage_states']} -> {st['third_stage_vectors']} directions, "
          f"gemm launches {st['gemm_launches']}, sweeps {st['sweeps']}, max|dp| vs oracle {d:.3e}", flush=True)
    assert st["third_stage_instances"] == 1 and st["adjoint2_instances"] == 1 and st["sweeps"] == 0
    assert np.array_equal(tt, tto)
    assert d < 1e-6


def test_third_stage_not_paying_continues_from_states(ctx, oracle):
    """32 samples: the second stage is shared but the samples' states barely compress, so the
    third stage runs as terms from the stored second-stage states."""
    n, m, samples = 16, 5, 32
    recs = M.generate_sk(n, 2, 611)
    _say("third stage that does not pay, 32 samples")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d = np.max(np.abs(sp - spo))
    print(f"[mid] 32 samples: {st['third_stage_instances']} compressed, gemm launches {st['gemm_launches']}, "
          f"sweeps {st['sweeps']}, max|dp| vs oracle {d:.3e}", flush=True)
    assert st["third_stage_instances"] == 0 and st["gemm_launches"] > 0 and st["sweeps"] > 0
    assert np.array_equal(tt, tto)
    assert d < 1e-6


def test_third_stage_in_uneven_batch_with_overrides(ctx, oracle):
    """Compressed third stages beside every other path, with schedule overrides."""
    inst = []
    rng = np.random.default_rng(29)
    for k, (n, m) in enumerate([(15, 5), (16, 3), (14, 5), (15, 1), (16, 4), (15, 6), (14, 2)]):
        rec = M.generate_sk(n, 1, 3100 + k)[0]
        s = M.schedule_heuristic(n, m, rec)
        g = (s.gamma * rng.uniform(0.7, 1.3, m)).astype(np.float32)
        inst.append((n, m, rec, g, s.short_t))
    _say("third stage in an uneven batch")
    p, tt, sp = ctx.solve_batch(inst, samples=100, with_samples=True)
    st = ctx.last_stats()
    print(f"[mid] uneven: {st['third_stage_instances']} third stages, {st['adjoint2_instances']} two-stage "
          f"adjoints", flush=True)
    assert st["third_stage_instances"] == 3
    for k, (n, m, rec, g, t) in enumerate(inst):
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=100, threads=THREADS, with_samples=True,
                                    gamma=g[None, :], short_t=t[None, :], norm="f64")
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < 1e-6, (n, m)


def test_third_stage_overlapped_batches_bit_identical(ctx):
    """Several batches of bases overlapped on two streams against one after the other."""
    n, m, samples = 16, 5, 100
    recs = M.generate_sk(n, 4, 715)
    cap = 6 << 30  # a quarter per batch of bases: one instance each
    _say("third stage, overlapped batches")
    # (one launch per recurrence step: the on-chip recurrences do not overlap batches)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, max_device_bytes=cap,
                          chain_recurrences=False)
    st = ctx.last_stats()
    p2, tt2, sp2 = ctx.solve(n, m, recs, samples=samples, with_samples=True, max_device_bytes=cap,
                             overlap_bases=False, chain_recurrences=False)
    print(f"[mid] batches {st['basis_batches']} (overlapped {st['overlap_batches']})", flush=True)
    assert st["third_stage_instances"] == 4 and st["basis_batches"] >= 2 and st["overlap_batches"] >= 1
    assert np.array_equal(tt, tt2) and np.array_equal(sp, sp2)


def test_third_stage_while_next_batch_runs_n21(ctx):
    """n = 21, m = 5: a batch's third-stage recurrences run on the main stream while the next
    batch's second-stage recurrences are still running on the second stream, each on its own job
    tables (the mixed-size benchmark's draws, whose n = 21 instances once read the other's)."""
    rng = np.random.default_rng(3)
    inst = []
    for k in range(512):
        nn, mm = int(rng.integers(12, 23)), int(rng.integers(1, 6))
        if (nn, mm) == (21, 5) and len(inst) < 3:
            inst.append((nn, mm, M.generate_sk(nn, 1, 3 * 100003 + k)[0]))
    assert len(inst) == 3
    _say("third stage n=21 m=5 beside the next batch's recurrences")
    p, tt, sp = ctx.solve_batch(inst, samples=100, with_samples=True)
    st = ctx.last_stats()
    p2, tt2, sp2 = ctx.solve_batch(inst, samples=100, with_samples=True, overlap_bases=False)
    p3, tt3, sp3 = ctx.solve_batch(inst, samples=100, with_samples=True, shared_third_stage=False)
    d = np.max(np.abs(sp - sp3))
    print(f"[mid] n=21: batches {st['basis_batches']} (overlapped {st['overlap_batches']}), third stages "
          f"{st['third_stage_instances']}, max|dp| vs terms {d:.3e}", flush=True)
    assert st["third_stage_instances"] == 3 and st["overlap_batches"] >= 1
    assert np.array_equal(tt, tt2) and np.array_equal(sp, sp2)
    assert np.array_equal(tt, tt3) and d < 2e-6
