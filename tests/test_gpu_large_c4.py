"""GPU: the large single instance (n = 28, m = 5) and the FP32 / FP64 precision study.

* n = 24, m = 3 SK instance: the FP32 path (complex FP32 state, the
  benchmark's arithmetic) against the all-FP64 GPU path, per sample.
* n = 18 and 20: the FP64 GPU path against the FP64 oracle, per sample.
* n = 28, m = 5 SK instance (the large-instance workload): FP32 against the
  all-FP64 GPU path on one sample; a probability in [0, 1].
* n = 28, m = 5 with J = 0: the closed-form product probability, which holds
  at any size (oracle/dense.py).
Reference semantics: MultiQW.cpp:432-454 (stages, renormalisation, readout).
"""
import numpy as np
import pytest

import multistageqw_amd as M
from oracle import dense as D

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def test_precision_study_n24_fp32_vs_fp64(ctx):
    n, m = 24, 3
    rec = M.generate_sk(n, 1, 5)
    _say("n=24 m=3 FP32 vs FP64, 2 samples")
    p32, tt32, s32 = ctx.solve(n, m, rec, samples=2, with_samples=True)
    p64, tt64, s64 = ctx.solve(n, m, rec, samples=2, with_samples=True, precision=M.FP64)
    d = float(np.max(np.abs(s32 - s64)))
    print(f"[precision] n=24 p32={s32.tolist()} p64={s64.tolist()} max|dp|={d:.3e}", flush=True)
    assert tt32[0] == tt64[0]
    assert d < PROB_TOL
    assert np.all((s64 > 0) & (s64 < 1))


@pytest.mark.parametrize("n", [18, 20])
def test_fp64_gpu_vs_fp64_oracle(ctx, oracle, n):
    m = 3
    recs = M.generate_sk(n, 1, 11 + n)
    _say("FP64 GPU vs FP64 oracle n", n)
    p64, tt, s64 = ctx.solve(n, m, recs, samples=2, with_samples=True, precision=M.FP64)
    po, tto, spo = oracle.solve(n, m, recs, samples=2, precision=1, threads=2, with_samples=True)
    assert np.array_equal(tt, tto)
    # both FP64, summed in different orders; the probabilities are rounded to float32 at the end
    d = float(np.max(np.abs(s64 - spo)))
    print(f"[fp64] n={n} max|dp|={d:.3e}", flush=True)
    assert d < 1e-9 + 1e-7


def test_large_instance_n28_m5_fp32_vs_fp64(ctx):
    n, m = 28, 5
    rec = M.generate_sk(n, 1, 4)  # the large-instance workload's seed
    _say("n=28 m=5 SK, 1 sample, FP32 and FP64")
    p32, tt32, s32 = ctx.solve(n, m, rec, samples=1, with_samples=True)
    st = ctx.last_stats()
    p64, tt64, s64 = ctx.solve(n, m, rec, samples=1, with_samples=True, precision=M.FP64)
    print(f"[large] n=28 m=5 p32={float(s32[0, 0]):.9f} p64={float(s64[0, 0]):.9f} "
          f"applies={st['term_units']} device_ms={st['gpu_ms']:.1f}", flush=True)
    assert tt32[0] == tt64[0]
    assert 0.0 <= s32[0, 0] <= 1.0
    assert abs(float(s32[0, 0]) - float(s64[0, 0])) < PROB_TOL


def test_large_instance_n28_m5_j0_closed_form(ctx, oracle):
    n, m = 28, 5
    rng = np.random.default_rng(128)
    rec = np.zeros(n * (n + 1) // 2)
    rec[n * (n - 1) // 2:] = rng.uniform(0.4, 1.6, n) * rng.choice([-1.0, 1.0], n)
    J, h = oracle.unpack(n, rec)
    sch = oracle.schedule(n, m, J, h)
    times = oracle.sample_times(m, 1, sch["short_t"])
    bits = [1 if hi < 0 else 0 for hi in h]
    _say("n=28 m=5 J=0 closed form")
    _, tt, sp = ctx.solve(n, m, rec[None, :], samples=1, with_samples=True)
    assert tt[0] == sch["total_t"]
    ref = D.product_probability(h.astype(np.float64), sch["gamma"].astype(np.float64), times[:, 0], bits)
    print(f"[large] n=28 m=5 J=0 p={float(sp[0, 0]):.9f} closed form={ref:.9f}", flush=True)
    assert abs(sp[0, 0] - ref) < PROB_TOL
