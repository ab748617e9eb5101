"""GPU: m == 1 instances evaluated directly (mqw_opts.direct_m1, the default).

The reference chains an m == 1 instance's samples: sorted stage times,
differenced, each sample propagated from the previous one by its increment
(MultiQW.cpp:426-430,447).  The direct path evaluates sample j as one series
exp(iH T_j)|+> = sum_k a_jk T_k(Ht)|+> at the summed time T_j, from the
instance's first-stage basis and its Gram matrix: no term sweep runs.  The
two differ by the chain's accumulated truncation (each step's tail, <= tol).
Checked per sample against the GPU chain (direct_m1 = False) and the oracle's
chain.
"""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,count", [(14, 4), (16, 3), (20, 2)])
def test_direct_m1_vs_chain_and_oracle(ctx, oracle, n, count):
    recs = M.generate_sk(n, count, 900 + n)
    _say("direct m=1 n", n)
    p, tt, sp = ctx.solve(n, 1, recs, samples=100, with_samples=True)
    st = ctx.last_stats()
    assert st["direct_chains"] == count and st["sweeps"] == 0
    pc, ttc, spc = ctx.solve(n, 1, recs, samples=100, with_samples=True, direct_m1=False)
    assert ctx.last_stats()["direct_chains"] == 0 and ctx.last_stats()["sweeps"] > 0
    po, tto, spo = oracle.solve(n, 1, recs, samples=100, threads=THREADS, with_samples=True, norm="f64")
    d_chain, d_or = float(np.max(np.abs(sp - spc))), float(np.max(np.abs(sp - spo)))
    print(f"[direct] n={n}: max|dp| vs GPU chain {d_chain:.3e}, vs oracle chain {d_or:.3e}, "
          f"chain vs oracle {np.max(np.abs(spc - spo)):.3e}", flush=True)
    assert np.array_equal(tt, ttc) and np.array_equal(tt, tto)
    assert d_chain < 1e-6 and d_or < 1e-6
    assert np.max(np.abs(p - po)) < 1e-6


def test_direct_m1_in_uneven_batch_with_override(ctx, oracle):
    """m = 1 instances of several n beside m > 1 ones, one with a schedule override."""
    inst = [(15, 1, M.generate_sk(15, 1, 951)[0]), (15, 3, M.generate_sk(15, 1, 952)[0]),
            (17, 1, M.generate_sk(17, 1, 953)[0], np.array([1.3], np.float32), np.array([0.7], np.float32)),
            (14, 2, M.generate_sk(14, 1, 954)[0])]
    _say("direct m=1 in an uneven batch")
    p, tt, sp = ctx.solve_batch(inst, samples=50, with_samples=True)
    assert ctx.last_stats()["direct_chains"] == 2
    for k, ins in enumerate(inst):
        n, m, rec = ins[:3]
        kw = dict(gamma=ins[3][None, :], short_t=ins[4][None, :]) if len(ins) > 3 else {}
        po, tto, spo = oracle.solve(n, m, rec[None, :], samples=50, threads=THREADS, with_samples=True,
                                    norm="f64", **kw)
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < 1e-6, (n, m)


def test_direct_m1_fixed_times_vs_chain(ctx):
    """fixed_times: every increment of the chain is exactly short_t, so sample j sits at (j+1) short_t
    (two samples keep the chain's series within the 128 terms of the direct path, as the
    Monte-Carlo times, whose chain ends below 2 short_t, do)."""
    n = 15
    recs = M.generate_sk(n, 2, 961)
    _say("direct m=1, fixed times")
    p, _, sp = ctx.solve(n, 1, recs, samples=2, with_samples=True, fixed_times=True)
    assert ctx.last_stats()["direct_chains"] == 2
    pc, _, spc = ctx.solve(n, 1, recs, samples=2, with_samples=True, fixed_times=True, direct_m1=False)
    assert np.max(np.abs(sp - spc)) < 1e-6
