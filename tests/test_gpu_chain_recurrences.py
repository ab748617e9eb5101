"""GPU: the shared bases' recurrences with their state on chip (chain_recurrences).

One launch runs every step of a Chebyshev chain s_{k+1} = 2 Ht s_k - s_{k-1}
(the recurrence the reference's Clenshaw loop evaluates backwards,
MultiQW.cpp:100-173) on a group of workgroups of one XCD; the partners of the
top n - 15 bits come from the other members' outputs.  Its sums run in the
order of the per-step kernel with 2^14-amplitude tiles, so with tile_bits=14
every probability and every sample is bit-identical to one launch per step:
first-stage bases (m = 1 direct evaluation, m = 2), second-stage chains from
the compressed basis (m = 3), the two-stage adjoint's chains (m = 4) and the
third stage's (m = 5), at n = 15 (one workgroup per chain), 16, 18 and 20 (a
whole XCD per chain).  And against the oracle at the bench's size.
"""
import os

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,count,samples", [
    (15, 3, 3, 64),
    (16, 1, 3, 40),
    (16, 2, 3, 40),
    (16, 4, 2, 64),
    (16, 5, 2, 100),
    (18, 3, 2, 64),
    (20, 3, 2, 100),
    (20, 5, 1, 100),
])
def test_chains_bit_identical_to_per_step(ctx, n, m, count, samples):
    recs = M.generate_sk(n, count, 900 + n + m)
    _say("chains n", n, "m", m, "instances", count, "samples", samples)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, tile_bits=14)
    st = ctx.last_stats()
    p2, tt2, sp2 = ctx.solve(n, m, recs, samples=samples, with_samples=True, tile_bits=14, chain_recurrences=False)
    st2 = ctx.last_stats()
    print(f"[chain] n={n} m={m}: {st['chain_launches']} launches, {st['chain_steps']} steps, basis "
          f"{st['basis_ms']:.2f} ms against {st2['basis_ms']:.2f} ms per step "
          f"({st2['basis_launches']} launches)", flush=True)
    assert st["chain_launches"] > 0 and st2["chain_launches"] == 0
    assert np.array_equal(tt, tt2)
    assert np.array_equal(sp, sp2)
    assert np.array_equal(p, p2)


def test_chains_headline_size_vs_oracle(ctx, oracle):
    """n = 20, m = 3, 100 samples with the tuned configuration (on-chip recurrences, no term sweep)."""
    n, m, samples = 20, 3, 100
    recs = M.generate_sk(n, 2, 2)
    _say("chains n=20 m=3 against the oracle")
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True)
    st = ctx.last_stats()
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=THREADS, with_samples=True, norm="f64")
    d = np.max(np.abs(sp - spo))
    print(f"[chain] n=20 m=3: {st['chain_launches']} launches, sweeps {st['sweeps']}, max|dp| vs oracle {d:.3e}",
          flush=True)
    assert st["chain_launches"] > 0 and st["sweeps"] == 0
    assert np.array_equal(tt, tto)
    assert d < 1e-6


def test_chains_several_batches_and_uneven(ctx, oracle):
    """Batches of bases one after the other (a device cap), and an uneven batch across n = 14..17
    (n = 14 runs one launch per step) with schedule overrides, against the oracle."""
    n, m = 16, 3
    recs = M.generate_sk(n, 4, 1234)
    _say("chains over several batches")
    p, tt, sp = ctx.solve(n, m, recs, samples=64, with_samples=True, max_device_bytes=2 << 30)
    st = ctx.last_stats()
    p2, tt2, sp2 = ctx.solve(n, m, recs, samples=64, with_samples=True, max_device_bytes=2 << 30,
                             chain_recurrences=False, tile_bits=14)
    assert st["basis_batches"] >= 2 and st["chain_launches"] >= 2
    assert np.array_equal(sp, sp2)
    inst = []
    rng = np.random.default_rng(41)
    for k, (nn, mm) in enumerate([(14, 3), (15, 2), (16, 4), (17, 1), (16, 5), (15, 3)]):
        rec = M.generate_sk(nn, 1, 4100 + k)[0]
        s = M.schedule_heuristic(nn, mm, rec)
        g = (s.gamma * rng.uniform(0.7, 1.3, mm)).astype(np.float32)
        inst.append((nn, mm, rec, g, s.short_t))
    _say("chains in an uneven batch")
    p, tt, sp = ctx.solve_batch(inst, samples=64, with_samples=True)
    assert ctx.last_stats()["chain_launches"] > 0
    for k, (nn, mm, rec, g, t) in enumerate(inst):
        po, tto, spo = oracle.solve(nn, mm, rec[None, :], samples=64, threads=THREADS, with_samples=True,
                                    gamma=g[None, :], short_t=t[None, :], norm="f64")
        assert tt[k] == tto[0]
        assert np.max(np.abs(sp[k] - spo[0])) < 1e-6, (nn, mm)


def test_chains_fall_back_when_a_wait_runs_out(ctx):
    """The fallback, forced from the host: with no read per wait (chain_wait_polls=-1) the first wait
    between a group's workgroups gives up, the launch ends without waiting (every access inside its
    regions) and reports it; the context then runs the recurrences one launch per step, bit for bit
    those of chain_recurrences=False, and keeps doing so."""
    n, m, samples = 16, 3, 64
    recs = M.generate_sk(n, 2, 1777)
    _say("chains: forced fallback")
    c = M.Context(0)
    try:
        p, tt, sp = c.solve(n, m, recs, samples=samples, with_samples=True, tile_bits=14, chain_wait_polls=-1)
        st = c.last_stats()
        p2, tt2, sp2 = c.solve(n, m, recs, samples=samples, with_samples=True, tile_bits=14)
        st2 = c.last_stats()
    finally:
        c.close()
    p3, tt3, sp3 = ctx.solve(n, m, recs, samples=samples, with_samples=True, tile_bits=14, chain_recurrences=False)
    print(f"[chain] forced fallback: {st['chain_launches']} and then {st2['chain_launches']} on-chip launches, "
          f"{st['basis_launches']} basis launches", flush=True)
    assert st["chain_launches"] == 0 and st2["chain_launches"] == 0
    assert st["basis_launches"] > 2  # one per step
    assert np.array_equal(sp, sp3) and np.array_equal(sp2, sp3)
    assert np.array_equal(tt, tt3) and np.array_equal(p, p3)


def test_chain_wait_polls_below_minus_one_is_an_error(ctx):
    recs = M.generate_sk(16, 1, 5)
    with pytest.raises(M.MQWError) as e:
        ctx.solve(16, 3, recs, samples=8, chain_wait_polls=-2)
    assert e.value.status == 1  # MQW_EINVAL


@pytest.mark.parametrize("n,m,count,cap,compress", [
    (16, 5, 2, 0, True),          # the compression's launch aborts; the third stage after the checks
    (16, 3, 4, 2 << 30, True),    # the same over several batches
    (16, 2, 2, 0, False),         # no compression: the batch's launch (checked after the host has
    (16, 2, 24, 1 << 30, False),  # built its plan) aborts, in one batch and in the first of several
])
def test_forced_fallback_third_stage_and_batches(ctx, n, m, count, cap, compress):
    """The fallback forced from the host (no read per wait): where the first on-chip launch is the
    compression's (checked at once), and where it is a batch's, whose error word is read after the
    host has built the batch's plan; before the compressed third stage, and over several batches of
    bases.  Bit for bit the results of chain_recurrences=False."""
    recs = M.generate_sk(n, count, 1900 + m)
    kw = dict(samples=100, with_samples=True, tile_bits=14, compress_bases=compress)
    if cap:
        kw["max_device_bytes"] = cap
    _say("forced fallback n", n, "m", m, "batches cap", cap)
    c = M.Context(0)
    try:
        p, tt, sp = c.solve(n, m, recs, chain_wait_polls=-1, **kw)
        st = c.last_stats()
    finally:
        c.close()
    p2, tt2, sp2 = ctx.solve(n, m, recs, chain_recurrences=False, **kw)
    st2 = ctx.last_stats()
    print(f"[chain] forced fallback n={n} m={m}: batches {st['basis_batches']}, third stages "
          f"{st['third_stage_instances']}, on-chip launches {st['chain_launches']}", flush=True)
    assert st["chain_launches"] == 0
    if m == 5:
        assert st["third_stage_instances"] == count and st2["third_stage_instances"] == count
    if cap:
        assert st["basis_batches"] >= 2
    assert np.array_equal(tt, tt2)
    assert np.array_equal(sp, sp2) and np.array_equal(p, p2)
