"""Measurement studies beyond the bench line.

  python -m multistageqw_amd.studies precision --n 24 --m 3 --samples 8
      FP32 state/recurrence against an all-FP64 GPU run of the same instance:
      per-sample ground-state probabilities, max |difference|.
  python -m multistageqw_amd.studies large --n 28 --m 5 --samples 2
      one large instance: complex H|psi> applies/s and algorithmic GB/s.
  python -m multistageqw_amd.studies sweep --ns 14 16 18 20 22 24 26
      apply rate of the sweep kernel per n (one instance, a few samples).
  python -m multistageqw_amd.studies batch --n 10 --m 2 --count 64 --samples 100 --seed 1
      one (n, m) batch (the n=10 case runs on the LDS-resident kernel), instances/s.
  python -m multistageqw_amd.studies uneven --count 4096 --seed 3 --fp64-every 16 --tune-cache tune.txt
      the uneven workload (n uniform in 12..22, m uniform in 1..5): instances/s,
      bytes-weighted GB/s, per-instance error against FP64 runs of a subset.
  python -m multistageqw_amd.studies moments --n 20 --m 3 --count 256 --seed 6 --samples 100
      the heuristic schedule against the same heuristic with a Cornish-Fisher
      ground-state depth from the diagonal's skewness and kurtosis: mean
      ground-state probability of each, and how often the corrected one wins.

Each prints JSON lines.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import FP32, FP64, Context, generate_sk, schedule_heuristic

PEAK = 8000.0


def _apply_line(ctx, n, wall, **extra):
    st = ctx.last_stats()
    # complex H-applies of the term sweeps over their own device time; the
    # shared bases, combines and readouts are reported as aux_ms
    t = st["sweep_ms"] if st["sweep_ms"] > 0 else st["gpu_ms"]
    applies = st["term_units"] / (t * 1e-3)
    bpa = 36.0 if extra.get("precision", "fp32") == "fp32" else 72.0
    gbs = applies * bpa * (1 << n) / 1e9
    return dict(n=n, applies_per_s=applies, algorithmic_GBps=gbs, frac=gbs / PEAK, device_ms=st["gpu_ms"],
                aux_ms=st["aux_ms"], wall_s=wall, applies=st["term_units"], **extra)


def precision(ctx, n, m, samples, seed):
    rec = generate_sk(n, 1, seed)
    t0 = time.time()
    p32, _, s32 = ctx.solve(n, m, rec, samples=samples, precision=FP32, with_samples=True)
    r32 = _apply_line(ctx, n, time.time() - t0, precision="fp32")
    t0 = time.time()
    p64, _, s64 = ctx.solve(n, m, rec, samples=samples, precision=FP64, with_samples=True)
    r64 = _apply_line(ctx, n, time.time() - t0, precision="fp64")
    return dict(study="precision", n=n, m=m, samples=samples, seed=seed, p_fp32=float(p32[0]), p_fp64=float(p64[0]),
                max_abs_diff_per_sample=float(np.max(np.abs(s32 - s64))), abs_diff_mean=float(abs(p32[0] - p64[0])),
                fp32=r32, fp64=r64)


def large(ctx, n, m, samples, seed, warm=True, **opts):
    rec = generate_sk(n, 1, seed)
    if warm:
        ctx.solve(n, m, rec, samples=1, **opts)
    t0 = time.time()
    p, _ = ctx.solve(n, m, rec, samples=samples, **opts)
    return dict(study="large", m=m, samples=samples, p=float(p[0]), **_apply_line(ctx, n, time.time() - t0))


def batch(ctx, n, m, count, samples, seed, reps=3):
    """count seeded instances of one (n, m): instances/s over reps solves after a warm-up."""
    recs = generate_sk(n, count, seed)
    ctx.solve(n, m, recs, samples=samples)
    t0 = time.time()
    dev = alg = 0.0
    for _ in range(reps):
        p, _ = ctx.solve(n, m, recs, samples=samples)
        st = ctx.last_stats()
        dev += st["gpu_ms"]
        alg += st["algorithmic_bytes"] + st["aux_bytes"]
    wall = time.time() - t0
    return dict(study="batch", n=n, m=m, count=count, samples=samples, seed=seed, instances_per_s=reps * count / wall,
                wall_s=wall / reps, device_ms=dev / reps, algorithmic_GBps=alg / (dev * 1e-3) / 1e9,
                lds_trajectories=st["lds_trajectories"], mean_p=float(np.mean(p)))


def uneven(ctx, count, samples, seed, chunk=256, fp64_every=0, tune_cache=None):
    """The uneven workload: n ~ U{12..22}, m ~ U{1..5}, solved chunk by chunk
    (each chunk one mqw_solve_batch call); every fp64_every-th instance is
    also run with an FP64 state for the per-instance error."""
    rng = np.random.default_rng(seed)
    inst = []
    for k in range(count):
        n = int(rng.integers(12, 23))
        m = int(rng.integers(1, 6))
        inst.append((n, m, generate_sk(n, 1, seed * 100003 + k)[0]))
    tuned = {}
    if tune_cache is not None:
        for n in sorted({i[0] for i in inst}):
            tuned[n] = ctx.tune(n, tune_cache)
        print(json.dumps(dict(study="uneven", tuned=tuned)), flush=True)
    probs = np.zeros(count, np.float32)
    keys = ("gpu_ms", "sweep_ms", "algorithmic_bytes", "basis_ms", "basis_bytes", "gemm_ms", "gemm_flops",
            "lds_trajectories")
    tot = {k: 0.0 for k in keys}
    wall = 0.0
    for c0 in range(0, count, chunk):
        part = inst[c0:c0 + chunk]
        t0 = time.time()
        p, _ = ctx.solve_batch(part, samples=samples)
        wall += time.time() - t0
        st = ctx.last_stats()
        for k in keys:
            tot[k] += st[k]
        probs[c0:c0 + len(part)] = p
        print(json.dumps(dict(study="uneven", done=c0 + len(part), wall_s=round(wall, 2))), flush=True)

    def rate(x, ms, scale):
        return x / (ms * 1e-3) / scale if ms > 0 else 0.0

    out = dict(study="uneven", count=count, samples=samples, instances_per_s=count / wall, wall_s=wall,
               device_ms=tot["gpu_ms"], term_sweep_ms=tot["sweep_ms"],
               term_sweep_GBps=rate(tot["algorithmic_bytes"], tot["sweep_ms"], 1e9), basis_ms=tot["basis_ms"],
               basis_GBps=rate(tot["basis_bytes"], tot["basis_ms"], 1e9), gemm_ms=tot["gemm_ms"],
               gemm_TFLOPs=rate(tot["gemm_flops"], tot["gemm_ms"], 1e12),
               lds_trajectories=int(tot["lds_trajectories"]), mean_p=float(np.mean(probs)))
    if fp64_every:
        idx = list(range(0, count, fp64_every))
        t0 = time.time()
        p64, _ = ctx.solve_batch([inst[i] for i in idx], samples=samples, precision=FP64)
        d = np.abs(probs[idx] - p64)
        out.update(fp64_instances=len(idx), fp64_wall_s=time.time() - t0, max_abs_err_vs_fp64=float(d.max()),
                   mean_abs_err_vs_fp64=float(d.mean()))
    return out


def moments(ctx, n, m, count, samples, seed, chunk=64):
    """Mean ground-state probability under the reference heuristic and under the
    heuristic with the Cornish-Fisher depth (each instance's own diagonal
    moments, computed on the GPU), same Monte-Carlo draws."""
    recs = generate_sk(n, count, seed)
    base, corr, skews, kurts, depth = [], [], [], [], []
    t0 = time.perf_counter()
    for c0 in range(0, count, chunk):
        inst_b, inst_c = [], []
        for rec in recs[c0:c0 + chunk]:
            mo = ctx.diag_moments(n, rec)
            sb = schedule_heuristic(n, m, rec)
            sc = schedule_heuristic(n, m, rec, skewness=mo["skewness"], excess_kurtosis=mo["excess_kurtosis"])
            inst_b.append((n, m, rec, sb.gamma, sb.short_t))
            inst_c.append((n, m, rec, sc.gamma, sc.short_t))
            skews.append(mo["skewness"])
            kurts.append(mo["excess_kurtosis"])
            depth.append((float(sb.a), float(sc.a)))
        pb, _ = ctx.solve_batch(inst_b, samples=samples)
        pc, _ = ctx.solve_batch(inst_c, samples=samples)
        base.extend(pb.tolist())
        corr.extend(pc.tolist())
    base, corr = np.array(base), np.array(corr)
    d = np.array(depth)
    return dict(study="moments", n=n, m=m, count=count, samples=samples, seed=seed,
                mean_p_heuristic=float(base.mean()), mean_p_corrected=float(corr.mean()),
                ratio=float(corr.mean() / base.mean()), corrected_wins=float(np.mean(corr > base)),
                skewness_mean=float(np.mean(skews)), skewness_sd=float(np.std(skews)),
                excess_kurtosis_mean=float(np.mean(kurts)), a_heuristic=float(d[:, 0].mean()),
                a_corrected_mean=float(d[:, 1].mean()), a_corrected_sd=float(d[:, 1].std()),
                wall_s=time.perf_counter() - t0)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["precision", "large", "sweep", "uneven", "batch", "moments"])
    ap.add_argument("--n", type=int, default=24)
    ap.add_argument("--ns", type=int, nargs="*", default=[14, 16, 18, 20, 22, 24, 26])
    ap.add_argument("--m", type=int, default=3)
    ap.add_argument("--samples", type=int, default=8)
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--chunk", type=int, default=256)
    ap.add_argument("--fp64-every", type=int, default=0)
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--no-warm", action="store_true", help="large: skip the warm-up solve (one solve to profile)")
    ap.add_argument("--high-pass", type=int, default=-1, help="large/sweep: mqw_opts.high_pass (-1 by n)")
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.study == "precision":
        print(json.dumps(precision(ctx, a.n, a.m, a.samples, a.seed)), flush=True)
    elif a.study == "large":
        print(json.dumps(large(ctx, a.n, a.m, a.samples, a.seed, warm=not a.no_warm, high_pass=a.high_pass)),
              flush=True)
    elif a.study == "sweep":
        for n in a.ns:
            print(json.dumps(large(ctx, n, a.m, a.samples, a.seed, high_pass=a.high_pass)), flush=True)
    elif a.study == "batch":
        print(json.dumps(batch(ctx, a.n, a.m, a.count, a.samples, a.seed)), flush=True)
    elif a.study == "moments":
        print(json.dumps(moments(ctx, a.n, a.m, a.count, a.samples, a.seed)), flush=True)
    else:
        print(json.dumps(uneven(ctx, a.count, a.samples, a.seed, a.chunk, a.fp64_every, a.tune_cache)), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
