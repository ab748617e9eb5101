#!/usr/bin/env python3
"""Benchmark of the multi-stage quantum-walk hot path on one MI355X.

Workload (the headline configuration): n = 20 spins, m = 3 stages, 256 seeded
SK instances (J_ij, h_i ~ N(0,1), seed 2), 100 Monte-Carlo samples each, FP32
state, FP64 energies / coefficients / probability sums.  One step solves a
slice of --per-step instances of that file end to end (energies, schedule,
Chebyshev coefficients, every Clenshaw sweep, renormalisation, readout), with
the input records already in host memory.  `value` is instances/s over the
timed steps.

Also reported: the n = 24 apply rate (H|psi> applies/s and algorithmic GB/s of
one n = 24, m = 3 instance), the roofline of the dominant kernel (the fused
Clenshaw term kernel) and a CPU baseline: the FP32 restatement of MultiQW.cpp
(oracle/) on the host cores for a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--per-step", type=int, default=16, help="instances solved per step")
    ap.add_argument("--n", type=int, default=20)
    ap.add_argument("--m", type=int, default=3)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--apply-n", type=int, default=24, help="n of the apply-rate measurement (0 = skip)")
    ap.add_argument("--apply-samples", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores); -1 = skip")
    ap.add_argument("--cpu-samples", type=int, default=8)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}). The (instance, sample) "
              "trajectories are independent and would shard across GPUs with no exchange; this benchmark "
              "measures one GPU's throughput.", file=sys.stderr)
        return 2

    import multistageqw_amd as M

    n, m = args.n, args.m
    recs = M.generate_sk(n, args.instances, 2)
    ctx = M.Context(0)

    def step(k):
        lo = (k * args.per_step) % args.instances
        idx = [(lo + i) % args.instances for i in range(args.per_step)]
        return ctx.solve(n, m, recs[idx])

    for k in range(args.warmup):
        step(k)
    agg = dict(amp_terms=0, gpu_ms=0.0, sweep_ms=0.0, sweeps=0, term_units=0, alg=0.0, req=0.0, aux_ms=0.0,
               aux_bytes=0.0, aux_launches=0)
    t0 = time.perf_counter()
    probs = []
    for k in range(args.steps):
        p, _ = step(args.warmup + k)
        st = ctx.last_stats()
        agg["amp_terms"] += st["amp_terms"]
        agg["gpu_ms"] += st["gpu_ms"]
        agg["sweeps"] += st["sweeps"]
        agg["term_units"] += st["term_units"]
        agg["alg"] += st["algorithmic_bytes"]
        agg["req"] += st["required_bytes"]
        for key in ("sweep_ms", "aux_ms", "aux_bytes", "aux_launches"):
            agg[key] += st[key]
        probs.append(p)
    wall = time.perf_counter() - t0
    inst = args.steps * args.per_step
    value = inst / wall if wall > 0 else 0.0
    # Roofline of the fused term kernel: algorithmic bytes (36 B per amplitude
    # per complex H-apply) over the event-timed device time of the term sweeps
    # (term launches and the small per-group reduce launches).
    achieved = agg["alg"] / (agg["sweep_ms"] * 1e-3) / 1e9 if agg["sweep_ms"] > 0 else 0.0
    # The same without the psi reads of terms whose psi is the constant |+>.
    achieved_req = agg["req"] / (agg["sweep_ms"] * 1e-3) / 1e9 if agg["sweep_ms"] > 0 else 0.0
    per_launch_bytes = agg["alg"] / max(1, agg["sweeps"])
    per_launch_ms = agg["sweep_ms"] / max(1, agg["sweeps"])
    # Every propagation kernel together: the shared bases, combines and adjoint
    # readouts (aux) with the term sweeps, over all event-timed device time.
    achieved_all = (agg["alg"] + agg["aux_bytes"]) / (agg["gpu_ms"] * 1e-3) / 1e9 if agg["gpu_ms"] > 0 else 0.0
    # HBM traffic per launch: the calibrated FETCH_SIZE + WRITE_SIZE ratio to
    # algorithmic bytes measured on this workload (profiles/, rocprofv3 --pmc).
    traffic, traffic_src = None, None
    prof = os.path.join(ROOT, "profiles", "half_term_kernel_n20_m3_traffic.md")
    if os.path.exists(prof) and n == 20 and m == 3:
        for line in open(prof):
            if line.startswith("traffic_per_algorithmic_byte:"):
                traffic = float(line.split(":")[1]) * per_launch_bytes
                traffic_src = "profiles/half_term_kernel_n20_m3_traffic.md"

    apply = None
    if args.apply_n > 0:
        na = args.apply_n
        ra = M.generate_sk(na, 1, 24)
        ctx.solve(na, m, ra, samples=1)  # warm the workspace
        ta = time.perf_counter()
        ctx.solve(na, m, ra, samples=args.apply_samples)
        wa = time.perf_counter() - ta
        sa = ctx.last_stats()
        # complex H-applies of the term sweeps over their device time (the shared
        # bases, combines and readouts are timed apart: aux_ms)
        applies = sa["term_units"] / (sa["sweep_ms"] * 1e-3)
        gbs = applies * 36.0 * (1 << na) / 1e9
        apply = dict(n=na, m=m, samples=args.apply_samples, applies_per_s=applies, gbs=gbs,
                     frac=gbs / HBM_PEAK_GBS, device_ms=sa["gpu_ms"], aux_ms=sa["aux_ms"], wall_s=wa)

    cpu = None
    if args.cpu_threads >= 0:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        O.lib()
        tc = time.perf_counter()
        O.solve(n, m, recs[:threads], samples=args.cpu_samples, threads=threads)
        wc = time.perf_counter() - tc
        cpu = dict(value=threads * args.cpu_samples / 100.0 / wc, unit="instances/s", cores=threads, kind="port",
                   sample=f"{threads} of the n={n} m={m} instances x {args.cpu_samples} of 100 samples each "
                          f"(FP32 restatement of MultiQW.cpp, one problem per thread, {wc:.1f} s), "
                          "scaled linearly to 100 samples")

    line = {
        "metric": "batched instances/s at n=20,m=3 (+ H|psi> applies/s and HBM GB/s at n=24)",
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / max(1, args.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"SK n={n} m={m}, {args.instances} instances seed 2, 100 samples, "
                               f"{args.per_step} instances per step", "n": n, "m": m,
                   "instances_per_step": args.per_step, "samples": 100},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "achieved_required": achieved_req, "frac_required": achieved_req / HBM_PEAK_GBS,
                     "kernel": "fused Clenshaw sweep (half_term_kernel / term_kernel)",
                     "alg_bytes_per_launch": per_launch_bytes, "avg_launch_ms": per_launch_ms,
                     "launches": agg["sweeps"], "achieved_all_kernels": achieved_all,
                     "frac_all_kernels": achieved_all / HBM_PEAK_GBS, "device_ms": agg["gpu_ms"],
                     "sweep_ms": agg["sweep_ms"], "aux_ms": agg["aux_ms"], "aux_bytes": agg["aux_bytes"],
                     "aux_launches": agg["aux_launches"]},
        "cpu_baseline": cpu,
        "apply_rate": apply,
        "mean_prob": float(np.mean(np.concatenate(probs))) if probs else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
