#!/usr/bin/env python3
"""Benchmark of the multi-stage quantum-walk hot path on one MI355X.

Workload (the headline configuration): n = 20 spins, m = 3 stages, 256 seeded
SK instances (J_ij, h_i ~ N(0,1), seed 2), 100 Monte-Carlo samples each, FP32
state, FP64 energies / coefficients / probability sums.  One step solves a
slice of --per-step instances of that file end to end (energies, schedule,
Chebyshev coefficients, every Clenshaw sweep, renormalisation, readout), with
the input records already in host memory; by default the 4 timed steps of 64
solve each of the 256 instances once.  `value` is instances/s over the timed
steps.

Also reported: the n = 24 apply rate (H|psi> applies/s and algorithmic GB/s of
one n = 24, m = 3 instance), the roofline of the kernel class with the most
device time (term sweeps, shared-basis recurrences or the second-stage GEMM;
all three are listed) and a CPU baseline: the FP32 restatement of
MultiQW.cpp (oracle/) on the host cores for a bounded sample of the same
workload.

The batches of shared bases overlap (the next batch's recurrences run on a
second stream beside this batch's GEMM), so the kernels' event times in the
timed steps include time they share the GPU.  "roofline" is measured over the
timed steps as they ran, with "roofline.all_kernels" the algorithmic bytes of
every kernel over the timed wall clock; "roofline_isolated" repeats the
measurement on one extra step after the timed region with the batches one
after the other, where each kernel has the GPU to itself.
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
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
MFMA_FP16_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 matrix


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--per-step", type=int, default=64, help="instances solved per step")
    ap.add_argument("--n", type=int, default=20)
    ap.add_argument("--m", type=int, default=3)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--apply-n", type=int, default=24, help="n of the apply-rate measurement (0 = skip)")
    ap.add_argument("--apply-samples", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores); -1 = skip")
    ap.add_argument("--cpu-samples", type=int, default=8)
    ap.add_argument("--gemm", default="split_fp16", choices=("split_fp16", "fp32"),
                    help="arithmetic of the second-stage GEMM")
    ap.add_argument("--overlap", type=int, default=1, help="overlap batches of shared bases (0/1)")
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

    def step(k, overlap=bool(args.overlap)):
        lo = (k * args.per_step) % args.instances
        idx = [(lo + i) % args.instances for i in range(args.per_step)]
        return ctx.solve(n, m, recs[idx], gemm=args.gemm, overlap_bases=overlap)

    for k in range(args.warmup):
        step(k)
    keys = ("amp_terms", "gpu_ms", "sweep_ms", "sweeps", "term_units", "algorithmic_bytes", "required_bytes",
            "aux_ms", "aux_bytes", "aux_launches", "basis_ms", "basis_bytes", "basis_launches", "gemm_ms",
            "gemm_flops", "gemm_useful_flops", "gemm_launches", "gemm_bytes")
    agg = {key: 0.0 for key in keys}
    t0 = time.perf_counter()
    probs = []
    for k in range(args.steps):
        p, _ = step(args.warmup + k)
        st = ctx.last_stats()
        for key in keys:
            agg[key] += st[key]
        probs.append(p)
    wall = time.perf_counter() - t0
    inst = args.steps * args.per_step
    value = inst / wall if wall > 0 else 0.0

    def rate(x, ms, scale=1e9):
        return x / (ms * 1e-3) / scale if ms > 0 else 0.0

    def roofline_of(agg):
        # Each kernel class, event-timed on the stream it runs on:
        #  - term sweeps: algorithmic bytes 36 B per amplitude per complex H-apply;
        #  - basis recurrences (the same term kernel on real halves): 16 B per vector and amplitude;
        #  - second-stage GEMM: basis rows read once per tile and outputs written (HBM), and the
        #    matrix-core flops it executes (three FP16 products per FP32 product when split).
        split = args.gemm == "split_fp16"
        gemm_peak = MFMA_FP16_PEAK_TFLOPS if split else MFMA_FP32_PEAK_TFLOPS
        gemm_gbs = rate(agg["gemm_bytes"], agg["gemm_ms"])
        gemm_tf = rate(agg["gemm_flops"], agg["gemm_ms"], 1e12)
        gemm_hbm_bound = split  # the split GEMM executes 192 FP16 flops per basis byte: HBM-bound
        classes = {
            "term_sweep": dict(kernel="half_term_kernel (mode 0-2) / term_kernel", bound="hbm", unit="GB/s",
                               peak=HBM_PEAK_GBS, ms=agg["sweep_ms"], launches=int(agg["sweeps"]),
                               work=agg["algorithmic_bytes"],
                               achieved=rate(agg["algorithmic_bytes"], agg["sweep_ms"]),
                               profile="half_term_kernel_n20_m3_traffic.md"),
            "basis": dict(kernel="half_term_kernel (mode 3, real halves)", bound="hbm", unit="GB/s",
                          peak=HBM_PEAK_GBS, ms=agg["basis_ms"], launches=int(agg["basis_launches"]),
                          work=agg["basis_bytes"], achieved=rate(agg["basis_bytes"], agg["basis_ms"]),
                          profile="basis_recurrence_n20_m3_traffic.md"),
            "gemm": dict(kernel=("combine2h_kernel (v_mfma_f32_32x32x16_f16, split FP16 operands)" if split
                                 else "combine2_kernel (v_mfma_f32_32x32x2f32)"),
                         bound="hbm" if gemm_hbm_bound else "mfma", unit="GB/s" if gemm_hbm_bound else "TFLOP/s",
                         peak=HBM_PEAK_GBS if gemm_hbm_bound else gemm_peak, ms=agg["gemm_ms"],
                         launches=int(agg["gemm_launches"]),
                         work=agg["gemm_bytes"] if gemm_hbm_bound else agg["gemm_flops"],
                         achieved=gemm_gbs if gemm_hbm_bound else gemm_tf,
                         mfma_tflops=gemm_tf, mfma_peak=gemm_peak, mfma_frac=gemm_tf / gemm_peak,
                         useful_fp32_tflops=rate(agg["gemm_useful_flops"], agg["gemm_ms"], 1e12), hbm_gbs=gemm_gbs,
                         profile=("combine2h_kernel_n20_m3_traffic.md" if split
                                  else "combine2_kernel_n20_m3_traffic.md")),
        }
        other_ms = agg["gpu_ms"] - agg["sweep_ms"] - agg["basis_ms"] - agg["gemm_ms"]
        dom_name = max(classes, key=lambda c: classes[c]["ms"])
        dom = classes[dom_name]
        per_launch_work = dom["work"] / max(1, dom["launches"])
        # HBM bytes per launch of the dominant kernel: its calibrated FETCH_SIZE +
        # WRITE_SIZE per algorithmic byte (or per flop), measured with rocprofv3
        # --pmc on this workload (profiles/).
        traffic, traffic_src = None, None
        prof = os.path.join(ROOT, "profiles", dom["profile"])
        if os.path.exists(prof) and n == 20 and m == 3:
            for line in open(prof):
                if line.startswith("traffic_per_unit:"):
                    traffic = float(line.split(":")[1]) * per_launch_work
                    traffic_src = "profiles/" + dom["profile"]
        roofline = {"bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"], "unit": dom["unit"],
                    "frac": dom["achieved"] / dom["peak"], "traffic": traffic, "traffic_source": traffic_src,
                    "kernel": dom["kernel"], "class": dom_name, "work_per_launch": per_launch_work,
                    "avg_launch_ms": dom["ms"] / max(1, dom["launches"]), "launches": dom["launches"],
                    "device_ms": agg["gpu_ms"], "other_ms": other_ms,
                    "classes": {c: {key: v for key, v in d.items() if key not in ("profile",)}
                                for c, d in classes.items()}}
        for c in roofline["classes"].values():
            c["frac"] = c["achieved"] / c["peak"] if c["peak"] else 0.0
        return roofline

    roofline = roofline_of(agg)
    # every kernel's algorithmic bytes (term sweeps, recurrences, GEMM, readout,
    # combines) over the timed wall clock: the device as a whole, both streams
    all_bytes = agg["algorithmic_bytes"] + agg["aux_bytes"]
    roofline["all_kernels"] = {"algorithmic_bytes": all_bytes, "wall_s": wall,
                               "gbs": all_bytes / wall / 1e9 if wall > 0 else 0.0,
                               "frac": all_bytes / wall / 1e9 / HBM_PEAK_GBS if wall > 0 else 0.0}
    # one extra step, not timed, with the batches of bases one after the other
    step(args.warmup + args.steps, overlap=False)
    st = ctx.last_stats()
    roofline_iso = roofline_of({key: float(st[key]) for key in keys})
    roofline_iso["step"] = "one step after the timed region, batches of bases not overlapped"

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
                   "instances_per_step": args.per_step, "samples": 100, "second_stage_gemm": args.gemm,
                   "overlap_bases": bool(args.overlap)},
        "roofline": roofline,
        "roofline_isolated": roofline_iso,
        "cpu_baseline": cpu,
        "apply_rate": apply,
        "mean_prob": float(np.mean(np.concatenate(probs))) if probs else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
