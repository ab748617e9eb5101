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


def _host_cpu() -> dict:
    """CPU model, sockets and logical CPUs of the host (from /proc/cpuinfo)."""
    model, sockets = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return dict(cpu_model=model, sockets=len(sockets) or None, host_logical_cpus=os.cpu_count())


def cpu_baseline(args, n, m, recs) -> dict:
    """The FP32 restatement of MultiQW.cpp (oracle/mqw_oracle.c) built as the reference
    is built (gcc -O3 -march=native, on this host), timed on this host's CPUs: whole
    instances at all 100 samples across every CPU this process may use (capped at 16, the
    host's CPU share per GPU), and one core on a few samples of one instance."""
    import tempfile
    from oracle import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = args.cpu_threads or min(16, avail)
    with tempfile.TemporaryDirectory() as d:
        L = O.load_native(O.build_native(d))
        cnt = args.cpu_instances
        tc = time.perf_counter()
        O.solve_native(L, n, m, recs[:cnt], samples=100, threads=threads)
        wc = time.perf_counter() - tc
        t1 = time.perf_counter()
        O.solve_native(L, n, m, recs[:1], samples=args.cpu_core_samples, threads=1)
        w1 = time.perf_counter() - t1
    one_core = args.cpu_core_samples / 100.0 / w1
    return dict(value=cnt / wc, unit="instances/s", cores=threads, kind="port",
                sample=f"{cnt} whole n={n} m={m} instances (100 samples each) in {wc:.1f} s on {threads} threads "
                       f"(samples of an instance split across threads); FP32 restatement of MultiQW.cpp built "
                       f"gcc -O3 -march=native on this host",
                one_core=dict(value=one_core, unit="instances/s",
                              sample=f"{args.cpu_core_samples} of 100 samples of one instance on 1 thread "
                                     f"({w1:.1f} s), scaled to 100 samples"),
                cpus_available=avail, **_host_cpu())


TUNE_CACHE = os.path.join(ROOT, "multistageqw_amd", "tune", "gfx950.txt")


def uneven_leg(ctx, args) -> dict:
    """The uneven workload: instances with n ~ U{12..22}, m ~ U{1..5} (seed 3, the first
    args.uneven of the 4096-instance set), 100 samples, solved in chunks of 256 instances per
    mqw_solve_batch call with the autotuner's configuration for every n (read from the shipped
    tuning table; measured where it has no line), after one untimed chunk that allocates the
    workspace.  Every (uneven // uneven_check)-th instance is also run at 8 samples in FP32 and
    with an all-FP64 state, for the per-sample error."""
    import multistageqw_amd as M
    rng = np.random.default_rng(3)
    inst = []
    for k in range(args.uneven):
        nn = int(rng.integers(12, 23))
        mm = int(rng.integers(1, 6))
        inst.append((nn, mm, M.generate_sk(nn, 1, 3 * 100003 + k)[0]))
    ns = sorted({i[0] for i in inst})
    tt0 = time.perf_counter()
    tuned = {nn: ctx.tune(nn, TUNE_CACHE) for nn in ns}
    tune_s = time.perf_counter() - tt0
    keys = ("gpu_ms", "sweep_ms", "algorithmic_bytes", "aux_bytes", "basis_ms", "gemm_ms", "lds_trajectories")
    tot = {k: 0.0 for k in keys}
    # warm-up: the first chunk, untimed, so the grow-only device workspace (up to
    # two basis buffers of about 64 GiB at n = 22) is allocated before the timing
    ctx.solve_batch(inst[:256], samples=100)
    wall = 0.0
    for c0 in range(0, len(inst), 256):
        t0 = time.perf_counter()
        ctx.solve_batch(inst[c0:c0 + 256], samples=100)
        wall += time.perf_counter() - t0
        st = ctx.last_stats()
        for k in keys:
            tot[k] += st[k]
    out = dict(workload=f"{len(inst)} instances, n in 12..22, m in 1..5, seed 3, 100 samples, 256 per call",
               instances_per_s=len(inst) / wall, wall_s=wall,
               gbs=(tot["algorithmic_bytes"] + tot["aux_bytes"]) / wall / 1e9,
               term_sweep_gbs=tot["algorithmic_bytes"] / (tot["sweep_ms"] * 1e-3) / 1e9 if tot["sweep_ms"] else 0.0,
               device_ms=tot["gpu_ms"], term_sweep_ms=tot["sweep_ms"], basis_ms=tot["basis_ms"],
               gemm_ms=tot["gemm_ms"], lds_trajectories=int(tot["lds_trajectories"]), tune_s=tune_s,
               tuned={str(k): v for k, v in tuned.items()})
    if args.uneven_check > 0:
        sub = inst[::max(1, len(inst) // args.uneven_check)][:args.uneven_check]
        _, _, s32 = ctx.solve_batch(sub, samples=8, with_samples=True)
        _, _, s64 = ctx.solve_batch(sub, samples=8, with_samples=True, precision=M.FP64)
        d = np.abs(s32 - s64)
        out.update(fp64_check=dict(instances=len(sub), samples=8, max_abs_err=float(d.max()),
                                   mean_abs_err=float(d.mean())))
    return out


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
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = every CPU this process may use, at most 16 (the host's share per GPU); -1 = skip")
    ap.add_argument("--cpu-instances", type=int, default=2, help="whole instances (100 samples) the CPU times")
    ap.add_argument("--cpu-core-samples", type=int, default=4, help="samples of the 1-core measurement")
    ap.add_argument("--uneven", type=int, default=512,
                    help="instances of the uneven workload (n in 12..22, m in 1..5, seed 3) timed after the "
                         "headline steps (0 = skip)")
    ap.add_argument("--uneven-check", type=int, default=16, help="of those, instances also run in FP64")
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
            "gemm_flops", "gemm_useful_flops", "gemm_launches", "gemm_bytes", "chain_launches", "chain_steps")
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
        #  - basis recurrences: on chip (chain_kernel, one launch per batch of chains: each output
        #    written, the diagonal and the seed read once per chain: 4 B per vector and amplitude
        #    and 4-8 B per chain and amplitude), or the term kernel on real halves, one launch per
        #    step (16 B per vector and amplitude: s_k, s_{k-1}, the diagonal read, s_{k+1} written);
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
            "basis": (dict(kernel="chain_kernel (on-chip recurrences, a group of one XCD per chain)", bound="hbm",
                           unit="GB/s", peak=HBM_PEAK_GBS, ms=agg["basis_ms"], launches=int(agg["basis_launches"]),
                           work=agg["basis_bytes"], achieved=rate(agg["basis_bytes"], agg["basis_ms"]),
                           profile="chain_kernel_n20_m3_traffic.md")
                      if agg["chain_launches"] > 0 else
                      dict(kernel="half_term_kernel (mode 3, real halves)", bound="hbm", unit="GB/s",
                           peak=HBM_PEAK_GBS, ms=agg["basis_ms"], launches=int(agg["basis_launches"]),
                           work=agg["basis_bytes"], achieved=rate(agg["basis_bytes"], agg["basis_ms"]),
                           profile="basis_recurrence_n20_m3_traffic.md")),
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
        # the counted HBM traffic per launch over the same launch time: what the
        # memory system moved, where `achieved` counts the algorithm's bytes
        if traffic is not None and dom["unit"] == "GB/s" and roofline["avg_launch_ms"] > 0:
            roofline["traffic_gbs"] = traffic / (roofline["avg_launch_ms"] * 1e-3) / 1e9
            roofline["traffic_frac"] = roofline["traffic_gbs"] / dom["peak"]
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

    uneven = uneven_leg(ctx, args) if args.uneven > 0 else None

    cpu = None
    if args.cpu_threads >= 0:
        cpu = cpu_baseline(args, n, m, recs)

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
        "uneven": uneven,
        "mean_prob": float(np.mean(np.concatenate(probs))) if probs else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
