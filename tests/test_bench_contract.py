"""bench.py's output: one JSON line with the fields the benchmark harness reads.

The CPU test checks that a request for several GPUs is refused with a message
(one GPU only); the GPU test runs a tiny bench (4 instances, one step, a
one-instance CPU baseline) and checks every field of its line.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=timeout)


def test_bench_refuses_several_gpus():
    r = _bench("--gpus", "2", timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert not r.stdout.strip()


@pytest.mark.gpu
def test_bench_line_fields():
    print("[case] bench.py tiny run", flush=True)
    r = _bench("--steps", "1", "--warmup", "0", "--per-step", "4", "--instances", "8", "--apply-n", "0",
               "--n", "14", "--cpu-threads", "1", "--cpu-instances", "1", "--cpu-core-samples", "1",
               "--uneven", "0")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for key, typ in [("metric", str), ("value", float), ("unit", str), ("n_gpus", int), ("steps", int),
                     ("warmup", int), ("ms_per_step", float), ("higher_is_better", bool), ("scaling", str),
                     ("dtype", str), ("data", str), ("config", dict), ("roofline", dict), ("cpu_baseline", dict)]:
        assert isinstance(d[key], typ), (key, d.get(key))
    assert "vs_baseline" in d
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 0
    assert d["value"] > 0 and d["higher_is_better"] is True
    assert "workload" in d["config"] and "model" not in d["config"]
    rl = d["roofline"]
    assert rl["bound"] in ("hbm", "mfma") and rl["unit"] in ("GB/s", "TFLOP/s")
    assert rl["peak"] > 0 and 0 < rl["achieved"] <= rl["peak"]
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-9
    assert rl["traffic"] is None or rl["traffic"] > 0
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] == 1 and cb["kind"] in ("port", "reference")
    assert isinstance(cb["unit"], str) and isinstance(cb["sample"], str)
    assert cb["one_core"]["value"] > 0 and cb["cpus_available"] >= 1 and isinstance(cb["cpu_model"], str)
    assert "uneven" in d and d["uneven"] is None
