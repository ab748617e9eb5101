"""GPU: what the shipped learned-schedule network gains over the heuristic.

The network in multistageqw_amd/weights/learned_schedule_m3.f32 was trained on
simulator-scored schedules of instances with seeds 1000003 + i (n = 12..18,
m = 3).  Here it proposes schedules for fresh n = 20, m = 3 SK instances
(seeds 7000021 + i, disjoint from the training and the recorded evaluation
seeds 2000006 + i), and both its proposals and the heuristic schedule of
MultiQW.cpp:386-423 are scored by the simulator at 100 samples, with the
same Monte-Carlo draws (MultiQW.cpp:387) and the same total evolution time.
The network's raw proposals are compared, without the fallback to the
heuristic, so the gain cannot come from taking the better of the two.
"""
import os

import numpy as np
import pytest

from multistageqw_amd import learned as Lr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "multistageqw_amd", "weights", "learned_schedule_m3.f32")
COUNT = 256
SEED = 7


def test_shipped_network_gain_on_fresh_instances(ctx):
    print(f"[case] shipped network, {COUNT} fresh n=20 m=3 instances, 100 samples", flush=True)
    params = np.fromfile(WEIGHTS, "<f4")
    mlp = Lr.MLP(ctx, seed=1)
    mlp.set_params(params)
    assert np.array_equal(mlp.params(), params)
    ev = Lr.evaluate(ctx, mlp, COUNT, seed=SEED, n=20, m=3, samples=100)
    mlp.close()
    print("[learned]", ev, flush=True)
    # every proposal ran for exactly the heuristic's total time (float32 sum of the stage times)
    assert ev["equal_time"]
    # the raw proposals, not max(heuristic, proposal): at least +30 % mean probability
    assert ev["gain_raw"] >= 0.30 * ev["p_heuristic"]
    assert ev["win_rate"] >= 0.90
    assert ev["p_mlp"] > ev["p_heuristic"]
