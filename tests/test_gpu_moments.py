"""Moments of the diagonal energies (the variance MultiQW.cpp:375 has in closed
form, the skewness and kurtosis :379-380 leaves commented out), on the GPU."""
import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_moments_vs_brute_force_and_closed_form(ctx, oracle):
    n = 12
    _say("diagonal moments n=12")
    rec = M.generate_sk(n, 1, 81)[0]
    got = ctx.diag_moments(n, rec)
    J, h = oracle.unpack(n, rec)
    hp, _, info = oracle.energy_compat(n, J, h)
    x = hp.astype(np.float64) + float(np.float32(info["shift"]))
    mu = x.mean()
    d = x - mu
    var = (d * d).mean()
    assert abs(got["mean"] - mu) < 1e-9 * np.sqrt(var)
    assert abs(got["variance"] - var) < 1e-12 * var
    assert abs(got["skewness"] - (d ** 3).mean() / var ** 1.5) < 1e-9
    assert abs(got["excess_kurtosis"] - ((d ** 4).mean() / var ** 2 - 3)) < 1e-9
    # mean_k E(k) = 0 and mean_k E(k)^2 = HP2 = sum_pairs t^2 + sum h^2 (exact energies)
    sch = oracle.schedule(n, 1, J, h)
    assert abs(got["mean"]) < 1e-5 * np.sqrt(var)
    assert abs(got["variance"] / float(sch["hp2"]) - 1) < 1e-5


def test_moments_j0_closed_form_at_n20(ctx):
    """J = 0: E = sum_i h_i s_i with independent signs, skewness 0 and excess
    kurtosis -2 sum h^4 / (sum h^2)^2 at any n."""
    n = 20
    _say("diagonal moments J=0 n=20")
    rng = np.random.default_rng(82)
    rec = np.zeros(n * (n + 1) // 2)
    rec[n * (n - 1) // 2:] = rng.uniform(0.4, 1.6, n) * rng.choice([-1.0, 1.0], n)
    got = ctx.diag_moments(n, rec)
    h = rec[n * (n - 1) // 2:].astype(np.float32).astype(np.float64)
    assert abs(got["skewness"]) < 1e-6
    assert abs(got["excess_kurtosis"] - (-2 * np.sum(h ** 4) / np.sum(h ** 2) ** 2)) < 1e-6
    assert abs(got["variance"] / np.sum(h ** 2) - 1) < 1e-6


def test_moment_corrected_schedule_end_to_end(ctx, oracle):
    """Diagonal moments on the GPU -> Cornish-Fisher schedule -> solve with it as an override,
    against the dense FP64 propagator with the same schedule and Monte-Carlo times."""
    from oracle import dense as D
    n, m, samples = 10, 3, 6
    rec = M.generate_sk(n, 1, 995)[0]
    mo = ctx.diag_moments(n, rec)
    sc = M.schedule_heuristic(n, m, rec, skewness=mo["skewness"], excess_kurtosis=mo["excess_kurtosis"])
    sb = M.schedule_heuristic(n, m, rec)
    assert sc.a != sb.a
    _, _, sp = ctx.solve_batch([(n, m, rec, sc.gamma, sc.short_t)], samples=samples, with_samples=True)
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    times = oracle.sample_times(m, samples, sc.short_t)
    ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, sc.gamma.astype(np.float64), times, info["eloc"])
    assert np.max(np.abs(sp[0] - ref)) < 1e-5
