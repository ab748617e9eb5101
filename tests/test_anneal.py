"""Time-dependent anneals compiled to stages (multistageqw_amd.anneal) and the
fixed-times option of the simulator.

The reference has no time-dependent integrator (its README.md lists it as a
to-do), so the pins are mathematical: the exact propagator of the
time-dependent Schroedinger equation dpsi/dt = +i H(t) psi (the sign of
MultiQW.cpp:92,206-217) from scipy's DOP853 at tight tolerance, the
4th-order convergence that fixes the integrator's coefficients and order, and
the dense FP64 stage product for the GPU path."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import multistageqw_amd as M
from multistageqw_amd import anneal as A
from oracle import dense as D


def _say(*a):
    print("[case]", *a, flush=True)


def _exact(E, masks_n, gamma_of_t, T):
    HG = D.hypercube(masks_n)
    N = E.size
    f = lambda t, y: 1j * (E * y - gamma_of_t(t) * (HG @ y))
    r = solve_ivp(f, [0.0, T], np.full(N, 1 / np.sqrt(N), np.complex128), method="DOP853", rtol=1e-12, atol=1e-13)
    assert r.success
    return r.y[:, -1]


def _stage_product(E, gammas, times, order=None):
    psi = np.full(E.size, 1 / np.sqrt(E.size), np.complex128)
    for i in (order if order is not None else range(len(gammas))):
        psi = D.evolve(E, float(gammas[i]), float(times[i]), psi)
    return psi


def test_cfet4_is_fourth_order():
    """Error against the exact propagator falls as h^4; the midpoint rule and the
    CFET4 pair applied in the wrong order (a planted slip) fall only as h^2."""
    n, T = 6, 3.0
    E = np.random.default_rng(5).normal(size=1 << n) * 2.0
    g = lambda t: 2.0 * (1.0 - t / T) ** 2
    ex = _exact(E, n, g, T)

    def err(stages, steps, swap=False):
        gs, ts = stages(g, T, steps)
        order = [k ^ 1 for k in range(len(gs))] if swap else None
        return np.linalg.norm(_stage_product(E, gs, ts, order) - ex)

    c = [err(A.cfet4_stages, s) for s in (8, 32)]
    mid = [err(A.midpoint_stages, s) for s in (8, 32)]
    sw = [err(A.cfet4_stages, s, swap=True) for s in (8, 32)]
    assert c[0] / c[1] > 200          # 4^4 = 256 (measured ~1000 in the pre-asymptotic range)
    assert 8 < mid[0] / mid[1] < 30   # 4^2 = 16
    assert sw[0] / sw[1] < 30
    assert c[1] < 1e-4


def test_cfet4_stage_algebra():
    """Two stages per step, each h/2 long, whose gammas average to the Gauss-rule mean."""
    g = A.linear_anneal(1.5, 0.1, 2.0)
    gs, ts = A.cfet4_stages(g, 2.0, 5)
    assert gs.shape == ts.shape == (10,)
    assert np.allclose(ts, 0.2)
    h = 0.4
    for k in range(5):
        mean = 0.5 * (g(k * h + A.C1 * h) + g(k * h + A.C2 * h))
        assert np.isclose(0.5 * (gs[2 * k] + gs[2 * k + 1]), mean, rtol=1e-6)
        assert gs[2 * k] > gs[2 * k + 1]          # the earlier Gauss point weighs more first
    with pytest.raises(ValueError):
        A.cfet4_stages(g, 2.0, 33)                # 66 stages exceed the simulator's 64


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(M.FP32, 1e-5), (M.FP64, 1e-6)])
@pytest.mark.parametrize("small_n_max", [13, 0])
def test_anneal_on_gpu_vs_dense(ctx, oracle, precision, tol, small_n_max):
    """CFET4 stages with fixed times on the GPU against the dense stage product and
    the exact time-dependent propagator (n = 8; LDS-resident and sweep paths)."""
    n, T, steps, samples = 8, 4.0, 16, 3
    _say("anneal vs dense precision", precision, "small_n_max", small_n_max)
    rec = M.generate_sk(n, 1, 91)[0]
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    E = e - (info["emax"] + info["e0"]) / 2
    g = lambda t: 1.2 * (1.0 - t / T) ** 2
    gs, ts = A.cfet4_stages(g, T, steps)
    p, tt, sp = ctx.solve_batch([(n, 2 * steps, rec, gs, ts)], samples=samples, with_samples=True,
                                precision=precision, fixed_times=True, small_n_max=small_n_max)
    assert np.isclose(tt[0], T, rtol=1e-6)
    assert np.max(sp[0]) - np.min(sp[0]) < tol            # every sample runs the same schedule
    dense = abs(_stage_product(E, gs, ts)[info["eloc"]]) ** 2
    assert abs(p[0] - dense) < tol
    exact = abs(_exact(E, n, g, T)[info["eloc"]]) ** 2
    assert abs(p[0] - exact) < 1e-4


@pytest.mark.gpu
def test_signed_gamma_override_vs_dense(ctx, oracle):
    """Override gammas of either sign (as CFET4 makes when gamma(t) crosses zero)."""
    n, samples = 10, 2
    _say("signed gamma override n=10")
    rec = M.generate_sk(n, 1, 92)[0]
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    E = e - (info["emax"] + info["e0"]) / 2
    gs = np.array([0.7, -0.3, -0.9, 0.2], np.float32)
    ts = np.array([0.6, 0.9, 0.4, 1.3], np.float32)
    for small in (13, 0):
        p, _, sp = ctx.solve_batch([(n, 4, rec, gs, ts)], samples=samples, with_samples=True, precision=M.FP64,
                                   fixed_times=True, small_n_max=small)
        dense = abs(_stage_product(E, gs, ts)[info["eloc"]]) ** 2
        assert np.max(np.abs(sp[0] - dense)) < 1e-6


@pytest.mark.gpu
def test_fixed_times_off_keeps_monte_carlo_draws(ctx, oracle):
    """Without fixed_times the stage times are the reference's MT19937 draws."""
    n, m, samples = 8, 3, 6
    _say("fixed_times off n=8")
    rec = M.generate_sk(n, 1, 93)[0]
    gs = np.array([0.9, 0.5, 0.2], np.float32)
    ts = np.array([0.5, 0.7, 0.9], np.float32)
    _, _, a = ctx.solve_batch([(n, m, rec, gs, ts)], samples=samples, with_samples=True, precision=M.FP64)
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    times = oracle.sample_times(m, samples, ts)
    ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, gs.astype(np.float64), times, info["eloc"])
    assert np.max(np.abs(a[0] - ref)) < 1e-6
    assert np.max(a[0]) - np.min(a[0]) > 1e-4
