"""Time-dependent anneals as multi-stage schedules.

The reference evolves a piecewise-constant H = H_P - gamma_i * sum_j X_j, one
exponential per stage (MultiQW.cpp:432-451).  A continuous anneal gamma(t)
with H(t) = H_P - gamma(t) * H_M is integrated here by a commutator-free
exponential integrator (the 4th-order two-exponential scheme of Blanes & Moan,
Appl. Numer. Math. 56 (2006), also used by Alvermann & Fehske, J. Comput.
Phys. 230 (2011) as CFET4:2).  Each step of length h applies

    exp(i h (a2 H(t1) + a1 H(t2)))   then   exp(i h (a1 H(t1) + a2 H(t2)))

with Gauss points t1,2 = t0 + (1/2 -+ sqrt(3)/6) h and a1,2 = (3 -+ 2 sqrt(3))/12.
Because a1 + a2 = 1/2 and H is affine in gamma, each exponential is one stage
of the simulator: H_P - gamma' * H_M for time h/2, with
gamma' = 2 (a2 gamma(t1) + a1 gamma(t2)) and 2 (a1 gamma(t1) + a2 gamma(t2)).
The schedules run through ``Context.solve_batch`` with ``fixed_times=True``
so every stage lasts exactly its time (no Monte-Carlo draw).  gamma' may
become negative near the end of an anneal that reaches zero; the simulator
accepts any finite gamma.
"""
from __future__ import annotations

from typing import Callable, Tuple

import numpy as np

MAX_STAGES = 64

_S3 = np.sqrt(3.0)
A1 = (3.0 - 2.0 * _S3) / 12.0
A2 = (3.0 + 2.0 * _S3) / 12.0
C1 = 0.5 - _S3 / 6.0
C2 = 0.5 + _S3 / 6.0


def _check(stages: int):
    if stages > MAX_STAGES:
        raise ValueError(f"{stages} stages exceed the simulator's {MAX_STAGES}")


def cfet4_stages(gamma_of_t: Callable[[float], float], total_time: float,
                 steps: int) -> Tuple[np.ndarray, np.ndarray]:
    """Stages (gamma', time) of the 4th-order commutator-free integrator: 2 per step."""
    if steps < 1 or not total_time > 0:
        raise ValueError("steps >= 1 and total_time > 0")
    _check(2 * steps)
    h = total_time / steps
    g, t = [], []
    for k in range(steps):
        g1 = float(gamma_of_t(k * h + C1 * h))
        g2 = float(gamma_of_t(k * h + C2 * h))
        g += [2.0 * (A2 * g1 + A1 * g2), 2.0 * (A1 * g1 + A2 * g2)]
        t += [h / 2.0, h / 2.0]
    return np.array(g, np.float32), np.array(t, np.float32)


def midpoint_stages(gamma_of_t: Callable[[float], float], total_time: float,
                    steps: int) -> Tuple[np.ndarray, np.ndarray]:
    """Stages of the 2nd-order exponential midpoint rule: gamma(t0 + h/2) for time h."""
    if steps < 1 or not total_time > 0:
        raise ValueError("steps >= 1 and total_time > 0")
    _check(steps)
    h = total_time / steps
    g = [float(gamma_of_t((k + 0.5) * h)) for k in range(steps)]
    return np.array(g, np.float32), np.full(steps, h, np.float32)


def linear_anneal(gamma0: float, gamma1: float, total_time: float) -> Callable[[float], float]:
    """gamma(t) falling linearly from gamma0 at t = 0 to gamma1 at t = total_time."""
    return lambda t: gamma0 + (gamma1 - gamma0) * (t / total_time)
