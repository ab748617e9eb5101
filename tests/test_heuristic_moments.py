"""Heuristic schedule with the diagonal's own shape: the ground-state depth a
from Cornish-Fisher corrected lower-tail quantiles (skewness, excess kurtosis)
instead of a Gaussian's (no counterpart in the reference: MultiQW.cpp:379-380
has these terms commented out).  Pinned by (1) zero moments giving the
reference heuristic exactly and (2) distributions with exact quantiles."""
import numpy as np
import pytest
from scipy import stats

import multistageqw_amd as M

EULER = 0.5772156649015329


def _depth_a(n, lower_quantile):
    """a = (1 - gamma_E) depth(1/N) + gamma_E depth(1/(eN)) from a standardized distribution's exact quantiles."""
    N = 2.0 ** n
    return (1 - EULER) * -lower_quantile(1 / N) + EULER * -lower_quantile(1 / (np.e * N))


@pytest.mark.parametrize("n,m", [(10, 2), (20, 3), (24, 5)])
def test_zero_moments_give_the_reference_schedule(n, m):
    rec = M.generate_sk(n, 1, 900 + n)[0]
    ref = M.schedule_heuristic(n, m, rec)
    got = M.schedule_heuristic(n, m, rec, skewness=0.0, excess_kurtosis=0.0)
    assert got.a == ref.a
    assert np.array_equal(got.gamma, ref.gamma) and np.array_equal(got.short_t, ref.short_t)


@pytest.mark.parametrize("k", [200, 50])
@pytest.mark.parametrize("n", [14, 20])
@pytest.mark.parametrize("mirror", [False, True])
def test_depth_matches_exact_quantiles_of_a_skewed_distribution(k, n, mirror):
    """Standardized chi-square(k) (skewness sqrt(8/k), excess kurtosis 12/k), and its mirror image
    (negative skewness): the corrected depth is within 1 % of the exact one, where the Gaussian's
    is 5 to 40 % off."""
    s, ek = np.sqrt(8 / k), 12 / k
    if mirror:
        q = lambda p: -(stats.chi2.ppf(1 - p, k) - k) / np.sqrt(2 * k)
        s = -s
    else:
        q = lambda p: (stats.chi2.ppf(p, k) - k) / np.sqrt(2 * k)
    exact = _depth_a(n, q)
    rec = M.generate_sk(n, 1, 950 + n)[0]
    gauss = M.schedule_heuristic(n, 2, rec).a
    cf = M.schedule_heuristic(n, 2, rec, skewness=s, excess_kurtosis=ek).a
    assert abs(cf - exact) < 0.01 * exact
    assert abs(gauss - exact) > 0.05 * exact


def test_right_skew_means_a_shallower_ground_state():
    rec = M.generate_sk(16, 1, 990)[0]
    base = M.schedule_heuristic(16, 3, rec)
    right = M.schedule_heuristic(16, 3, rec, skewness=0.3, excess_kurtosis=0.0)
    assert right.a < base.a and np.all(right.gamma < base.gamma)


def test_non_finite_moments_are_errors():
    rec = M.generate_sk(12, 1, 991)[0]
    with pytest.raises(M.MQWError):
        M.schedule_heuristic(12, 2, rec, skewness=float("nan"), excess_kurtosis=0.0)
