"""Pins of the CPU restatement (oracle/) against sources independent of it:
golden vectors (libstdc++, scipy), brute force, closed forms and invariants.
Each check is chosen so a plausible slip (sign, index, dropped term,
transposed operand) fails it.  Reference: MultiQW.cpp."""
import os

import numpy as np
import pytest

from oracle import dense as D

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rng(seed):
    return np.random.default_rng(seed)


# ---- a-7: MT19937 + uniform_real_distribution<float> (MultiQW.cpp:387,417-421)
def test_mt19937_standard_kat(oracle):
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937
    assert oracle.mt_outputs(5489, 10000)[-1] == 4123659995


def test_mt19937_raw_matches_libstdcxx(oracle):
    g = np.load(os.path.join(GOLD, "mt19937_uniform.npz"))
    assert np.array_equal(oracle.mt_outputs(29552825458725 % 2**32, 1000), g["raw"])


def test_uniform_draws_match_reference_build(oracle):
    """The oracle's draws equal libstdc++ built with g++ -O3 -march=native (FMA-contracted),
    and differ from the uncontracted build, so the fused form is what is pinned."""
    g = np.load(os.path.join(GOLD, "mt19937_uniform.npz"))
    s = g["s"]
    m = s.size
    times = oracle.sample_times(m, 100, s) if m > 1 else None
    assert np.array_equal(times, g["draws_native"])
    assert not np.array_equal(g["draws_native"], g["draws_plain"])


def test_m1_sort_and_difference(oracle):
    """MultiQW.cpp:426-430: m == 1 sorts the draws and differences them in place backwards."""
    g = np.load(os.path.join(GOLD, "mt19937_uniform.npz"))
    s = g["s"][:1]
    t = oracle.sample_times(1, 100, s)[0]
    srt = np.sort(g["draws_native"][0])
    assert t[0] == srt[0]
    assert np.array_equal(t[1:], (srt[1:] - srt[:-1]).astype(np.float32))
    assert abs(np.sum(t.astype(np.float64)) - srt[-1]) < 1e-5


# ---- a-4: inverse normal CDF (semantics of my_normcdfinvf, MultiQW.cpp:80-86)
def test_ppnd16_vs_scipy_ndtri(oracle):
    g = np.load(os.path.join(GOLD, "scipy_special.npz"))
    got = np.array([oracle.ppnd16(p) for p in g["p"]])
    assert np.max(np.abs(got - g["ndtri"]) / np.maximum(1, np.abs(g["ndtri"]))) < 1e-14


def test_neg_normcdfinv_sign_and_rounding(oracle):
    # -Phi^{-1}(1/N) is the positive expected-maximum scale (b in MultiQW.cpp:367)
    for n in (5, 10, 20, 28):
        p = np.float32(1) / np.float32(2 ** n)
        v = oracle.neg_normcdfinv_f32(p)
        assert v > 0
        assert v == np.float32(-oracle.ppnd16(float(p)))


# ---- a-8: Chebyshev coefficients (replaces ApproxTools, MultiQW.cpp:438-440)
def test_cheb_coeffs_vs_scipy_jv(oracle):
    g = np.load(os.path.join(GOLD, "scipy_special.npz"))
    for si, s in enumerate(g["s"]):
        c = oracle.cheb_coeffs(s, 1e-6)
        k = np.arange(c.size)
        ref = np.where(k == 0, 1.0, 2.0) * np.where((k >> 1) & 1, -1.0, 1.0) * g["jv"][si, : c.size]
        assert np.max(np.abs(c - ref)) < 1e-13, s


def test_cheb_series_reproduces_cos_plus_sin(oracle):
    """Jacobi-Anger: sum c_k T_k(x) = cos(sx) + sin(sx) on [-1, 1] within the tail bound."""
    x = np.linspace(-1, 1, 2001)
    for s in (0.01, 0.9, 6.0, 40.0, 333.0):
        c = oracle.cheb_coeffs(s, 1e-6)
        err = np.max(np.abs(np.polynomial.chebyshev.chebval(x, c) - (np.cos(s * x) + np.sin(s * x))))
        assert err < 1.5e-6, (s, err)
        # truncation: dropping one more term breaks the tolerance
        assert np.sum(np.abs(c[-1:])) > 0


def test_cheb_truncation_is_minimal(oracle):
    g = np.load(os.path.join(GOLD, "scipy_special.npz"))
    for si, s in enumerate(g["s"][:8]):
        c = oracle.cheb_coeffs(s, 1e-6)
        k = np.arange(g["jv"].shape[1])
        full = 2 * np.abs(g["jv"][si])
        tail = np.cumsum(full[::-1])[::-1]
        K = c.size
        assert tail[K] <= 1e-6 * (1 + 1e-9)
        assert K == 1 or tail[K - 1] > 1e-6 * (1 - 1e-9)


# ---- a-1: record layout (MultiQW.cpp:303-323)
def test_unpack_k_order_n4(oracle):
    n = 4
    rec = np.arange(1, 11, dtype=np.float64)  # 6 couplings then 4 fields
    J, h = oracle.unpack(n, rec)
    # k-order: (1,0), (2,0), (2,1), (3,0), (3,1), (3,2); halved and symmetric
    expect = {(1, 0): 1, (2, 0): 2, (2, 1): 3, (3, 0): 4, (3, 1): 5, (3, 2): 6}
    for (i, j), v in expect.items():
        assert J[i, j] == v / 2 and J[j, i] == v / 2
    assert np.all(np.diag(J) == 0)
    assert np.array_equal(h, np.array([7, 8, 9, 10], np.float32))


# ---- a-2/a-3/a-5: energies and moments (MultiQW.cpp:325-375,395)
@pytest.mark.parametrize("n", [4, 9, 14])
def test_energy_walk_vs_brute_force(oracle, n):
    rec = _rng(n).standard_normal(n * (n + 1) // 2)
    J, h = oracle.unpack(n, rec)
    hp, e, info = oracle.energy_compat(n, J, h)
    ex = oracle.energy_exact(n, J, h)
    # float deltas accumulate a small drift; the walk itself is the brute-force energy
    assert np.max(np.abs(e - ex)) < 1e-4
    assert info["eloc"] == int(np.argmin(ex))
    assert info["e0"] == e.min() and info["emax"] == e.max()
    shift = np.float32((info["emax"] + info["e0"]) / 2)
    assert np.array_equal(hp, (e.astype(np.float32) - shift).astype(np.float32))
    assert info["e_abs"] == np.float32((info["emax"] - info["e0"]) / 2)


def test_energy_index_convention(oracle):
    """Bit i of the state index set <=> spin i = +1; index 0 is all -1 (MultiQW.cpp:331-346)."""
    n = 5
    rec = np.zeros(15)
    rec[10 + 2] = 1.5  # h_2 only
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    k = np.arange(32)
    assert np.array_equal(e, np.where((k >> 2) & 1, 1.5, -1.5))
    assert info["eloc"] == 0  # first index in Gray order among the ties


def test_moments_invariants(oracle):
    """HP2 = mean_k E(k)^2 and delta2 = sum_f mean_k (E(k^2^f) - E(k))^2 (MultiQW.cpp:375,395)."""
    for n in (6, 11):
        rec = _rng(50 + n).standard_normal(n * (n + 1) // 2)
        J, h = oracle.unpack(n, rec)
        ex = oracle.energy_exact(n, J, h)
        sch = oracle.schedule(n, 3, J, h)
        k = np.arange(1 << n)
        assert abs(np.mean(ex)) < 1e-9
        assert abs(sch["hp2"] - np.mean(ex ** 2)) < 2e-6 * np.mean(ex ** 2)
        d2 = sum(np.mean((ex[k ^ (1 << f)] - ex) ** 2) for f in range(n))
        assert abs(sch["delta2"] - d2) < 2e-6 * d2


def test_schedule_closed_forms(oracle):
    """gamma_i and the stage times against an FP64 evaluation of the same heuristic."""
    from scipy.special import ndtri
    n, m = 12, 4
    rec = _rng(9).standard_normal(n * (n + 1) // 2)
    J, h = oracle.unpack(n, rec)
    s = oracle.schedule(n, m, J, h)
    N = 2.0 ** n
    em = 0.5772156649015329
    a = (1 - em) * -ndtri(1 / N) + em * -ndtri(1 / (np.e * N))
    hp2 = 2 * np.sum(J.astype(np.float64) ** 2) + np.sum(h.astype(np.float64) ** 2)
    d2 = 16 * np.sum(J.astype(np.float64) ** 2) + 4 * np.sum(h.astype(np.float64) ** 2)
    g = a * np.sqrt(hp2) / np.tan(np.pi * np.arange(1, m + 1) / (2 * m + 2)) / n
    assert np.allclose(s["gamma"], g, rtol=3e-6)
    assert np.all(np.diff(s["gamma"]) < 0)
    w = g / np.sqrt(1 + g ** 2)
    prev = np.concatenate([[1.0], w[:-1]])
    nxt = np.concatenate([w[1:], [0.0]])
    dH = prev - nxt
    first = np.sqrt(4 * n * dH / d2)
    last = np.sqrt(2 * n * dH / (np.sqrt(2 / np.pi) * g * np.sqrt(n * d2)))
    st = np.where(np.arange(m) == 0, first, np.maximum(first, last))
    assert np.allclose(s["short_t"], st, rtol=3e-6)
    assert abs(s["total_t"] - st.sum()) < 3e-6 * st.sum()


# ---- a-9/a-10: Clenshaw_step and Clenshaw (MultiQW.cpp:88-236)
def test_clenshaw_step_is_two_x_minus_b2_plus_coef_psi(oracle):
    """Brute-force H_G: b2' = (2/scale)(hp b1 - gamma sum_i b1[k^2^i]) - b2 + coef psi."""
    for n in (4, 9, 16):
        N = 1 << n
        r = _rng(n)
        b1, b2, hp, psi = (r.standard_normal(N).astype(np.float32) for _ in range(4))
        scale, gamma, coef = np.float32(17.5), np.float32(0.8), np.float32(-0.35)
        got = b2.copy()
        oracle.clenshaw_step_f32(b1, got, hp, psi, n, scale, gamma, coef)
        k = np.arange(N)
        hg = sum(b1[k ^ (1 << i)].astype(np.float64) for i in range(n))
        ref = (2 / scale) * (hp * b1.astype(np.float64) - gamma * hg) - b2 + coef * psi
        assert np.max(np.abs(got - ref)) < 2e-5 * max(1, np.max(np.abs(ref))), n
        # coef == 0 skips psi entirely
        got0 = b2.copy()
        oracle.clenshaw_step_f32(b1, got0, hp, psi * np.nan, n, scale, gamma, 0.0)
        assert np.all(np.isfinite(got0))


def test_walsh_eigenpairs():
    """H_G w_S = (n - 2|S|) w_S for the Walsh functions w_S(k) = (-1)^{popcount(k & S)}."""
    n = 6
    H = D.hypercube(n)
    k = np.arange(1 << n)
    for S in (0, 1, 5, 63, 42):
        w = (-1.0) ** np.array([bin(x & S).count("1") for x in k])
        assert np.allclose(H @ w, (n - 2 * bin(S).count("1")) * w)


@pytest.mark.parametrize("n", [6, 10])
def test_clenshaw_vs_dense_expm(oracle, n):
    N = 1 << n
    rec = _rng(70 + n).standard_normal(n * (n + 1) // 2)
    J, h = oracle.unpack(n, rec)
    hp, _, info = oracle.energy_compat(n, J, h)
    gamma, t = 1.1, 1.3
    scale = float(np.float32(info["e_abs"] + np.float32(gamma) * n))
    c = oracle.cheb_coeffs(scale * t)
    z = _rng(n).standard_normal(N) + 1j * _rng(n + 1).standard_normal(N)
    z /= np.linalg.norm(z)
    psi = np.concatenate([z.real, z.imag])
    ref = D.evolve(hp.astype(np.float64), gamma, t, z)
    o64 = oracle.clenshaw_f64(c, psi, hp.astype(np.float64), n, gamma, scale)
    o32 = oracle.clenshaw_f32(c.astype(np.float32), psi.astype(np.float32), hp, n, gamma, scale, False)
    assert np.max(np.abs(o64[:N] + 1j * o64[N:] - ref)) < 2e-6
    assert np.max(np.abs(o32[:N] + 1j * o32[N:] - ref)) < 5e-6
    # the sign of gamma matters for |+>: exp(+i(H_P + gamma H_G)t) is a different walk
    wrong = D.evolve(hp.astype(np.float64), -gamma, t, z)
    assert np.max(np.abs(o64[:N] + 1j * o64[N:] - wrong)) > 1e-2


def test_clenshaw_is_unitary_beyond_dense_sizes(oracle):
    """exp(iHt) preserves the norm: at n = 14 (no dense check) the FP64 and FP32
    restatements keep ||psi|| = 1, and a wrong odd-term phase would not."""
    n = 14
    N = 1 << n
    rec = _rng(90).standard_normal(n * (n + 1) // 2)
    J, h = oracle.unpack(n, rec)
    hp, _, info = oracle.energy_compat(n, J, h)
    gamma, t = 0.9, 2.1
    scale = float(np.float32(info["e_abs"] + np.float32(gamma) * n))
    c = oracle.cheb_coeffs(scale * t)
    z = _rng(91).standard_normal(N) + 1j * _rng(92).standard_normal(N)
    z /= np.linalg.norm(z)
    psi = np.concatenate([z.real, z.imag])
    o64 = oracle.clenshaw_f64(c, psi, hp.astype(np.float64), n, gamma, scale)
    o32 = oracle.clenshaw_f32(c.astype(np.float32), psi.astype(np.float32), hp, n, gamma, scale, False)
    assert abs(np.linalg.norm(o64) - 1.0) < 1e-9
    assert abs(np.linalg.norm(o32.astype(np.float64)) - 1.0) < 2e-5
    # real coefficients only (cos(sx) + sin(sx) without the i on odd terms) is not unitary
    bad = oracle.clenshaw_f64(c * np.where(np.arange(c.size) % 2 == 1, 0.0, 1.0), psi, hp.astype(np.float64), n,
                              gamma, scale)
    assert abs(np.linalg.norm(bad) - 1.0) > 1e-3


def test_group_property_pins_m1_chain(oracle):
    """e^{iHt2} e^{iHt1} = e^{iH(t1+t2)}: chained stages equal one long stage (the m == 1 path)."""
    n = 8
    N = 1 << n
    rec = _rng(4).standard_normal(n * (n + 1) // 2)
    J, h = oracle.unpack(n, rec)
    hp, _, info = oracle.energy_compat(n, J, h)
    gamma = 0.9
    scale = float(np.float32(info["e_abs"] + np.float32(gamma) * n))
    psi = np.concatenate([np.full(N, 1 / np.sqrt(N)), np.zeros(N)])
    a = oracle.clenshaw_f64(oracle.cheb_coeffs(scale * 0.3), psi, hp.astype(np.float64), n, gamma, scale)
    ab = oracle.clenshaw_f64(oracle.cheb_coeffs(scale * 0.5), a, hp.astype(np.float64), n, gamma, scale)
    one = oracle.clenshaw_f64(oracle.cheb_coeffs(scale * 0.8), psi, hp.astype(np.float64), n, gamma, scale)
    assert np.max(np.abs(ab - one)) < 3e-6


def test_product_closed_form_j0(oracle):
    """J = 0: the walk factorises; the oracle's full driver reproduces the product formula."""
    n, m, samples = 9, 3, 5
    rec = np.zeros(n * (n + 1) // 2)
    rec[n * (n - 1) // 2:] = _rng(12).uniform(0.4, 1.6, n) * _rng(13).choice([-1, 1], n)
    J, h = oracle.unpack(n, rec)
    _, e, info = oracle.energy_compat(n, J, h)
    sch = oracle.schedule(n, m, J, h)
    times = oracle.sample_times(m, samples, sch["short_t"])
    _, _, sp = oracle.solve(n, m, rec[None, :], samples=samples, precision=1, with_samples=True)
    bits = [(info["eloc"] >> i) & 1 for i in range(n)]
    for j in range(samples):
        ref = D.product_probability(h.astype(np.float64), sch["gamma"].astype(np.float64), times[:, j], bits)
        assert abs(sp[0, j] - ref) < 3e-6


def test_driver_fp32_vs_fp64_and_dense(oracle):
    """The FP32 restatement of the whole per-problem loop against FP64 and dense expm."""
    n = 7
    rec = _rng(21).standard_normal((2, n * (n + 1) // 2))
    for m in (1, 2):
        p32, tt, s32 = oracle.solve(n, m, rec, samples=20, with_samples=True)
        p64, _, s64 = oracle.solve(n, m, rec, samples=20, precision=1, with_samples=True)
        assert np.max(np.abs(s32 - s64)) < 2e-6
        for k in range(2):
            J, h = oracle.unpack(n, rec[k])
            _, e, info = oracle.energy_compat(n, J, h)
            sch = oracle.schedule(n, m, J, h)
            times = oracle.sample_times(m, 20, sch["short_t"])
            ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, sch["gamma"].astype(np.float64), times,
                                     info["eloc"], reset=(m != 1))
            assert np.max(np.abs(s64[k] - ref)) < 3e-6
            assert tt[k] == sch["total_t"]


@pytest.mark.parametrize("size", [1, 15, 16, 17, 31, 32, 33, 100, (1 << 20) + 3])
def test_eigen_sum_f32_error_bound(oracle, size):
    """oracle_eigen_sum_f32 (the float reductions of MultiQW.cpp:443,453) against the exact sum.

    Positive data, with the elements past the last 16-lane packet and the second
    accumulator's lanes made larger, so a dropped tail or a missing accumulator
    merge moves the sum by far more than rounding can.  Two bounds hold for any
    summation order that adds every element once: the general (size - 1) u sum|x|,
    and the tighter one for 32-lane accumulation, (size/32 + 8 + tail) u sum|x|.
    """
    rng = np.random.default_rng(size)
    x = rng.uniform(1.0, 2.0, size).astype(np.float32)
    a1 = (size // 16) * 16
    a2 = (size // 32) * 32
    lane = np.arange(size) % 32
    x[:a2][lane[:a2] >= 16] *= np.float32(3.0)  # the second accumulator's lanes
    if size > a1:  # the tail past the last packet: a tenth of the total
        x[a1:] = np.float32(0.1 * float(np.sum(x[:a1], dtype=np.float64)) / (size - a1)) if a1 else x[a1:]
    u = 2.0 ** -24
    exact = float(np.sum(x.astype(np.float64)))
    total = float(np.sum(np.abs(x.astype(np.float64))))
    loose = max(size - 1, 1) * u * total
    tight = (size // 32 + 8 + (size - a1)) * u * total
    s = float(oracle.eigen_sum(x))
    assert abs(s - exact) <= loose
    assert abs(s - exact) <= tight
    # the mistakes the bounds must catch, made on the CPU from the same data
    if size > a1:
        dropped_tail = float(oracle.eigen_sum(x[:a1])) if a1 else 0.0
        assert abs(dropped_tail - exact) > loose
    if a2 >= 64:
        no_merge = exact - float(np.sum(x[:a2][lane[:a2] >= 16].astype(np.float64)))
        assert abs(no_merge - exact) > tight


def test_float_norm_rounding_at_large_n(oracle):
    """The reference renormalises each stage with a float sum of 2^(n+1) squares
    (MultiQW.cpp:443, Eigen's order).  At n = 18 that sum's rounding moves the
    probabilities by far more than the rest of the FP32 arithmetic does: with the
    squares summed in double the FP32 restatement is within 2e-7 of FP64, and the
    float sum alone moves it many times further.  This is why GPU results (FP64
    norms) are compared with both forms of the oracle."""
    import multistageqw_amd as M
    n, m = 18, 4
    recs = M.generate_sk(n, 1, 6)
    kw = dict(samples=4, threads=4, with_samples=True)
    eig = oracle.solve(n, m, recs, **kw)[2]
    f64n = oracle.solve(n, m, recs, norm="f64", **kw)[2]
    exact = oracle.solve(n, m, recs, precision=1, **kw)[2]
    rest = float(np.max(np.abs(f64n - exact)))
    float_sum = float(np.max(np.abs(eig - exact)))
    assert rest < 2e-7
    assert float_sum > 5 * rest
