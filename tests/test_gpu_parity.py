"""GPU parity: libmqw's HIP path against the CPU restatement (oracle/) and the
dense FP64 reference, through the C ABI.  Reference: MultiQW.cpp."""
import numpy as np
import pytest

import multistageqw_amd as M
from oracle import dense as D

pytestmark = pytest.mark.gpu

# Ground-state probability bar against the reference restatement (FP32 vs FP32).
PROB_TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n", [4, 5, 9, 12, 13, 16, 20])
def test_energy_compat_bit_exact(ctx, oracle, n):
    """MultiQW.cpp:325-362: the GPU walk reproduces the serial float-delta walk."""
    _say("energy compat n", n)
    rec = M.generate_sk(n, 1, 100 + n)[0]
    hp, e, info = ctx.energy(n, rec, M.ENERGY_COMPAT, want_e64=True)
    J, h = oracle.unpack(n, rec)
    hp_o, e_o, info_o = oracle.energy_compat(n, J, h)
    assert info["eloc"] == info_o["eloc"]
    assert info["e0"] == info_o["e0"] and info["emax"] == info_o["emax"]
    assert info["shift"] == info_o["shift"] and info["e_abs"] == info_o["e_abs"]
    assert np.array_equal(e, e_o)
    assert np.array_equal(hp, hp_o)


@pytest.mark.parametrize("n", [4, 10, 15])
def test_energy_exact_matches_brute_force(ctx, oracle, n):
    _say("energy exact n", n)
    rec = M.generate_sk(n, 1, 200 + n)[0]
    hp, e, info = ctx.energy(n, rec, M.ENERGY_EXACT, want_e64=True)
    J, h = oracle.unpack(n, rec)
    ex = oracle.energy_exact(n, J, h)
    assert np.max(np.abs(e - ex)) < 1e-10
    assert info["eloc"] == int(np.argmin(ex))


def test_clenshaw_step_seam(ctx, oracle):
    """The device-pointer seam of MultiQW.cpp:88-173 on one real half."""
    from tests.hipmem import DeviceArray
    for n in (4, 11, 16):
        N = 1 << n
        _say("clenshaw_step n", n)
        rng = np.random.default_rng(3 + n)
        b1, b2, hp, psi = (rng.standard_normal(N).astype(np.float32) for _ in range(4))
        hp *= 20
        for coef in (0.3, 0.0):
            ref = b2.copy()
            oracle.clenshaw_step_f32(b1, ref, hp, psi, n, 45.0, 1.7, coef)
            dev = [DeviceArray(x) for x in (b1, b2, hp, psi)]
            ctx.clenshaw_step_ptr(dev[0].ptr.value, dev[1].ptr.value, dev[2].ptr.value, dev[3].ptr.value, n, 45.0,
                                  1.7, coef)
            got = dev[1].numpy()
            for d in dev:
                d.free()
            assert np.max(np.abs(got - ref)) < 2e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n,psi_real", [(10, True), (13, False), (17, False)])
def test_clenshaw_seam_vs_oracle_and_dense(ctx, oracle, n, psi_real):
    """MultiQW.cpp:177-236: one Chebyshev evaluation of exp(iHt) psi."""
    _say("clenshaw n", n)
    N = 1 << n
    rec = M.generate_sk(n, 1, 300 + n)[0]
    J, h = oracle.unpack(n, rec)
    hp, _, info = oracle.energy_compat(n, J, h)
    gamma, t = 1.3, 0.8
    scale = float(np.float32(info["e_abs"] + np.float32(gamma) * n))
    c = oracle.cheb_coeffs(scale * t).astype(np.float32)
    rng = np.random.default_rng(n)
    if psi_real:
        psi = np.concatenate([np.full(N, 1 / np.sqrt(N)), np.zeros(N)]).astype(np.float32)
    else:
        z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        z /= np.linalg.norm(z)
        psi = np.concatenate([z.real, z.imag]).astype(np.float32)
    got = ctx.clenshaw(c, psi, hp, gamma, scale, psi_real)
    ref = oracle.clenshaw_f32(c, psi, hp, n, gamma, scale, psi_real)
    assert np.max(np.abs(got - ref)) < 1e-5
    if n <= 10:
        zz = D.evolve(hp.astype(np.float64), gamma, t, psi[:N] + 1j * psi[N:])
        assert np.max(np.abs((got[:N] + 1j * got[N:]) - zz)) < 1e-5


def test_solve_c1_vs_oracle(ctx, oracle, c1_recs):
    """n=10, m=2, 64 instances: the reference's CPU-runnable case."""
    _say("solve C1 n=10 m=2 x64")
    p, tt = ctx.solve(10, 2, c1_recs)
    po, tto = oracle.solve(10, 2, c1_recs, threads=8)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(p - po)) < PROB_TOL
    pg, _ = ctx.solve(10, 2, c1_recs, small_n_max=0, use_graph=False)
    assert np.max(np.abs(pg - po)) < PROB_TOL


SOLVE_CASES = [(4, 2, 3, 20), (8, 1, 6, 100), (11, 3, 4, 100), (13, 2, 2, 30), (14, 5, 2, 12), (17, 2, 1, 6)]


PATHS = {
    "default": {},                                        # n <= 13: LDS-resident; else tuned sweeps
    "global": dict(small_n_max=0),                        # sweeps for every n
    "complex": dict(small_n_max=0, tile_bits=12),         # complex-tile sweep kernel
    "half13": dict(small_n_max=0, tile_bits=13, variant=1),   # real-half sweep kernel, 2^13 tiles
    "half14": dict(small_n_max=0, tile_bits=14),          # real-half sweep kernel, 2^14 tiles
    "halfdirect": dict(small_n_max=0, tile_bits=13, use_graph=False, group_bytes=48 << 20),  # small groups
    "halfpaired": dict(small_n_max=0, tile_bits=13, variant=10),  # tile pairs across the top tile bit
    "halfboth": dict(small_n_max=0, tile_bits=13, variant=16),    # both halves of a tile per workgroup
}


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("n,m,count,samples", SOLVE_CASES)
def test_solve_vs_oracle(ctx, oracle, n, m, count, samples, path):
    if path == "global" and n > 13:
        pytest.skip("same as default above n = 13")
    if path.startswith("half") and n < 13:
        pytest.skip("real-half tiles need n >= 13")
    if path == "half14" and n < 14:
        pytest.skip("2^14 real-half tiles need n >= 14")
    _say("solve n", n, "m", m, "count", count, "samples", samples, "path", path)
    recs = M.generate_sk(n, count, 1000 + 10 * n + m)
    kw = PATHS[path]
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, **kw)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=8, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    assert np.max(np.abs(p - po)) < PROB_TOL


def test_solve_fp64_vs_dense(ctx, oracle):
    """FP64 GPU path against exact dense evolution (n=8, both m==1 and m>1 paths)."""
    n = 8
    recs = M.generate_sk(n, 2, 77)
    for m in (1, 3):
        _say("fp64 vs dense m", m)
        p, tt, sp = ctx.solve(n, m, recs, samples=10, precision=M.FP64, with_samples=True)
        for k in range(2):
            J, h = oracle.unpack(n, recs[k])
            hp, e, info = oracle.energy_compat(n, J, h)
            sch = oracle.schedule(n, m, J, h)
            times = oracle.sample_times(m, 10, sch["short_t"])
            ref = D.walk_probability(e - (info["emax"] + info["e0"]) / 2, sch["gamma"].astype(np.float64), times,
                                     info["eloc"], reset=(m != 1))
            assert np.max(np.abs(sp[k] - ref)) < 5e-6


@pytest.mark.parametrize("variant", [10, 16, 7, 13])
def test_workgroup_modes_are_bit_identical(ctx, variant):
    """Tile pairs (variants 6-11) and both-halves workgroups (12-17) keep every sum in the same order."""
    _say("sweep variant", variant, "vs variant 4, n=17 m=2")
    recs = M.generate_sk(17, 2, 77)
    _, _, a = ctx.solve(17, 2, recs, samples=6, with_samples=True, small_n_max=0, tile_bits=13, variant=4)
    _, _, b = ctx.solve(17, 2, recs, samples=6, with_samples=True, small_n_max=0, tile_bits=13, variant=variant)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,m,tb,samples,hp", [(17, 2, 13, 6, 1), (19, 3, 13, 3, 1), (21, 2, 14, 2, 1),
                                                (18, 1, 14, 6, 1), (19, 2, 13, 3, 2), (21, 3, 13, 2, 3),
                                                (19, 2, 13, 3, 4), (21, 3, 14, 2, 4)])
def test_high_bit_pass_vs_oracle(ctx, oracle, n, m, tb, samples, hp):
    """The sweep with the high-bit pass: hp 1..3, the top min(11 - hp, n - tb) bits summed in LDS
    pencils with rows of 32 * 2^hp bytes; hp 4, the register variant (4..7 bits, 128 B rows)."""
    _say("high-bit pass n", n, "m", m, "tb", tb, "samples", samples, "rows", 32 << hp)
    recs = M.generate_sk(n, 2, 500 + n)
    p, tt, sp = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, tile_bits=tb,
                          high_pass=hp)
    po, tto, spo = oracle.solve(n, m, recs, samples=samples, threads=2, with_samples=True)
    assert np.array_equal(tt, tto)
    assert np.max(np.abs(sp - spo)) < PROB_TOL
    # the same sums in another order: close to the sweep without the pass
    _, _, sq = ctx.solve(n, m, recs, samples=samples, with_samples=True, small_n_max=0, tile_bits=tb, high_pass=0)
    assert np.max(np.abs(sp - sq)) < 1e-6


def test_clenshaw_seam_is_unitary_n18(ctx, oracle):
    """exp(iHt) on the GPU keeps ||psi|| = 1 at n = 18 (FP32 state)."""
    n = 18
    N = 1 << n
    _say("unitarity n=18")
    rec = M.generate_sk(n, 1, 61)[0]
    hp, info = ctx.energy(n, rec, M.ENERGY_COMPAT)
    gamma, t = 1.2, 1.7
    scale = float(np.float32(info["e_abs"] + np.float32(gamma) * n))
    c = M.chebyshev_coeffs(scale * t).astype(np.float32)
    rng = np.random.default_rng(62)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    z /= np.linalg.norm(z)
    out = ctx.clenshaw(c, np.concatenate([z.real, z.imag]).astype(np.float32), hp, gamma, scale)
    assert abs(np.linalg.norm(out.astype(np.float64)) - 1.0) < 2e-5
