"""GPU: mqw_clenshaw_noalloc, the Clenshaw seam (MultiQW.cpp:177-236) without
device allocation, against mqw_clenshaw on the same inputs.

Both run the same jobs through the same kernels, so the outputs must be equal
bit for bit.  Two calls in a row on one context, the second smaller, since the
fixed scratch keeps the first call's contents; sizes past the limit are
rejected; and the device's free memory does not move across a call.
"""
import ctypes as C

import numpy as np
import pytest

import multistageqw_amd as M

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    rec = M.generate_sk(n, 1, 500 + n)[0]
    psi = rng.normal(size=2 << n).astype(np.float32)
    psi /= np.float32(np.sqrt(np.sum(psi.astype(np.float64) ** 2)))
    gamma, t = 0.9, 0.35
    return rec, psi, gamma, t


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_matches_clenshaw_two_calls():
    with M.Context(0) as ctx:
        for n in (20, 14):  # the largest size allowed, then a smaller call on the same scratch
            print(f"[case] no-alloc Clenshaw n={n}", flush=True)
            rec, psi, gamma, t = _case(n, n)
            hp, info = ctx.energy(n, rec)
            scale = float(np.float32(info["e_abs"] + gamma * n))
            cf = M.chebyshev_coeffs(scale * t).astype(np.float32)
            ref = ctx.clenshaw(cf, psi, hp, gamma, scale)
            free0 = _free_bytes()
            out = ctx.clenshaw_noalloc(cf, psi, hp, gamma, scale)
            free1 = _free_bytes()
            assert np.array_equal(out, ref), (n, float(np.max(np.abs(out - ref))))
            assert np.all(np.isfinite(out)) and abs(np.sum(out.astype(np.float64) ** 2) - 1.0) < 1e-4
            assert free1 >= free0 - (2 << 20), (free0, free1)


def test_noalloc_alone_on_a_fresh_context():
    """The no-allocation call first, before any other call has sized a workspace."""
    with M.Context(0) as ctx:
        n = 16
        rec, psi, gamma, t = _case(n, 3)
        hp, info = ctx.energy(n, rec)
        scale = float(np.float32(info["e_abs"] + gamma * n))
        cf = M.chebyshev_coeffs(scale * t).astype(np.float32)
        free0 = _free_bytes()
        out = ctx.clenshaw_noalloc(cf, psi, hp, gamma, scale)
        free1 = _free_bytes()
        assert free1 >= free0 - (2 << 20), (free0, free1)
        ref = ctx.clenshaw(cf, psi, hp, gamma, scale)
        assert np.array_equal(out, ref)


def test_noalloc_rejects_sizes_past_its_limit():
    with M.Context(0) as ctx:
        n = 21
        hp = np.zeros(1 << n, np.float32)
        psi = np.zeros(2 << n, np.float32)
        psi[0] = 1.0
        with pytest.raises(M.MQWError) as e:
            ctx.clenshaw_noalloc(np.ones(8, np.float32), psi, hp, 0.5, 30.0)
        assert e.value.status == 1  # MQW_EINVAL
        n = 20  # a job table of 60000 terms does not fit beside the n = 20 states
        hp = np.zeros(1 << n, np.float32)
        psi = np.zeros(2 << n, np.float32)
        psi[0] = 1.0
        with pytest.raises(M.MQWError) as e:
            ctx.clenshaw_noalloc(np.full(60000, 1e-9, np.float32), psi, hp, 0.5, 30.0)
        assert e.value.status == 1
        assert "no-allocation limit" in str(e.value)
