"""ctypes view of the CPU restatement in ``mqw_oracle.c``.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or, for ``cpu_baseline``, as the timed CPU program).  The product
package ``multistageqw_amd`` never imports it.

Each wrapper names the MultiQW.cpp lines (Asa-Hopkins/MultiStageQW) that the C
function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmqw_oracle.so")
_lib = None

MAX_M = 64


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


class EnergyInfo(C.Structure):
    _fields_ = [("e0", C.c_double), ("emax", C.c_double), ("eloc", C.c_uint32),
                ("shift", C.c_float), ("e_abs", C.c_float), ("e_init", C.c_double)]


class Schedule(C.Structure):
    _fields_ = [("a", C.c_float), ("hp2", C.c_float), ("delta2", C.c_float), ("total_t", C.c_float),
                ("gamma", C.c_float * MAX_M), ("short_t", C.c_float * MAX_M)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


def build_native(outdir: str) -> str:
    """Compile the same source as the reference is built (gcc -O3 -march=native)
    into outdir, on the machine that will run it: the CPU baseline's program."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE_DIR={os.path.abspath(outdir)}"],
                   check=True)
    return os.path.join(os.path.abspath(outdir), "libmqw_oracle_native.so")


def load_native(path: str):
    """ctypes handle of a native build (build_native), bound like lib()."""
    return _bind(C.CDLL(path))


def solve_native(L, n: int, m: int, recs, samples: int = 100, threads: int = 1):
    """oracle_solve through the handle L of a native build: (prob, total_t)."""
    recs = np.ascontiguousarray(recs, np.float64).reshape(-1, n * (n + 1) // 2)
    cnt = recs.shape[0]
    prob = np.zeros(cnt, np.float32)
    tt = np.zeros(cnt, np.float32)
    if L.oracle_solve(n, m, _d(recs), cnt, samples, 1e-6, 0, threads, _f(prob), _f(tt), None) != 0:
        raise ValueError("oracle_solve rejected its arguments")
    return prob, tt


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    """Argument and result types of the exported functions."""
    fp, dp, u = C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_uint
    L.oracle_eigen_sum_f32.argtypes = [fp, C.c_long]
    L.oracle_eigen_sum_f32.restype = C.c_float
    L.oracle_unpack.argtypes = [C.c_int, dp, fp, fp]
    L.oracle_energy_compat.argtypes = [C.c_int, fp, fp, fp, dp, C.POINTER(EnergyInfo)]
    L.oracle_energy_exact.argtypes = [C.c_int, fp, fp, dp]
    L.oracle_ppnd16.argtypes = [C.c_double]
    L.oracle_ppnd16.restype = C.c_double
    L.oracle_neg_normcdfinv_f32.argtypes = [C.c_float]
    L.oracle_neg_normcdfinv_f32.restype = C.c_float
    L.oracle_schedule_heuristic.argtypes = [C.c_int, C.c_int, fp, fp, C.POINTER(Schedule)]
    L.oracle_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
    L.oracle_mt_next.argtypes = [C.POINTER(MT)]
    L.oracle_mt_next.restype = C.c_uint32
    L.oracle_sample_times.argtypes = [C.c_int, C.c_int, fp, fp]
    L.oracle_cheb_coeffs.argtypes = [C.c_double, C.c_double, dp, C.c_int]
    L.oracle_cheb_coeffs.restype = C.c_int
    L.oracle_clenshaw_step_f32.argtypes = [fp, fp, fp, fp, u, C.c_float, C.c_float, C.c_float]
    L.oracle_clenshaw_f32.argtypes = [fp, C.c_int, fp, fp, u, C.c_float, C.c_float, C.c_int, fp]
    L.oracle_clenshaw_f64.argtypes = [dp, C.c_int, dp, dp, u, C.c_double, C.c_double, dp]
    L.oracle_solve.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                               fp, fp, fp]
    L.oracle_solve.restype = C.c_int
    L.oracle_solve_sched.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                     fp, fp, fp, fp, fp]
    L.oracle_solve_sched.restype = C.c_int
    L.oracle_solve_ex.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                  fp, fp, fp, fp, fp, C.c_int]
    L.oracle_solve_ex.restype = C.c_int
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def eigen_sum(x) -> np.float32:
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.float32(lib().oracle_eigen_sum_f32(_f(x), x.size))


def unpack(n: int, rec):
    """MultiQW.cpp:303-323 -> (J symmetric float32 n*n with halving, h float32 n)."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    J = np.zeros((n, n), np.float32)
    h = np.zeros(n, np.float32)
    lib().oracle_unpack(n, _d(rec), _f(J), _f(h))
    return J, h


def energy_compat(n: int, J, h):
    """MultiQW.cpp:325-362 -> (H_P float32 shifted, E float64 unshifted, info dict)."""
    N = 1 << n
    hp = np.zeros(N, np.float32)
    e = np.zeros(N, np.float64)
    info = EnergyInfo()
    J = np.ascontiguousarray(J, np.float32)
    h = np.ascontiguousarray(h, np.float32)
    lib().oracle_energy_compat(n, _f(J), _f(h), _f(hp), _d(e), C.byref(info))
    return hp, e, dict(e0=info.e0, emax=info.emax, eloc=int(info.eloc), shift=info.shift,
                       e_abs=info.e_abs, e_init=info.e_init)


def energy_exact(n: int, J, h):
    e = np.zeros(1 << n, np.float64)
    J = np.ascontiguousarray(J, np.float32)
    h = np.ascontiguousarray(h, np.float32)
    lib().oracle_energy_exact(n, _f(J), _f(h), _d(e))
    return e


def ppnd16(p: float) -> float:
    return lib().oracle_ppnd16(p)


def neg_normcdfinv_f32(a: float) -> np.float32:
    return np.float32(lib().oracle_neg_normcdfinv_f32(a))


def schedule(n: int, m: int, J, h) -> dict:
    """MultiQW.cpp:365-423."""
    s = Schedule()
    J = np.ascontiguousarray(J, np.float32)
    h = np.ascontiguousarray(h, np.float32)
    lib().oracle_schedule_heuristic(n, m, _f(J), _f(h), C.byref(s))
    return dict(a=np.float32(s.a), hp2=np.float32(s.hp2), delta2=np.float32(s.delta2),
                total_t=np.float32(s.total_t),
                gamma=np.array(s.gamma[:m], np.float32), short_t=np.array(s.short_t[:m], np.float32))


def mt_outputs(seed: int, count: int) -> np.ndarray:
    g = MT()
    lib().oracle_mt_seed(C.byref(g), seed)
    return np.array([lib().oracle_mt_next(C.byref(g)) for _ in range(count)], np.uint32)


def sample_times(m: int, samples: int, short_t) -> np.ndarray:
    """MultiQW.cpp:387,417-430 -> times[m, samples] float32."""
    st = np.ascontiguousarray(short_t, np.float32)
    t = np.zeros(m * samples, np.float32)
    lib().oracle_sample_times(m, samples, _f(st), _f(t))
    return t.reshape(m, samples)


def cheb_coeffs(s: float, tol: float = 1e-6, cap: int = 1 << 16) -> np.ndarray:
    c = np.zeros(cap, np.float64)
    K = lib().oracle_cheb_coeffs(s, tol, _d(c), cap)
    if K < 0:
        raise ValueError("coefficient capacity too small")
    return c[:K].copy()


def clenshaw_step_f32(b1, b2, hp, psi, n, scale, gamma, coef):
    """MultiQW.cpp:88-173; b2 is updated in place (float32 arrays of 2^n)."""
    lib().oracle_clenshaw_step_f32(_f(b1), _f(b2), _f(hp), _f(psi), n, scale, gamma, coef)


def clenshaw_f32(coeffs, psi, hp, n, gamma, scale, psi_real):
    """MultiQW.cpp:177-236; returns the new planar state (float32, 2N)."""
    coeffs = np.ascontiguousarray(coeffs, np.float32)
    out = np.ascontiguousarray(psi, np.float32).copy()
    hp = np.ascontiguousarray(hp, np.float32)
    scratch = np.zeros(4 << n, np.float32)
    lib().oracle_clenshaw_f32(_f(coeffs), coeffs.size, _f(out), _f(hp), n, gamma, scale, int(psi_real),
                              _f(scratch))
    return out


def clenshaw_f64(coeffs, psi, hp, n, gamma, scale):
    coeffs = np.ascontiguousarray(coeffs, np.float64)
    out = np.ascontiguousarray(psi, np.float64).copy()
    hp = np.ascontiguousarray(hp, np.float64)
    scratch = np.zeros(4 << n, np.float64)
    lib().oracle_clenshaw_f64(_d(coeffs), coeffs.size, _d(out), _d(hp), n, gamma, scale, _d(scratch))
    return out


def solve(n: int, m: int, recs, samples: int = 100, tol: float = 1e-6, precision: int = 0,
          threads: int = 1, with_samples: bool = False, gamma=None, short_t=None, norm: str = "eigen"):
    """Per-problem loop of MultiQW.cpp:296-455.  recs: (count, n(n+1)/2) float64.
    gamma/short_t: optional (count, m) schedules replacing the heuristic.
    norm (FP32 path): "eigen" = each stage's norm as the reference takes it, a float
    sum in Eigen's order (MultiQW.cpp:443); "f64" = the same squares summed in double."""
    if norm not in ("eigen", "f64"):
        raise ValueError(f"norm must be 'eigen' or 'f64', got {norm!r}")
    recs = np.ascontiguousarray(recs, np.float64).reshape(-1, n * (n + 1) // 2)
    cnt = recs.shape[0]
    prob = np.zeros(cnt, np.float32)
    tt = np.zeros(cnt, np.float32)
    sp = np.zeros(cnt * samples, np.float32) if with_samples else None
    g = np.ascontiguousarray(gamma, np.float32).reshape(cnt, m) if gamma is not None else None
    st = np.ascontiguousarray(short_t, np.float32).reshape(cnt, m) if short_t is not None else None
    rc = lib().oracle_solve_ex(n, m, _d(recs), cnt, samples, tol, precision, threads, _f(prob), _f(tt),
                               _f(sp) if sp is not None else None, _f(g) if g is not None else None,
                               _f(st) if st is not None else None, int(norm == "f64"))
    if rc != 0:
        raise ValueError("oracle_solve rejected its arguments")
    if with_samples:
        return prob, tt, sp.reshape(cnt, samples)
    return prob, tt
