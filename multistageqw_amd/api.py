"""Python host API over libmqw, mirroring the pieces of MultiQW.cpp.

Reference (Asa-Hopkins/MultiStageQW, MultiQW.cpp)      here
  instance read + J/h unpack   :268-323                  read_instances, unpack
  Gray-code energies, E_loc    :325-362                  Context.energy
  heuristic gamma / times      :365-430                  schedule_heuristic, sample_times
  ApproxTools RCF_odd_even     :438-440                  chebyshev_coeffs
  Clenshaw_step                :88-173                   Context.clenshaw_step_ptr (device pointers)
  Clenshaw                     :177-236                  Context.clenshaw
  per-problem loop of main     :296-455                  Context.solve / solve_batch
  output_<n>_<m> + stdout      :264-266,423,455,457     write_output, format_line

Everything runs through the C ABI of libmqw.so; the GPU work runs in its
gfx950 kernels.  Nothing here computes on the host what the library does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


def params_per_instance(n: int) -> int:
    return n * (n + 1) // 2


def count_instances(path: str, n: int) -> int:
    c = C.c_long(0)
    L.check(L.load().mqw_count_instances(path.encode(), n, C.byref(c)))
    return c.value


def read_instances(path: str, n: int, start: int = 0, count: Optional[int] = None) -> np.ndarray:
    """Read records of n(n+1)/2 little-endian float64 (MultiQW.cpp:268-271,303-304).

    Raises on a missing file or when fewer than ``count`` records remain, where
    the reference would silently repeat its last record.
    """
    if count is None:
        count = count_instances(path, n) - start
    out = np.zeros(max(count, 0) * params_per_instance(n), np.float64)
    L.check(L.load().mqw_read_instances(path.encode(), n, start, count, L.f64p(out)))
    return out.reshape(count, params_per_instance(n))


def unpack(n: int, rec) -> tuple[np.ndarray, np.ndarray]:
    rec = np.ascontiguousarray(rec, np.float64)
    J = np.zeros((n, n), np.float32)
    h = np.zeros(n, np.float32)
    L.check(L.load().mqw_unpack(n, L.f64p(rec), L.f32p(J), L.f32p(h)))
    return J, h


@dataclass
class Schedule:
    gamma: np.ndarray
    short_t: np.ndarray
    total_t: np.float32
    a: np.float32
    hp2: np.float32
    delta2: np.float32


def schedule_heuristic(n: int, m: int, rec, skewness=None, excess_kurtosis=None) -> Schedule:
    """MultiQW.cpp:365-423; with the diagonal's skewness and excess kurtosis
    (Context.diag_moments), the Cornish-Fisher corrected ground-state depth."""
    rec = np.ascontiguousarray(rec, np.float64)
    s = L.Schedule()
    if skewness is None and excess_kurtosis is None:
        L.check(L.load().mqw_schedule_heuristic(n, m, L.f64p(rec), C.byref(s)))
    else:
        L.check(L.load().mqw_schedule_heuristic_moments(n, m, L.f64p(rec), C.c_double(skewness or 0.0),
                                                        C.c_double(excess_kurtosis or 0.0), C.byref(s)))
    return Schedule(np.array(s.gamma[:m], np.float32), np.array(s.short_t[:m], np.float32),
                    np.float32(s.total_t), np.float32(s.a), np.float32(s.hp2), np.float32(s.delta2))


def sample_times(m: int, samples: int, short_t) -> np.ndarray:
    st = np.ascontiguousarray(short_t, np.float32)
    out = np.zeros(m * samples, np.float32)
    L.check(L.load().mqw_sample_times(m, samples, L.f32p(st), L.f32p(out)))
    return out.reshape(m, samples)


def chebyshev_coeffs(s: float, tol: float = L.DEFAULT_TOL, cap: int = 1 << 16) -> np.ndarray:
    c = np.zeros(cap, np.float64)
    K = C.c_int(0)
    L.check(L.load().mqw_chebyshev_coeffs(float(s), float(tol), L.f64p(c), cap, C.byref(K)))
    return c[:K.value].copy()


def generate_sk(n: int, count: int, seed: int) -> np.ndarray:
    out = np.zeros(count * params_per_instance(n), np.float64)
    L.check(L.load().mqw_generate_sk(n, count, seed, L.f64p(out)))
    return out.reshape(count, params_per_instance(n))


def format_line(total_t: float, prob: float) -> str:
    """One stdout line of the reference: iostream defaults (6 significant digits)."""
    return f"{float(np.float32(total_t)):.6g} {float(np.float32(prob)):.6g}"


def write_output(n: int, m: int, probs, directory: str = ".") -> str:
    """Append float32 results to output_<n>_<m> (MultiQW.cpp:264-266,457)."""
    p = np.ascontiguousarray(probs, np.float32)
    L.check(L.load().mqw_append_output(directory.encode(), n, m, L.f32p(p), p.size))
    return os.path.join(directory, f"output_{n}_{m}")


def make_opts(samples=L.DEFAULT_SAMPLES, tol=L.DEFAULT_TOL, precision=L.FP32, energy_mode=L.ENERGY_COMPAT,
              tile_bits=0, chunk_bits=-1, small_n_max=-1, use_graph=True, max_device_bytes=0,
              group_bytes=0, variant=-1, high_pass=-1, mixer_masks=None, fixed_times=False,
              shared_first_stage=True, final_readout=True, shared_second_stage=True,
              overlap_bases=True, gemm="split_fp16", mixer="xstring", compress_bases=True,
              direct_m1=True, adjoint_two_stages=True, shared_third_stage=True,
              chain_recurrences=True, chain_wait_polls=0) -> L.Opts:
    o = L.Opts()
    L.load().mqw_opts_default(C.byref(o))
    o.samples = samples
    o.tol = tol
    o.precision = precision
    o.energy_mode = energy_mode
    o.tile_bits = tile_bits
    o.chunk_bits = chunk_bits
    o.small_n_max = small_n_max
    o.use_graph = int(bool(use_graph))
    o.max_device_bytes = max_device_bytes
    o.group_bytes = group_bytes
    o.variant = variant
    o.high_pass = high_pass
    o.fixed_times = int(bool(fixed_times))
    o.shared_first_stage = int(bool(shared_first_stage))
    o.final_readout = int(bool(final_readout))
    o.shared_second_stage = int(bool(shared_second_stage))
    o.overlap_bases = int(bool(overlap_bases))
    o.compress_bases = int(bool(compress_bases))
    o.direct_m1 = int(bool(direct_m1))
    o.adjoint_two_stages = int(bool(adjoint_two_stages))
    o.shared_third_stage = int(bool(shared_third_stage))
    o.chain_recurrences = int(bool(chain_recurrences))
    o.chain_wait_polls = int(chain_wait_polls)
    gemms = {"split_fp16": 0, "fp32": 1}
    if gemm not in gemms:
        raise ValueError(f"gemm must be one of {sorted(gemms)}, got {gemm!r}")
    o.second_stage_gemm = gemms[gemm]
    mixers = {"xstring": 0, "xy": 1}
    if mixer not in mixers:
        raise ValueError(f"mixer must be one of {sorted(mixers)}, got {mixer!r}")
    o.mixer_kind = mixers[mixer]
    if mixer_masks is not None:
        masks = np.ascontiguousarray(mixer_masks, np.uint32)
        o._masks = masks  # keeps the array alive as long as the options
        o.mixer_masks = masks.ctypes.data_as(C.POINTER(C.c_uint32))
        o.mixer_count = masks.size
    return o


class Context:
    """One device, one HIP stream, a grow-only device workspace (mqw_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = L.load()
        h = C.c_void_p()
        st = self._lib.mqw_create(device, C.byref(h))
        if st != L.OK:
            raise L.MQWError(st, f"mqw_create(device={device}) failed: "
                             f"{self._lib.mqw_status_string(st).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mqw_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        L.check(st, self._h)

    # -- energies (MultiQW.cpp:325-362) -------------------------------------
    def energy(self, n: int, rec, mode: int = L.ENERGY_COMPAT, want_e64: bool = False):
        rec = np.ascontiguousarray(rec, np.float64)
        hp = np.zeros(1 << n, np.float32)
        e = np.zeros(1 << n, np.float64) if want_e64 else None
        info = L.EnergyInfo()
        self._check(self._lib.mqw_energy(self._h, n, L.f64p(rec), mode, L.f32p(hp),
                                         L.f64p(e) if e is not None else None, C.byref(info)))
        d = dict(e0=info.e0, emax=info.emax, eloc=int(info.eloc), shift=info.shift, e_abs=info.e_abs,
                 e_init=info.e_init)
        return (hp, e, d) if want_e64 else (hp, d)

    def diag_moments(self, n: int, rec, mode: int = L.ENERGY_COMPAT) -> dict:
        """Mean, variance, skewness and excess kurtosis of the diagonal energies
        (the moments MultiQW.cpp:375,379-380 has in closed form or commented out)."""
        rec = np.ascontiguousarray(rec, np.float64)
        out = np.zeros(4, np.float64)
        self._check(self._lib.mqw_diag_moments(self._h, n, L.f64p(rec), mode, L.f64p(out)))
        return dict(mean=out[0], variance=out[1], skewness=out[2], excess_kurtosis=out[3])

    # -- Clenshaw_step on device buffers (MultiQW.cpp:88-173) ---------------
    def clenshaw_step_ptr(self, b1: int, b2: int, hp: int, psi: int, n: int, scale: float, gamma: float,
                          coef: float):
        """Device addresses (from the HIP runtime libmqw links) of 2^n float32 each; b2 is updated."""
        self._check(self._lib.mqw_clenshaw_step_dev(self._h, b1, b2, hp, psi, n, scale, gamma, coef))

    # -- Clenshaw (MultiQW.cpp:177-236) -------------------------------------
    def clenshaw(self, coeffs, psi, hp, gamma: float, scale: float, psi_real: bool = False) -> np.ndarray:
        hp = np.ascontiguousarray(hp, np.float32)
        n = int(np.log2(hp.size))
        out = np.ascontiguousarray(psi, np.float32).copy()
        cf = np.ascontiguousarray(coeffs, np.float32)
        self._check(self._lib.mqw_clenshaw(self._h, n, L.f32p(cf), cf.size, L.f32p(out), L.f32p(hp), gamma,
                                           scale, int(psi_real)))
        return out

    def clenshaw_noalloc(self, coeffs, psi, hp, gamma: float, scale: float, psi_real: bool = False) -> np.ndarray:
        """clenshaw() through mqw_clenshaw_noalloc: no device allocation, buffers in a fixed
        per-device array (n <= 20)."""
        hp = np.ascontiguousarray(hp, np.float32)
        n = int(np.log2(hp.size))
        out = np.ascontiguousarray(psi, np.float32).copy()
        cf = np.ascontiguousarray(coeffs, np.float32)
        self._check(self._lib.mqw_clenshaw_noalloc(self._h, n, L.f32p(cf), cf.size, L.f32p(out), L.f32p(hp),
                                                   gamma, scale, int(psi_real)))
        return out

    # -- the per-problem loop (MultiQW.cpp:296-455) -------------------------
    def solve(self, n: int, m: int, recs, with_samples: bool = False, **opts):
        recs = np.ascontiguousarray(recs, np.float64).reshape(-1, params_per_instance(n))
        cnt = recs.shape[0]
        o = make_opts(**opts)
        prob = np.zeros(cnt, np.float32)
        tt = np.zeros(cnt, np.float32)
        sp = None
        if with_samples:
            sp = np.zeros(cnt * o.samples, np.float32)
            o.sample_probs = L.f32p(sp)
        self._check(self._lib.mqw_solve(self._h, n, m, L.f64p(recs), cnt, C.byref(o), L.f32p(prob), L.f32p(tt)))
        if with_samples:
            return prob, tt, sp.reshape(cnt, o.samples)
        return prob, tt

    def solve_batch(self, instances: Sequence[tuple], with_samples: bool = False, **opts):
        """instances: (n, m, rec) or (n, m, rec, gamma, short_t) tuples of uneven size."""
        keep = []
        arr = (L.Instance * max(1, len(instances)))()
        for k, ins in enumerate(instances):
            n, m, rec = ins[0], ins[1], np.ascontiguousarray(ins[2], np.float64)
            keep.append(rec)
            arr[k].n, arr[k].m, arr[k].rec = n, m, L.f64p(rec)
            if len(ins) > 3 and ins[3] is not None:
                g = np.ascontiguousarray(ins[3], np.float32)
                t = np.ascontiguousarray(ins[4], np.float32)
                keep += [g, t]
                arr[k].gamma, arr[k].short_t = L.f32p(g), L.f32p(t)
        o = make_opts(**opts)
        cnt = len(instances)
        prob = np.zeros(cnt, np.float32)
        tt = np.zeros(cnt, np.float32)
        sp = None
        if with_samples:
            sp = np.zeros(cnt * o.samples, np.float32)
            o.sample_probs = L.f32p(sp)
        self._check(self._lib.mqw_solve_batch(self._h, arr, cnt, C.byref(o), L.f32p(prob), L.f32p(tt)))
        if with_samples:
            return prob, tt, sp.reshape(cnt, o.samples)
        return prob, tt

    def tune(self, n: int, cache_path: Optional[str] = None) -> dict:
        """Autotune the sweep for n (mqw_tune); returns the chosen configuration."""
        tb, lc, var, hp = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        self._check(self._lib.mqw_tune(self._h, n, cache_path.encode() if cache_path else None, C.byref(tb),
                                       C.byref(lc), C.byref(var), C.byref(hp)))
        return dict(tile_bits=tb.value, chunk_bits=lc.value, variant=var.value, high_pass=hp.value)

    def last_stats(self) -> dict:
        s = L.Stats()
        self._check(self._lib.mqw_last_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Stats._fields_}
