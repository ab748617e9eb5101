"""ctypes binding of libmqw.so (include/mqw.h).

There is no fallback: if the library is missing or a call fails, the call
raises.  The library is built in-tree by ``multistageqw_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmqw.so")
CLI_PATH = os.path.join(HERE, "mqw")

ABI_VERSION = 18  # MQW_ABI_VERSION of include/mqw.h this binding matches
MAX_STAGES = 64
DEFAULT_SAMPLES = 100
DEFAULT_TOL = 1e-6

OK, EINVAL, ENOMEM, EIO, EDEVICE, ENOTSUP = range(6)
ENERGY_COMPAT, ENERGY_EXACT = 0, 1
FP32, FP64 = 0, 1

# Symbols declared in include/mqw.h (checked by the CPU test suite).
EXPORTS = [
    "mqw_create", "mqw_destroy", "mqw_last_error", "mqw_status_string", "mqw_abi_version", "mqw_opts_default",
    "mqw_unpack", "mqw_schedule_heuristic", "mqw_schedule_heuristic_moments", "mqw_sample_times", "mqw_chebyshev_coeffs", "mqw_generate_sk",
    "mqw_energy", "mqw_clenshaw_step_dev", "mqw_clenshaw", "mqw_clenshaw_noalloc", "mqw_solve", "mqw_solve_batch", "mqw_tune",
    "mqw_last_stats", "mqw_count_instances", "mqw_read_instances", "mqw_append_output", "mqw_features",
    "mqw_mlp_create", "mqw_mlp_destroy", "mqw_mlp_get_params", "mqw_mlp_set_params", "mqw_mlp_forward",
    "mqw_mlp_loss_grad", "mqw_mlp_adam_step", "mqw_diag_moments",
]
FEATURES = 32
MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 32, 128, 10, 22026


class MQWError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mqw status {status}: {msg}")
        self.status = status


class EnergyInfo(C.Structure):
    _fields_ = [("e0", C.c_double), ("emax", C.c_double), ("eloc", C.c_uint32), ("shift", C.c_float),
                ("e_abs", C.c_float), ("e_init", C.c_double)]


class Schedule(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("a", C.c_float), ("hp2", C.c_float), ("delta2", C.c_float),
                ("total_t", C.c_float), ("gamma", C.c_float * MAX_STAGES), ("short_t", C.c_float * MAX_STAGES)]


class Opts(C.Structure):
    _fields_ = [("samples", C.c_int), ("tol", C.c_double), ("precision", C.c_int), ("energy_mode", C.c_int),
                ("tile_bits", C.c_int), ("chunk_bits", C.c_int), ("small_n_max", C.c_int),
                ("use_graph", C.c_int), ("max_device_bytes", C.c_size_t), ("group_bytes", C.c_size_t), ("variant", C.c_int),
                ("sample_probs", C.POINTER(C.c_float)), ("high_pass", C.c_int),
                ("mixer_masks", C.POINTER(C.c_uint32)), ("mixer_count", C.c_int), ("fixed_times", C.c_int),
                ("shared_first_stage", C.c_int), ("final_readout", C.c_int), ("shared_second_stage", C.c_int),
                ("overlap_bases", C.c_int), ("second_stage_gemm", C.c_int), ("mixer_kind", C.c_int),
                ("compress_bases", C.c_int), ("direct_m1", C.c_int), ("adjoint_two_stages", C.c_int),
                ("shared_third_stage", C.c_int),
                ("chain_recurrences", C.c_int),
                ("chain_wait_polls", C.c_int)]


class Instance(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("rec", C.POINTER(C.c_double)),
                ("gamma", C.POINTER(C.c_float)), ("short_t", C.POINTER(C.c_float))]


class Stats(C.Structure):
    _fields_ = [("sweeps", C.c_uint64), ("term_units", C.c_uint64), ("amp_terms", C.c_uint64),
                ("algorithmic_bytes", C.c_double), ("gpu_ms", C.c_double), ("sweep_ms", C.c_double),
                ("lds_trajectories", C.c_uint64), ("psi_const_amp_terms", C.c_uint64),
                ("required_bytes", C.c_double), ("aux_ms", C.c_double), ("aux_bytes", C.c_double),
                ("aux_launches", C.c_uint64), ("basis_ms", C.c_double), ("basis_bytes", C.c_double),
                ("gemm_ms", C.c_double), ("gemm_flops", C.c_double), ("gemm_useful_flops", C.c_double),
                ("basis_launches", C.c_uint64), ("gemm_launches", C.c_uint64), ("gemm_bytes", C.c_double),
                ("basis_batches", C.c_uint64), ("overlap_batches", C.c_uint64),
                ("compressed_instances", C.c_uint64), ("compressed_vectors", C.c_uint64),
                ("uncompressed_vectors", C.c_uint64), ("direct_chains", C.c_uint64),
                ("first_stage_chunks", C.c_uint64), ("adjoint2_instances", C.c_uint64),
                ("adjoint2_vectors", C.c_uint64), ("adjoint2_uncompressed", C.c_uint64),
                ("third_stage_instances", C.c_uint64), ("third_stage_vectors", C.c_uint64),
                ("third_stage_states", C.c_uint64),
                ("chain_launches", C.c_uint64), ("chain_steps", C.c_uint64)]


_lib = None


def load():
    """Load libmqw.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m multistageqw_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.mqw_abi_version.restype = C.c_int
    if L.mqw_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {L.mqw_abi_version()}, this binding needs {ABI_VERSION}: "
                          "rebuild it with `python -m multistageqw_amd.build`")
    fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "mqw_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "mqw_destroy": ([vp], None),
        "mqw_last_error": ([vp], C.c_char_p),
        "mqw_status_string": ([C.c_int], C.c_char_p),
        "mqw_abi_version": ([], C.c_int),
        "mqw_opts_default": ([C.POINTER(Opts)], None),
        "mqw_unpack": ([C.c_int, dp, fp, fp], C.c_int),
        "mqw_schedule_heuristic": ([C.c_int, C.c_int, dp, C.POINTER(Schedule)], C.c_int),
        "mqw_schedule_heuristic_moments": ([C.c_int, C.c_int, dp, C.c_double, C.c_double, C.POINTER(Schedule)],
                                           C.c_int),
        "mqw_sample_times": ([C.c_int, C.c_int, fp, fp], C.c_int),
        "mqw_chebyshev_coeffs": ([C.c_double, C.c_double, dp, C.c_int, ip], C.c_int),
        "mqw_generate_sk": ([C.c_int, C.c_int, C.c_uint64, dp], C.c_int),
        "mqw_energy": ([vp, C.c_int, dp, C.c_int, fp, dp, C.POINTER(EnergyInfo)], C.c_int),
        "mqw_diag_moments": ([vp, C.c_int, dp, C.c_int, dp], C.c_int),
        "mqw_clenshaw_step_dev": ([vp, vp, vp, vp, vp, C.c_uint, C.c_float, C.c_float, C.c_float], C.c_int),
        "mqw_clenshaw": ([vp, C.c_int, fp, C.c_int, fp, fp, C.c_float, C.c_float, C.c_int], C.c_int),
        "mqw_clenshaw_noalloc": ([vp, C.c_int, fp, C.c_int, fp, fp, C.c_float, C.c_float, C.c_int], C.c_int),
        "mqw_solve": ([vp, C.c_int, C.c_int, dp, C.c_int, C.POINTER(Opts), fp, fp], C.c_int),
        "mqw_solve_batch": ([vp, C.POINTER(Instance), C.c_int, C.POINTER(Opts), fp, fp], C.c_int),
        "mqw_tune": ([vp, C.c_int, C.c_char_p, ip, ip, ip, ip], C.c_int),
        "mqw_last_stats": ([vp, C.POINTER(Stats)], C.c_int),
        "mqw_count_instances": ([C.c_char_p, C.c_int, C.POINTER(C.c_long)], C.c_int),
        "mqw_read_instances": ([C.c_char_p, C.c_int, C.c_long, C.c_long, dp], C.c_int),
        "mqw_append_output": ([C.c_char_p, C.c_int, C.c_int, fp, C.c_long], C.c_int),
        "mqw_features": ([vp, C.POINTER(Instance), C.c_int, fp], C.c_int),
        "mqw_mlp_create": ([vp, C.c_uint64, C.POINTER(vp)], C.c_int),
        "mqw_mlp_destroy": ([vp], None),
        "mqw_mlp_get_params": ([vp, fp, C.c_size_t], C.c_int),
        "mqw_mlp_set_params": ([vp, fp, C.c_size_t], C.c_int),
        "mqw_mlp_forward": ([vp, fp, C.c_int, fp], C.c_int),
        "mqw_mlp_loss_grad": ([vp, fp, fp, fp, C.c_int, fp, fp], C.c_int),
        "mqw_mlp_adam_step": ([vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check(status, ctx=None):
    if status != OK:
        L = load()
        msg = L.mqw_last_error(ctx).decode() if ctx else L.mqw_status_string(status).decode()
        raise MQWError(status, msg)
