"""MI355X (gfx950) multi-stage quantum-walk simulator: drop-in for the hot path
of Asa-Hopkins/MultiStageQW (MultiQW.cpp).  The compute runs in libmqw.so
(hand-written HIP kernels behind the C ABI of include/mqw.h)."""
from ._lib import (ENERGY_COMPAT, ENERGY_EXACT, FP32, FP64, MQWError, load, LIB_PATH, CLI_PATH)  # noqa: F401
from .api import (Context, chebyshev_coeffs, count_instances, format_line, generate_sk, make_opts, params_per_instance,  # noqa: F401
                  read_instances, sample_times, schedule_heuristic, unpack, write_output)

__all__ = ["Context", "chebyshev_coeffs", "count_instances", "format_line", "generate_sk", "make_opts", "params_per_instance",
           "read_instances", "sample_times", "schedule_heuristic", "unpack", "write_output", "MQWError", "load",
           "ENERGY_COMPAT", "ENERGY_EXACT", "FP32", "FP64"]
