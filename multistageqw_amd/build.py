"""In-tree build of libmqw.so (gfx950 kernels + C ABI) and the drop-in CLI ``mqw``.

Run ``python -m multistageqw_amd.build`` (or ``__graft_entry__.build()``).
hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmqw.so")
CLI = os.path.join(HERE, "mqw")
ARCH = os.environ.get("MQW_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 VALU ops are kept out of every kernel (they are disabled for the
# device pass; hipcc's host pass warns that it does not know the feature).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-Wno-unused-value", f"-I{INCLUDE}", f"-I{CSRC}",
          f"-I{ROCM}/include"]

KERNEL_SOURCES = ["mqw_kernels.hip", "mqw_lds.hip", "mqw_mlp.hip", "mqw_noalloc.hip", "mqw_gram.hip", "mqw_chain.hip"]
HOST_SOURCES = ["mqw_engine.cpp", "mqw_host.cpp", "mqw_tune.cpp", "mqw_io.cpp", "mqw_learn.cpp", "mqw_basis.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _obj(src):
    return os.path.join(OBJ, os.path.splitext(src)[0] + ".o")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    steps = []
    for src in KERNEL_SOURCES:
        path = os.path.join(CSRC, src)
        if os.path.exists(path):
            steps.append([HIPCC, "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, "-c", path,
                          "-o", _obj(src)])
    for src in HOST_SOURCES:
        path = os.path.join(CSRC, src)
        if os.path.exists(path):
            # host-only C++ (no device pass) against the HIP runtime API
            steps.append([HIPCC, "-x", "c++", *COMMON, "-ffp-contract=off", "-pthread", "-D__HIP_PLATFORM_AMD__", "-c",
                          path, "-o", _obj(src)])
    with ThreadPoolExecutor(max_workers=min(8, len(steps))) as ex:
        outs = list(ex.map(_run, steps))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [s[-1] for s in steps]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "mqw_cli.cpp")
    if os.path.exists(cli_src):
        _run([HIPCC, "-x", "c++", *COMMON, "-D__HIP_PLATFORM_AMD__", "-o", CLI, cli_src, f"-L{HERE}", "-lmqw",
              f"-Wl,-rpath,$ORIGIN"])
    return LIB


def build_copy_probe() -> str:
    """The plain-copy kernel whose counter readings calibrate the HBM traffic
    profiles (hbm_copy): measurement infrastructure, not part of libmqw."""
    out = os.path.join(HERE, "hbm_copy")
    _run([HIPCC, "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, "-o", out,
          os.path.join(CSRC, "hbm_copy.hip")])
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_copy_probe())
