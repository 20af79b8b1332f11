"""Build recipe for libsrs_amd.so and the cru3d_amd driver (gfx950 only).

Host C++ files are compiled with g++, the HIP translation unit with hipcc for
gfx950, and everything is linked by hipcc into ``srs_amd/libsrs_amd.so``.
FP contraction is off everywhere so that the arithmetic follows the written
operation order; packed-FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsrs_amd.so")
EXE = os.path.join(HERE, "cru3d_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_SRCS = ["host_common.cpp", "host_params.cpp", "host_mesh.cpp", "host_run.cpp"]
HIP_SRCS = ["device.hip", "linescan.hip", "adjoint.hip"]
DEVICE_UNITS = (0, 1, 2, 3, 4)

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I", INCLUDE, "-I", CSRC]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SRCS:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, ["g++", *COMMON, "-Wall", "-I", "/opt/rocm/include", "-c", os.path.join(CSRC, src), "-o", obj]))
    for src in HIP_SRCS:
        # device.hip holds the sweep kernels' many instantiations: it is
        # compiled as five units (SRS_UNIT) that build side by side
        for unit in (DEVICE_UNITS if src == "device.hip" else (None,)):
            obj = os.path.join(BUILD, src + ("" if unit is None else f".{unit}") + ".o")
            defs = [] if unit is None else [f"-DSRS_UNIT={unit}"]
            jobs.append((obj, [HIPCC, *COMMON, *HIP_FLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1, 8)) as ex:
        list(ex.map(lambda j: _run(j[1]), jobs))
    objs = [obj for obj, _ in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    _run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, os.path.join(CSRC, "cru3d_main.cpp"), "-o", EXE,
          "-L", HERE, "-lsrs_amd", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
