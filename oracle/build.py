"""Build recipe for the oracle (test infrastructure only).

* ``build_oracle()`` compiles this directory's C restatement
  (``srs_oracle.c``) into ``oracle/liboracle.so`` with gcc.
* ``build_ref()`` compiles the reference's own CPU step,
  ``cpu/RunSimulation.cpp`` read where it lies under ``/root/reference``,
  together with ``ref_harness.cpp`` into ``oracle/_ref/libsrsref.so``.
  It is built with the reference's own flags (g++ -O3, cpu/Makefile:8) and is
  skipped when the reference tree is absent (on a machine without it, a
  previously built ``_ref`` library is used as is).

Only ``cpu/RunSimulation.cpp`` is compiled.  ``cpu/InitFuncs.cpp`` needs the
xerces-c library and ``mpi.h``, neither of which is installed here, so the
loader and operator assembly are pinned by the restatement plus the
mathematical checks in tests/ instead.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference"
REF_CPU = os.path.join(REF_ROOT, "cpu")
REF_DIR = os.path.join(HERE, "_ref")
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(REF_DIR, "libsrsref.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_oracle() -> str:
    _run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-std=c11",
          "-Wall", "-o", ORACLE_SO, os.path.join(HERE, "srs_oracle.c"), "-lm"])
    return ORACLE_SO


def build_ref() -> str | None:
    src = os.path.join(REF_CPU, "RunSimulation.cpp")
    if not os.path.exists(src):
        print("reference tree absent; keeping existing oracle/_ref build", flush=True)
        return REF_SO if os.path.exists(REF_SO) else None
    os.makedirs(REF_DIR, exist_ok=True)
    obj_rs = os.path.join(REF_DIR, "RunSimulation.o")
    obj_h = os.path.join(REF_DIR, "ref_harness.o")
    _run(["g++", "-O3", "-fPIC", "-fvisibility=hidden", "-ffunction-sections", "-fdata-sections",
          "-w", "-I", REF_CPU, "-c", src, "-o", obj_rs])
    _run(["g++", "-O3", "-fPIC", "-fvisibility=hidden", "-ffunction-sections", "-fdata-sections",
          "-I", REF_CPU, "-c", os.path.join(HERE, "ref_harness.cpp"), "-o", obj_h])
    _run(["g++", "-shared", "-Wl,--gc-sections", "-Wl,-z,defs", "-o", REF_SO, obj_h, obj_rs, "-lm"])
    for o in (obj_rs, obj_h):
        os.remove(o)
    return REF_SO


if __name__ == "__main__":
    build_oracle()
    build_ref()
    sys.exit(0)
