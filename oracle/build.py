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


def build_bridge_check() -> str | None:
    """oracle/_ref/bridge_check: the per-step bridge of INTEGRATION.md (its
    first C block, extracted to _ref/srs_bridge.inc) in a driver that steps it
    beside the reference's own ExplicitStep (cpu/RunSimulation.cpp compiled
    where it lies), linked against srs_amd/libsrs_amd.so.  Needs the reference
    tree and a built library; skipped otherwise (an existing build is kept)."""
    import re

    src = os.path.join(REF_CPU, "RunSimulation.cpp")
    lib_dir = os.path.join(os.path.dirname(HERE), "srs_amd")
    exe = os.path.join(REF_DIR, "bridge_check")
    if not os.path.exists(src) or not os.path.exists(os.path.join(lib_dir, "libsrs_amd.so")):
        print("reference tree or library absent; keeping existing bridge_check", flush=True)
        return exe if os.path.exists(exe) else None
    os.makedirs(REF_DIR, exist_ok=True)
    with open(os.path.join(os.path.dirname(HERE), "INTEGRATION.md")) as f:
        blocks = re.findall(r"```c\n(.*?)```", f.read(), re.S)
    with open(os.path.join(REF_DIR, "srs_bridge.inc"), "w") as f:
        f.write("#include <cstdio>\n#include <cstdlib>\n#include <cstring>\n" + blocks[0])
    inc = os.path.join(os.path.dirname(HERE), "include")
    obj_rs = os.path.join(REF_DIR, "RunSimulation_exe.o")
    obj_b = os.path.join(REF_DIR, "bridge_check.o")
    _run(["g++", "-O3", "-ffunction-sections", "-fdata-sections", "-w", "-I", REF_CPU, "-c", src, "-o", obj_rs])
    _run(["g++", "-O2", "-std=c++17", "-ffunction-sections", "-fdata-sections", "-I", REF_CPU, "-I", inc, "-I",
          REF_DIR, "-c", os.path.join(HERE, "bridge_check.cpp"), "-o", obj_b])
    _run(["g++", "-Wl,--gc-sections", "-o", exe, obj_b, obj_rs, "-L", lib_dir, "-lsrs_amd",
          "-Wl,-rpath,$ORIGIN/../../srs_amd", "-lm"])
    for o in (obj_rs, obj_b):
        os.remove(o)
    return exe


if __name__ == "__main__":
    build_oracle()
    build_ref()
    build_bridge_check()
    sys.exit(0)
