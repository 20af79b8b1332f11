"""The C-ABI library loads and exports every entry point include/srs_abi.h
declares (no device needed), and host-side entry points fail loudly."""
import ctypes
import os
import re

import pytest

import srs_amd as S
from srs_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(srs_\w+)\s*\(", src, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_abi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert declared_symbols() <= set(_abi.EXPORTED)


def test_version_and_error_strings():
    assert b"gfx950" in S.lib().srs_version()
    with pytest.raises(S.SrsError) as e:
        S.read_parameters("/nonexistent/file.xml")
    assert "cannot open" in str(e.value)


def test_context_without_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(S.SrsError):
        S.Context(0)


def test_header_is_plain_c(tmp_path):
    """include/srs_abi.h compiles as C99 and links from C: a caller in the
    reference's own language family (its cpu/ build is C-style C++) needs no
    C++ types.  Host-only calls run; the device call reports no device (or,
    with a GPU present, creates and destroys a context)."""
    import shutil
    import subprocess

    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "bridge.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "srs_abi.h"
int main(void) {
  int32_t ints[SRS_N_INT]; double reals[SRS_N_REAL]; char title[256];
  if (srs_read_parameters(NULL, ints, reals, title) != SRS_OK) return 1;
  if (reals[SRS_R_D_CA] <= 0.0) return 2;
  srs_box_spec spec; memset(&spec, 0, sizeof spec);
  if (srs_read_parameters("/nonexistent.xml", ints, reals, title) == SRS_OK) return 3;
  if (strstr(srs_last_error(), "cannot open") == NULL) return 4;
  srs_ctx* ctx = NULL;
  int rc = srs_ctx_create(0, &ctx);
  if (rc == SRS_OK) srs_ctx_destroy(ctx);
  else if (rc != SRS_E_NODEV && rc != SRS_E_HIP) return 5;
  printf("%s\n", srs_version());
  return 0;
}
''')
    lib_dir = os.path.dirname(_abi.LIB_PATH)
    exe = tmp_path / "bridge"
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", lib_dir, "-lsrs_amd", f"-Wl,-rpath,{lib_dir}"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stderr)
    assert "gfx950" in out.stdout


REF_CPU = "/root/reference/cpu"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_CPU, "parameters.h")), reason="reference sources not present")
def test_integration_snippets_compile(tmp_path):
    """The C bridge of the integration notes compiles against the reference's own
    cpu/parameters.h (its SimData/GridStruct/ParamStruct field names) and our
    header, and the fitting snippets compile as C99."""
    import shutil
    import subprocess

    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    blocks = re.findall(r"```c\n(.*?)```", open(os.path.join(ROOT, "INTEGRATION.md")).read(), re.S)
    assert len(blocks) >= 4
    bridge = tmp_path / "srs_bridge.cpp"
    bridge.write_text("#include <cstdio>\n#include <cstdlib>\n#include <cstring>\n" + blocks[0] +
                      "\nvoid SrsDeviceLoop(FILE* states_file) {\n" + blocks[1] + "\n}\n")
    fit = tmp_path / "fit.c"
    fit.write_text("#include <stdlib.h>\n#include \"srs_abi.h\"\n"
                   "void fit_one(srs_ctx* ctx, int64_t n_steps, int32_t n_ryr, int32_t scan_id, int32_t n_samples,\n"
                   "             const int64_t* sample_steps, const double* target, const double* new_reals) {\n"
                   + blocks[2] + "}\n"
                   "void fit_batch(srs_ctx* ctx, int64_t n_steps, int32_t scan_id, int32_t n_samples,\n"
                   "               const int64_t* sample_steps, const double* targets, int32_t* schedules,\n"
                   "               const int32_t* ids) {\n  enum { N = 4 };\n" + blocks[3] + "}\n")
    inc = os.path.join(ROOT, "include")
    for cmd in (["g++", "-fsyntax-only", "-Wall", "-Werror", "-I", REF_CPU, "-I", inc, str(bridge)],
                ["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, str(fit)]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
