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
