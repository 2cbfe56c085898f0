"""The C-ABI library loads without a GPU and exports every entry point include/bem2d_hip.h
declares (no compute calls are made)."""
import ctypes
import os
import re

from bem2d_amd import capi
from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bem2d_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bem2d_\w+)\s*\(", text)))


def test_library_exports_header_symbols():
    L = ctypes.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(capi.SIGNATURES)


def test_error_path_without_gpu():
    L = capi.lib()
    rc = L.bem2d_hankel01_host(None, -1, None)
    assert rc == capi.ERR_ARG
    assert b"bad argument" in L.bem2d_last_error()
