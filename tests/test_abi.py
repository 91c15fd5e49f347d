"""The C-ABI library loads and exports every entry point of include/polyspec_hip.h,
and the Python bindings declare exactly those entry points (CPU only: no
compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "polyspec_hip.h")
LIB = os.path.join(ROOT, "polyspec_amd", "libpolyspec_hip.so")


def header_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "ps_sht_synthesis" in syms and "ps_gnl_loc_fused_sum" in syms and len(syms) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_header():
    from polyspec_amd._abi import SIGNATURES
    assert sorted(SIGNATURES) == header_symbols()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_path_without_device_work():
    """Argument validation returns an error code and a message instead of aborting."""
    lib = ctypes.CDLL(LIB)
    lib.ps_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    rc = lib.ps_sht_plan_create(0, 10, 10, 0, 64, ctypes.byref(h))
    assert rc == -1
    assert b"nside" in lib.ps_last_error()
    rc = lib.ps_sht_plan_create(4, 10, 10, 0, 16, ctypes.byref(h))
    assert rc == -1 and b"precision" in lib.ps_last_error()
