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


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_bspec_multi_validation_without_device_work():
    """ps_bspec_multi_scratch_bytes validates the segment table on the host (no device work):
    bad modes, group mismatches and oversized stacks return -1 with a message."""
    from polyspec_amd._abi import BspecSegment
    lib = ctypes.CDLL(LIB)
    lib.ps_last_error.restype = ctypes.c_char_p
    lib.ps_bspec_multi_scratch_bytes.restype = ctypes.c_longlong
    assert lib.ps_bspec_multi_shape_count(32) >= 4 and lib.ps_bspec_multi_shape_count(64) >= 4
    assert lib.ps_bspec_multi_shape_count(16) == -1

    def table(specs):
        arr = (BspecSegment * len(specs))()
        for e, (mode, nmap, npix, ntrip, group) in zip(arr, specs):
            e.mode, e.nmap, e.npix, e.ntrip, e.group = mode, nmap, npix, ntrip, group
            for k in range(4):
                e.X[k] = 4096  # never dereferenced: sizing only
            e.triplets = 4096
        return ctypes.cast(arr, ctypes.c_void_p)
    ok = lib.ps_bspec_multi_scratch_bytes(32, 2, table([(1, 7, 1000, 50, 0), (2, 7, 1000, 50, 1)]), 2)
    assert ok >= 2 * 50 * 16  # at least one block of FP64 complex partials per segment
    assert lib.ps_bspec_multi_scratch_bytes(32, 1, table([(3, 7, 1000, 50, 0)]), 1) == -1
    assert b"mode" in lib.ps_last_error()
    assert lib.ps_bspec_multi_scratch_bytes(32, 2, table([(1, 7, 1000, 50, 0), (2, 7, 1000, 60, 0)]), 1) == -1
    assert b"same number of triplets" in lib.ps_last_error()
    assert lib.ps_bspec_multi_scratch_bytes(32, 1, table([(1, 7, 1000, 50, 0)]), 2) == -1
    assert b"no segment" in lib.ps_last_error()
    assert lib.ps_bspec_multi_scratch_bytes(64, 1, table([(2, 129, 1000, 50, 0)]), 1) == -1
    assert b"nmap" in lib.ps_last_error()
