"""CPU-side checks of the C ABI library: it loads and exports every function
include/vsmcx.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "vsmcx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vsmcx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "vsmcx_resample" in names and "vsmcx_filter_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    so = os.path.join(ROOT, "vsmc_amd", "libvsmcx.so")
    if not os.path.exists(so):
        pytest.skip("libvsmcx.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(so)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.vsmcx_version.restype = ctypes.c_char_p
    assert lib.vsmcx_version().startswith(b"vsmcx")
