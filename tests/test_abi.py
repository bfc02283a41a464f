"""The C-ABI library loads and exports every entry point include/dpx.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import dpx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dpx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "dpx_zuker" in names and "dpx_swat" in names and "dpx_zuker_batch" in names


def test_library_exports_all_declared():
    lib = ctypes.CDLL(dpx.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_exports():
    assert set(dpx.EXPORTS) <= set(_declared())
