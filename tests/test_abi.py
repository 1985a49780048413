"""The C-ABI library loads and exports every symbol include/fesom_gpu.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "fesom_gpu.h")).read()
    return sorted(set(re.findall(r"\b(fgpu_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding():
    from fesom2_amd import gpu
    assert _declared() == gpu.EXPORTED


def test_library_exports_all_symbols():
    from fesom2_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libfesom_gpu.so not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    lib.fgpu_version.restype = ctypes.c_int
    assert lib.fgpu_version() >= 1
