"""The C-ABI library loads and exports every entry point that
include/salmon_gpu.h declares (no compute: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from salmon_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "salmon_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(salmon_gpu_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "salmon_gpu_init" in names and "salmon_gpu_rt_run" in names
    assert set(names) == set(gpu.SIGNATURES), "the Python binding covers the header exactly"


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.fail(f"HIP library not built: {gpu.LIB_PATH}")
    L = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_errstr_without_device():
    L = gpu.lib()
    assert L.salmon_gpu_errstr(0) == b"ok"
    assert L.salmon_gpu_errstr(1) == b"invalid argument"
