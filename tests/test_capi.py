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


def test_timer_entry_points_reject_null_context():
    """The region timers' argument checks, without a device."""
    L = gpu.lib()
    s, n = ctypes.c_double(), ctypes.c_int()
    assert L.salmon_gpu_timer_enable(None, 1) == 1
    assert L.salmon_gpu_timer_reset(None) == 1
    assert L.salmon_gpu_timer_get(None, gpu.LOG["DYNAMICS"], ctypes.byref(s), ctypes.byref(n)) == 1


def test_timer_region_numbers_are_the_reference_ones():
    """misc/timer.f90:21-54 numbers the regions; the header's enum and the
    Python table carry the same numbers."""
    src = open(os.path.join(ROOT, "include", "salmon_gpu.h")).read()
    hdr = {m.group(1): int(m.group(2)) for m in re.finditer(r"SALMON_LOG_(\w+) = (\d+)", src)}
    ref = dict(ALL=0, GROUND_STATE=2, DYNAMICS=3, CG=10, DIAG=11, SP_ENERGY=12, GRAM_SCHMIDT=13, DT_EVOLVE=20,
               HPSI=21, PSI_RHO=22, HARTREE=23, CURRENT=25, TOTAL_ENERGY=26)
    assert {k: v for k, v in hdr.items() if k != "MAX"} == ref
    assert gpu.LOG == ref
