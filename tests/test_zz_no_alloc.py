"""The allocation-free Hartree + XC entry (salmon_gpu_hartree_xc_noalloc)
against the allocating one (salmon_gpu_hartree_xc) on the same densities:
two calls in a row on two contexts, the second on a smaller grid, so the
shared device scratch is reused with the previous call's contents in it.
Output buffers start as NaN, so an unwritten point fails the comparison.

Tolerances: V_xc is pointwise from the same density, so equal to the last
bit; V_H 1e-13 relative (the twiddles come from the device's sincospi here
and from the host's cos/sin there, which may differ in the last bit).
"""
import ctypes

import numpy as np
import pytest

from helpers import rel
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers for the test through libamdhip64 (the entry point takes device pointers)."""

    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        self.L = L
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, p, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _compare(c, s, seed, xc):
    rng = np.random.default_rng(seed)
    rho = 0.02 + rng.random(s.NL) * 0.3
    vh_ref, vx_ref = c.hartree_xc(rho, xc)
    hip = _Hip()
    try:
        nan = np.full(s.NL, np.nan)
        d_rho, d_vh, d_vx = hip.put(rho), hip.put(nan), hip.put(nan)
        c.hartree_xc_noalloc(d_rho, xc, d_vh, d_vx)
        c.synchronize()
        vh, vx = hip.get(d_vh, s.NL), hip.get(d_vx, s.NL)
    finally:
        hip.free()
    print(f"[gpu] hartree_xc_noalloc {s.NLx}x{s.NLy}x{s.NLz} xc={xc}: vh {rel(vh, vh_ref):.2e} "
          f"vxc {rel(vx, vx_ref) if np.any(vx_ref) else 0.0:.2e}", flush=True)
    assert np.all(np.isfinite(vh)) and np.all(np.isfinite(vx))
    assert rel(vh, vh_ref) < 1e-13
    assert np.array_equal(vx, vx_ref)


def test_noalloc_hartree_xc_matches_allocating_entry():
    big = synth.make_system(N=(20, 20, 28), L=(12.8, 12.8, 17.9), NB=1, seed=3, orbitals=False)
    small = synth.make_system(N=(12, 12, 12), L=8.0, NB=1, seed=4, orbitals=False)
    big.zu = synth.random_orbitals(np.random.default_rng(1), big.NK, big.NB, big.NL, big.Hxyz)
    small.zu = synth.random_orbitals(np.random.default_rng(2), small.NK, small.NB, small.NL, small.Hxyz)
    with gpu.Context(big, gpu.FP64) as c:
        _compare(c, big, 11, gpu.XC_PZ)
    with gpu.Context(small, gpu.FP64) as c:
        _compare(c, small, 12, gpu.XC_PZ)
        _compare(c, small, 13, gpu.XC_NONE)


def test_noalloc_hartree_xc_rejects_grids_beyond_its_scratch():
    s = synth.make_system(N=(72, 64, 64), L=(46.2, 41.0, 41.0), NB=1, seed=5, natoms=1, atoms="random",
                          orbitals=False)
    s.zu = synth.random_orbitals(np.random.default_rng(3), s.NK, s.NB, s.NL, s.Hxyz)
    hip = _Hip()
    try:
        # real buffers of the grid's size: the call is refused before any launch
        d_rho, d_vh = hip.put(np.full(s.NL, 0.1)), hip.put(np.zeros(s.NL))
        with gpu.Context(s, gpu.FP64) as c:
            with pytest.raises(gpu.SalmonGPUError, match="allocation-free limit"):
                c.hartree_xc_noalloc(d_rho, gpu.XC_PZ, d_vh, 0)
    finally:
        hip.free()
