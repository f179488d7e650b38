"""ctypes wrapper of ``_ref/libsalmon_ref.so`` -- the reference's own Fortran
for the ARTED real-time path, built from its sources by ``oracle/ref/Makefile``.
Test infrastructure only (golden vectors and the timed CPU baseline).

The reference keeps its state in Fortran module globals, so one process holds
one system at a time; ``Ref(sys)`` re-initialises only in a fresh process
(``ref_setup`` allocates module arrays once).  Callers that need several
systems run each in its own subprocess.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libsalmon_ref.so")
# the same reference code with RT/hamiltonian.f90 replaced by the GPU drop-in
LIB_GPU = os.path.join(HERE, "_ref", "libsalmon_ref_gpu.so")
REF_SRC = "/root/reference/src"


def available():
    return os.path.exists(LIB)


def build(gpu_dropin=True):
    """Compile the reference hot path where the reference tree is present
    (and, when the HIP library is built, its variant with the GPU drop-in)."""
    if not os.path.isdir(REF_SRC):
        return available()
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "ref")])
    hip = os.path.join(os.path.dirname(HERE), "salmon_amd", "lib", "libsalmon_gpu.so")
    if gpu_dropin and os.path.exists(hip):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "ref"), "gpu"])
    return available()


_lib = None
_owner = None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib(path=None):
    global _lib
    if _lib is None:
        L = ctypes.CDLL(path or LIB)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.ref_setup.argtypes = [ci, ci, ci, cd, cd, cd, ci, ci, vp, vp, vp, ci, ci, vp, vp, vp, vp, vp,
                                ci, vp, vp, vp, vp, cd, cd, vp]
        L.ref_get_fd.argtypes = [vp]
        L.ref_put_orbitals.argtypes = [vp]
        L.ref_get_orbitals.argtypes = [vp]
        L.ref_hpsi.argtypes = [ci, vp, vp, vp, vp]
        L.ref_hamiltonian.argtypes = [vp]
        L.ref_current.argtypes = [vp, vp]
        L.ref_psi_rho.argtypes = [vp]
        L.ref_energy.argtypes = [vp, vp, vp]
        L.ref_gram_schmidt.argtypes = []
        L.ref_sp_energy.argtypes = [vp, vp]
        L.ref_cg.argtypes = [vp, ci, vp]
        L.ref_rt_run.argtypes = [vp, ci, vp, vp, ci]
        L.ref_move_atoms.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ref_num_threads.restype = ci
        _lib = L
    return _lib


class Ref:
    def __init__(self, sys, lib_path=None):
        global _owner
        if _owner is not None:
            raise RuntimeError("the reference keeps module globals: one system per process")
        _owner = self
        lib(lib_path)
        self.sys = sys
        s = sys
        self._keep = [np.ascontiguousarray(x) for x in (s.kAc0, s.occ, s.vloc)]
        i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
        f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        arrs = dict(kac0=f64(s.kAc0), occ=f64(s.occ), vloc=f64(s.vloc), mps=i32(s.Mps), jxyz=i32(s.Jxyz),
                    jxx=i32(s.Jxx), jyy=i32(s.Jyy), jzz=i32(s.Jzz), atbl=i32(s.a_tbl), uv=f64(s.uV),
                    iuv=i32(s.iuV), rion=f64(s.Rion))
        self._keep += list(arrs.values())
        st = ctypes.c_int(-1)
        lib().ref_setup(s.NLx, s.NLy, s.NLz, s.aLx, s.aLy, s.aLz, s.NK, s.NB, _p(arrs["kac0"]), _p(arrs["occ"]),
                        _p(arrs["vloc"]), s.NI, s.Nps, _p(arrs["mps"]), _p(arrs["jxyz"]), _p(arrs["jxx"]),
                        _p(arrs["jyy"]), _p(arrs["jzz"]), s.Nlma, _p(arrs["atbl"]), _p(arrs["uv"]),
                        _p(arrs["iuv"]), _p(arrs["rion"]), float(s.rc), float(s.dt), ctypes.byref(st))
        if st.value != 0:
            raise RuntimeError("reference padded projector index disagrees with the tables")
        self.put(s.zu)

    def fd(self):
        out = np.zeros((6, 5))
        lib().ref_get_fd(_p(out))
        return out

    def put(self, zu):
        zu = np.ascontiguousarray(zu, dtype=np.complex128)
        lib().ref_put_orbitals(_p(zu))

    def get(self):
        zu = np.zeros((self.sys.NK, self.sys.NB, self.sys.NL), np.complex128)
        lib().ref_get_orbitals(_p(zu))
        return zu

    def hpsi(self, ik, k, psi):
        k = np.ascontiguousarray(k, dtype=np.float64)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        h = np.zeros_like(psi)
        t = np.zeros_like(psi)
        lib().ref_hpsi(int(ik) + 1, _p(k), _p(psi), _p(h), _p(t))
        return h, t

    def hamiltonian(self, kAc):
        kAc = np.ascontiguousarray(kAc, dtype=np.float64)
        lib().ref_hamiltonian(_p(kAc))

    def current(self, kAc):
        kAc = np.ascontiguousarray(kAc, dtype=np.float64)
        j = np.zeros(3)
        lib().ref_current(_p(kAc), _p(j))
        return j

    def density(self):
        r = np.zeros(self.sys.NL)
        lib().ref_psi_rho(_p(r))
        return r

    def energy(self, kAc):
        kAc = np.ascontiguousarray(kAc, dtype=np.float64)
        ek, en = ctypes.c_double(), ctypes.c_double()
        lib().ref_energy(_p(kAc), ctypes.byref(ek), ctypes.byref(en))
        return ek.value, en.value

    def gram_schmidt(self):
        lib().ref_gram_schmidt()

    def sp_energy(self, kAc):
        kAc = np.ascontiguousarray(kAc, dtype=np.float64)
        esp = np.zeros((self.sys.NK, self.sys.NB))
        lib().ref_sp_energy(_p(kAc), _p(esp))
        return esp

    def cg(self, kAc, ncg):
        kAc = np.ascontiguousarray(kAc, dtype=np.float64)
        ev = np.zeros((self.sys.NK, self.sys.NB))
        lib().ref_cg(_p(kAc), int(ncg), _p(ev))
        return ev

    def rt_run(self, ac, nsteps, energy_every=1):
        ac = np.ascontiguousarray(ac[: nsteps + 1], dtype=np.float64)
        jt = np.zeros((nsteps, 3))
        et = np.zeros(nsteps)
        lib().ref_rt_run(_p(ac), int(nsteps), _p(jt), _p(et), int(energy_every))
        return dict(J=jt, E=et)

    def move_atoms(self, Rion, Mps, Jxyz, Jxx, Jyy, Jzz, uV, vloc):
        """New atom positions with their sphere tables, uV and local potential
        (same atoms and projectors), as an Ehrenfest step leaves them."""
        i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
        f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        a = [i32(Mps), i32(Jxyz), i32(Jxx), i32(Jyy), i32(Jzz), f64(uV), f64(Rion), f64(vloc)]
        assert a[1].shape[0] == self.sys.NI and a[5].shape[0] == self.sys.Nlma and a[7].size == self.sys.NL
        lib().ref_move_atoms(int(a[1].shape[1]), *[_p(x) for x in a])

    def num_threads(self):
        return lib().ref_num_threads()
