"""Python host binding of the gfx950 ARTED hot path (ctypes over the C-ABI
declared in include/salmon_gpu.h).

The class mirrors the reference's call interface for this path: the module
state of ``Global_Variables`` becomes a ``Context`` and the routines keep their
names and argument meaning:

=====================  ==============================================  ===========================
method                 reference routine                               file:line
=====================  ==============================================  ===========================
``hpsi``               hpsi_omp_KB_GS / hpsi_omp_KB_RT                 ARTED/common/hpsi.f90:33,89
``hamiltonian``        hamiltonian(zu, flag_current)                   ARTED/RT/hamiltonian.f90:27
``current``            current_RT -> current('RT')                     ARTED/RT/current.f90:65,78
``psi_rho_RT``         psi_rho_RT(zu)                                  ARTED/common/psi_rho.f90:40
``total_energy``       Total_Energy_omp (Ekin, Enl) + frozen-V Eloc    ARTED/common/total_energy.f90:17
``gram_schmidt``       Gram_Schmidt                                    ARTED/GS/Gram_Schmidt.f90:20
``sp_energy``          sp_energy_omp                                   ARTED/GS/sp_energy.f90:20
``rt_run``             the tddft_sc time loop (middlepoint, frozen V)  ARTED/control/control_sc.f90:150
``cg``                 CG_omp -> CG_ompk                               ARTED/GS/CG.f90:20,34
``diag``               diag_omp (subspace diagonalisation)             ARTED/GS/diag.f90:20
``ground_state``       the GS sweep loop (frozen V)                    ARTED/control/ground_state.f90:136
=====================  ==============================================  ===========================

The reference aborts through ``err_finalize`` on any error; here every failed
call raises ``SalmonGPUError`` with the library's message.  There is no CPU
fallback: without the built HIP library the import of ``lib()`` fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsalmon_gpu.so")

FP64 = 0
MIXED = 1
PATH_AUTO, PATH_STREAMED, PATH_FUSED, PATH_SWEEP = 0, 1, 2, 3
BATCH_AUTO, BATCH_PHASES, BATCH_SWEEP = 0, 1, 2
PROP_MIDPOINT, PROP_ETRS = 0, 1
XC_NONE, XC_PZ = 0, 1
PC_NONE, PC_KINETIC, PC_LEARNED = 0, 1, 2
PCNET_WEIGHTS = 27 * 2 * 16 + 27 * 16 * 16 + 27 * 16 * 2
# region timers: the reference's region numbers (misc/timer.f90:21-54)
LOG = dict(ALL=0, GROUND_STATE=2, DYNAMICS=3, CG=10, DIAG=11, SP_ENERGY=12, GRAM_SCHMIDT=13, DT_EVOLVE=20, HPSI=21,
           PSI_RHO=22, HARTREE=23, CURRENT=25, TOTAL_ENERGY=26)


class SalmonGPUError(RuntimeError):
    pass


class GridT(ctypes.Structure):
    _fields_ = [("NLx", ctypes.c_int), ("NLy", ctypes.c_int), ("NLz", ctypes.c_int),
                ("aLx", ctypes.c_double), ("aLy", ctypes.c_double), ("aLz", ctypes.c_double),
                ("lapx", ctypes.c_double * 5), ("lapy", ctypes.c_double * 5), ("lapz", ctypes.c_double * 5),
                ("nabx", ctypes.c_double * 5), ("naby", ctypes.c_double * 5), ("nabz", ctypes.c_double * 5)]


# entry points declared in include/salmon_gpu.h (name -> argtypes)
_vp, _ci, _cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
SIGNATURES = {
    "salmon_gpu_init": [ctypes.POINTER(GridT), _ci, ctypes.POINTER(_vp)],
    "salmon_gpu_set_potential": [_vp, _vp],
    "salmon_gpu_set_projectors": [_vp, _ci, _ci, _vp, _vp, _vp, _vp, _vp, _ci, _vp, _vp, _vp],
    "salmon_gpu_set_kpoints": [_vp, _ci, _ci, _vp, _vp],
    "salmon_gpu_put_orbitals": [_vp, _vp],
    "salmon_gpu_get_orbitals": [_vp, _vp],
    "salmon_gpu_hpsi": [_vp, _vp, _vp, _vp, _vp],
    "salmon_gpu_hamiltonian": [_vp, _vp, _cd],
    "salmon_gpu_current": [_vp, _vp, _vp],
    "salmon_gpu_density": [_vp, _vp],
    "salmon_gpu_energy": [_vp, _vp, _vp, _vp, _vp],
    "salmon_gpu_norm": [_vp, _vp],
    "salmon_gpu_gram_schmidt": [_vp],
    "salmon_gpu_sp_energy": [_vp, _vp, _vp],
    "salmon_gpu_cg": [_vp, _vp, _ci, _ci, _vp, _vp],
    "salmon_gpu_subspace_diag": [_vp, _vp, _vp],
    "salmon_gpu_ground_state": [_vp, _vp, _ci, _ci, _ci, _cd, _vp, _vp, _vp, _vp],
    "salmon_gpu_pcnet_create": [ctypes.c_uint, ctypes.POINTER(_vp)],
    "salmon_gpu_pcnet_set_weights": [_vp, _vp],
    "salmon_gpu_pcnet_get_weights": [_vp, _vp],
    "salmon_gpu_pcnet_finalize": [_vp],
    "salmon_gpu_pcnet_apply": [_vp, _vp, _vp, _vp],
    "salmon_gpu_set_graphs": [_vp, _ci],
    "salmon_gpu_set_propagator": [_vp, _ci],
    "salmon_gpu_set_scf": [_vp, _vp, _ci],
    "salmon_gpu_hartree_xc": [_vp, _vp, _ci, _vp, _vp],
    "salmon_gpu_hartree_xc_noalloc": [_vp, _vp, _ci, _vp, _vp],
    "salmon_gpu_set_pcnet": [_vp, _vp],
    "salmon_gpu_set_pc_handover": [_vp, _ci, _cd],
    "salmon_gpu_pc_record": [_vp, _ci, _ci],
    "salmon_gpu_pc_samples": [_vp, _vp, _vp],
    "salmon_gpu_pc_get_sample": [_vp, _ci, _vp, _vp, _vp, _vp, _vp],
    "salmon_gpu_pc_set_targets": [_vp, _vp],
    "salmon_gpu_pcnet_gradient": [_vp, _vp, _ci, _vp, _vp, _vp],
    "salmon_gpu_pcnet_train": [_vp, _vp, _ci, _ci, _cd, ctypes.c_uint, _vp],
    "salmon_gpu_rt_run": [_vp, _vp, _ci, _cd, _ci, _vp, _vp, _vp, _vp],
    "salmon_gpu_batch_create": [_ci, ctypes.POINTER(_vp)],
    "salmon_gpu_batch_add": [_vp, _vp],
    "salmon_gpu_batch_set_kernel": [_vp, _ci],
    "salmon_gpu_batch_kernel": [_vp],
    "salmon_gpu_batch_tune": [_vp],
    "salmon_gpu_batch_zblock": [_vp, _ci],
    "salmon_gpu_batch_rt_run": [_vp, _vp, _ci, _cd, _vp, _vp, _vp],
    "salmon_gpu_batch_rt_run_ms": [_vp, _vp, _ci, _cd, _vp, _vp, _vp],
    "salmon_gpu_batch_last_error": [_vp],
    "salmon_gpu_batch_finalize": [_vp],
    "salmon_gpu_set_path": [_vp, _ci],
    "salmon_gpu_path": [_vp],
    "salmon_gpu_synchronize": [_vp],
    "salmon_gpu_timer_enable": [_vp, _ci],
    "salmon_gpu_timer_reset": [_vp],
    "salmon_gpu_timer_get": [_vp, _ci, _vp, _vp],
    "salmon_gpu_stream": [_vp],
    "salmon_gpu_errstr": [_ci],
    "salmon_gpu_last_error": [_vp],
    "salmon_gpu_finalize": [_vp],
}

_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SalmonGPUError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = argt
            if name in ("salmon_gpu_errstr", "salmon_gpu_last_error", "salmon_gpu_batch_last_error"):
                f.restype = ctypes.c_char_p
            elif name == "salmon_gpu_stream":
                f.restype = ctypes.c_void_p
            else:
                f.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """Device-resident state of one periodic cell (the reference's module globals)."""

    def __init__(self, sys, precision=FP64, orbitals=True):
        self.sys = sys
        self.precision = precision
        self._h = ctypes.c_void_p()
        g = GridT()
        g.NLx, g.NLy, g.NLz = sys.NLx, sys.NLy, sys.NLz
        g.aLx, g.aLy, g.aLz = sys.aLx, sys.aLy, sys.aLz
        for name in ("lapx", "lapy", "lapz", "nabx", "naby", "nabz"):
            setattr(g, name, (ctypes.c_double * 5)(*[float(v) for v in getattr(sys, name)]))
        rc = lib().salmon_gpu_init(ctypes.byref(g), int(precision), ctypes.byref(self._h))
        if rc != 0:
            raise SalmonGPUError(f"salmon_gpu_init: {lib().salmon_gpu_errstr(rc).decode()} "
                                 f"({lib().salmon_gpu_last_error(None).decode()})")
        self._check(lib().salmon_gpu_set_potential(self._h, _p(_f64(sys.vloc))), "set_potential")
        self.set_projectors(sys)
        self._check(lib().salmon_gpu_set_kpoints(self._h, sys.NK, sys.NB, _p(_f64(sys.kAc0)), _p(_f64(sys.occ))),
                    "set_kpoints")
        if orbitals:
            self.put_orbitals(sys.zu)

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().salmon_gpu_last_error(self._h).decode() if self._h else ""
            raise SalmonGPUError(f"{what}: {lib().salmon_gpu_errstr(rc).decode()}: {msg}")

    def set_potential(self, vloc):
        """A new local potential Vloc(NL) (the Hartree/XC update of a step)."""
        self._check(lib().salmon_gpu_set_potential(self._h, _p(_f64(vloc))), "set_potential")

    def set_projectors(self, sys):
        """Uploads the KB tables of sys (Mps, Jxyz, Jxx/Jyy/Jzz, a_tbl, uV, iuV);
        called again when the atoms move (prep_ps_periodic)."""
        keep = [_i32(sys.Mps), _i32(sys.Jxyz), _i32(sys.Jxx), _i32(sys.Jyy), _i32(sys.Jzz), _i32(sys.a_tbl),
                _f64(sys.uV), _i32(sys.iuV)]
        self._check(lib().salmon_gpu_set_projectors(self._h, sys.NI, sys.Nps, _p(keep[0]), _p(keep[1]), _p(keep[2]),
                                                    _p(keep[3]), _p(keep[4]), sys.Nlma, _p(keep[5]), _p(keep[6]),
                                                    _p(keep[7])), "set_projectors")

    def close(self):
        if self._h:
            lib().salmon_gpu_finalize(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- data movement --------------------------------------------------
    def put_orbitals(self, zu):
        s = self.sys
        zu = np.ascontiguousarray(zu, dtype=np.complex128)
        if zu.shape != (s.NK, s.NB, s.NL):
            raise SalmonGPUError(f"zu must have shape (NK, NB, NL) = {(s.NK, s.NB, s.NL)}")
        self._check(lib().salmon_gpu_put_orbitals(self._h, _p(zu)), "put_orbitals")

    def get_orbitals(self):
        s = self.sys
        zu = np.zeros((s.NK, s.NB, s.NL), np.complex128)
        self._check(lib().salmon_gpu_get_orbitals(self._h, _p(zu)), "get_orbitals")
        return zu

    # --- operators ------------------------------------------------------
    def hpsi(self, kvec, psi, with_tt=False):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = np.zeros_like(psi)
        tt = np.zeros_like(psi) if with_tt else None
        self._check(lib().salmon_gpu_hpsi(self._h, _p(_f64(kvec)), _p(psi), _p(out), _p(tt) if with_tt else None),
                    "hpsi")
        return (out, tt) if with_tt else out

    hpsi_omp_KB_RT = hpsi

    def hamiltonian(self, kAc, dt):
        self._check(lib().salmon_gpu_hamiltonian(self._h, _p(_f64(kAc)), float(dt)), "hamiltonian")

    def current(self, kAc):
        j = np.zeros(3)
        self._check(lib().salmon_gpu_current(self._h, _p(_f64(kAc)), _p(j)), "current")
        return j

    current_RT = current

    def psi_rho_RT(self):
        rho = np.zeros(self.sys.NL)
        self._check(lib().salmon_gpu_density(self._h, _p(rho)), "psi_rho_RT")
        return rho

    def total_energy(self, kAc):
        ek, en, el = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().salmon_gpu_energy(self._h, _p(_f64(kAc)), ctypes.byref(ek), ctypes.byref(en),
                                            ctypes.byref(el)), "total_energy")
        return ek.value, en.value, el.value

    def norm(self):
        n = ctypes.c_double()
        self._check(lib().salmon_gpu_norm(self._h, ctypes.byref(n)), "norm")
        return n.value

    def gram_schmidt(self):
        self._check(lib().salmon_gpu_gram_schmidt(self._h), "gram_schmidt")

    def sp_energy(self, kAc):
        esp = np.zeros((self.sys.NK, self.sys.NB))
        self._check(lib().salmon_gpu_sp_energy(self._h, _p(_f64(kAc)), _p(esp)), "sp_energy")
        return esp

    def rt_run(self, ac, nsteps, dt=None, reortho_every=0):
        ac = _f64(np.asarray(ac)[: nsteps + 1])
        if ac.shape != (nsteps + 1, 3):
            raise SalmonGPUError("ac must hold nsteps+1 rows of 3")
        jt = np.zeros((nsteps, 3))
        et = np.zeros(nsteps)
        nt = np.zeros(nsteps)
        rho = np.zeros(self.sys.NL)
        self._check(lib().salmon_gpu_rt_run(self._h, _p(ac), int(nsteps), float(self.sys.dt if dt is None else dt),
                                            int(reortho_every), _p(jt), _p(et), _p(nt), _p(rho)), "rt_run")
        return dict(J=jt, E=et, norm=nt, rho=rho)

    def cg(self, kAc, iter_cg_max, precond=PC_NONE):
        """CG_ompk on the resident orbitals; returns (esp_var, iters), each (NK, NB)."""
        s = self.sys
        ev = np.zeros((s.NK, s.NB))
        it = np.zeros((s.NK, s.NB), np.int32)
        self._check(lib().salmon_gpu_cg(self._h, _p(_f64(kAc)), int(iter_cg_max), int(precond), _p(ev), _p(it)), "cg")
        return ev, it

    CG_omp = cg

    def diag(self, kAc):
        """diag_omp: rotates the resident orbitals to the subspace eigenvectors; returns esp (NK, NB)."""
        esp = np.zeros((self.sys.NK, self.sys.NB))
        self._check(lib().salmon_gpu_subspace_diag(self._h, _p(_f64(kAc)), _p(esp)), "subspace_diag")
        return esp

    diag_omp = diag

    def ground_state(self, kAc, max_sweeps, iter_cg_max=4, precond=PC_KINETIC, tol=1e-10):
        """Sweeps of Gram-Schmidt, diag, Gram-Schmidt, CG, Gram-Schmidt until no band energy moves by tol
        between two sweeps; returns dict(esp, esp_var, sweeps, history (sweeps, NK, NB))."""
        s = self.sys
        esp = np.zeros((s.NK, s.NB))
        ev = np.zeros((s.NK, s.NB))
        hist = np.zeros((int(max_sweeps), s.NK, s.NB))
        n = ctypes.c_int()
        self._check(lib().salmon_gpu_ground_state(self._h, _p(_f64(kAc)), int(max_sweeps), int(iter_cg_max),
                                                  int(precond), float(tol), _p(esp), _p(ev), ctypes.byref(n),
                                                  _p(hist)), "ground_state")
        return dict(esp=esp, esp_var=ev, sweeps=n.value, history=hist[: n.value])

    # --- learned preconditioner: network, training samples -----------------
    def set_pcnet(self, net):
        self._pcnet = net
        self._check(lib().salmon_gpu_set_pcnet(self._h, net._h if net is not None else None), "set_pcnet")

    def pc_record(self, max_samples, stride=1):
        self._check(lib().salmon_gpu_pc_record(self._h, int(max_samples), int(stride)), "pc_record")

    def pc_samples(self):
        st, va = ctypes.c_int(), ctypes.c_int()
        self._check(lib().salmon_gpu_pc_samples(self._h, ctypes.byref(st), ctypes.byref(va)), "pc_samples")
        return st.value, va.value

    def set_propagator(self, propagator):
        """PROP_MIDPOINT (default) or PROP_ETRS for rt_run (dt_evolve.f90:28-59)."""
        if isinstance(propagator, str):
            propagator = {"middlepoint": PROP_MIDPOINT, "etrs": PROP_ETRS}[propagator]
        self._check(lib().salmon_gpu_set_propagator(self._h, int(propagator)), "set_propagator")

    def set_scf(self, vpsl, xc=XC_PZ):
        """Self-consistent Vloc = V_H[rho] + vpsl + V_xc[rho] after every RT
        step (vpsl None: back to the frozen potential)."""
        if vpsl is None:
            self._check(lib().salmon_gpu_set_scf(self._h, None, int(xc)), "set_scf")
            return
        v = _f64(vpsl)
        self._check(lib().salmon_gpu_set_scf(self._h, _p(v), int(xc)), "set_scf")

    def hartree_xc(self, rho, xc=XC_PZ):
        """(V_H, V_xc) of the density rho(NL) on this context's grid."""
        r = _f64(rho)
        vh, vx = np.zeros(self.sys.NL), np.zeros(self.sys.NL)
        self._check(lib().salmon_gpu_hartree_xc(self._h, _p(r), int(xc), _p(vh), _p(vx)), "hartree_xc")
        return vh, vx

    def hartree_xc_noalloc(self, rho_dev, xc, vh_dev, vxc_dev):
        """hartree_xc on device buffers (addresses as ints, vh/vxc may be 0),
        enqueued on the context's stream without allocating; call
        synchronize() before reading the results."""
        self._check(lib().salmon_gpu_hartree_xc_noalloc(self._h, _vp(rho_dev), int(xc), _vp(vh_dev or None),
                                                        _vp(vxc_dev or None)), "hartree_xc_noalloc")

    def set_pc_handover(self, max_learned_sweeps, stall_ratio=0.3):
        """Learned -> kinetic handover of ground_state: after at most
        max_learned_sweeps sweeps (0: no limit), earlier when a sweep's drop of
        the band-energy sum falls below stall_ratio times the previous one."""
        self._check(lib().salmon_gpu_set_pc_handover(self._h, int(max_learned_sweeps), float(stall_ratio)),
                    "set_pc_handover")

    def set_graphs(self, enable):
        """CG / Gram-Schmidt through the graph cache (default) or eagerly; a
        program driving several contexts from several threads turns it off."""
        self._check(lib().salmon_gpu_set_graphs(self._h, int(bool(enable))), "set_graphs")

    def pc_get_sample(self, i):
        g = np.zeros(self.sys.NL, np.complex128)
        x = np.zeros(self.sys.NL, np.complex128)
        ib, va = ctypes.c_int(), ctypes.c_int()
        th = ctypes.c_double()
        self._check(lib().salmon_gpu_pc_get_sample(self._h, int(i), _p(g), _p(x), ctypes.byref(ib), ctypes.byref(th),
                                                   ctypes.byref(va)), "pc_get_sample")
        return dict(g=g, x=x, band=ib.value, theta=th.value, valid=bool(va.value))

    def pc_set_targets(self, kAc):
        self._check(lib().salmon_gpu_pc_set_targets(self._h, _p(_f64(kAc))), "pc_set_targets")

    def stream(self):
        """The context's hipStream_t (as an integer handle)."""
        return lib().salmon_gpu_stream(self._h)

    def set_path(self, path):
        self._check(lib().salmon_gpu_set_path(self._h, int(path)), "set_path")

    def path(self):
        return lib().salmon_gpu_path(self._h)

    def synchronize(self):
        self._check(lib().salmon_gpu_synchronize(self._h), "synchronize")

    def timer_enable(self, on=True):
        """Per-region GPU timers (timer_begin/timer_end of misc/timer.f90)."""
        self._check(lib().salmon_gpu_timer_enable(self._h, int(bool(on))), "timer_enable")

    def timer_reset(self):
        self._check(lib().salmon_gpu_timer_reset(self._h), "timer_reset")

    def timer_get(self, region):
        """(seconds, calls) of a region, by number or by the name in LOG."""
        rid = LOG[region] if isinstance(region, str) else int(region)
        sec, calls = ctypes.c_double(), ctypes.c_int()
        self._check(lib().salmon_gpu_timer_get(self._h, rid, ctypes.byref(sec), ctypes.byref(calls)), "timer_get")
        return sec.value, calls.value

    def timers(self):
        """{name: (seconds, calls)} of every region timed so far."""
        out = {}
        for name in LOG:
            t = self.timer_get(name)
            if t[1]:
                out[name] = t
        return out


class PcNet:
    """The learned CG preconditioner's network (device weights + Adam state)."""

    def __init__(self, seed=0, weights=None):
        self._h = ctypes.c_void_p()
        rc = lib().salmon_gpu_pcnet_create(int(seed), ctypes.byref(self._h))
        if rc != 0:
            raise SalmonGPUError(f"salmon_gpu_pcnet_create: {lib().salmon_gpu_errstr(rc).decode()}")
        if weights is not None:
            self.set_weights(weights)

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != PCNET_WEIGHTS:
            raise SalmonGPUError(f"expected {PCNET_WEIGHTS} weights")
        rc = lib().salmon_gpu_pcnet_set_weights(self._h, _p(w))
        if rc != 0:
            raise SalmonGPUError("pcnet_set_weights failed")

    def weights(self):
        w = np.zeros(PCNET_WEIGHTS, np.float32)
        if lib().salmon_gpu_pcnet_get_weights(self._h, _p(w)) != 0:
            raise SalmonGPUError("pcnet_get_weights failed")
        return w

    def gradient(self, ctx, idx):
        idx = _i32(idx)
        g = np.zeros(PCNET_WEIGHTS)
        loss = ctypes.c_double()
        ctx._check(lib().salmon_gpu_pcnet_gradient(self._h, ctx._h, int(idx.size), _p(idx), ctypes.byref(loss), _p(g)),
                   "pcnet_gradient")
        return loss.value, g

    def apply(self, ctx, g):
        """The learned direction of one residual (BF16 matrix-core inference, as inside CG)."""
        g = np.ascontiguousarray(g, dtype=np.complex128)
        d = np.zeros_like(g)
        ctx._check(lib().salmon_gpu_pcnet_apply(self._h, ctx._h, _p(g), _p(d)), "pcnet_apply")
        return d

    def train(self, ctx, nsteps, batch=4, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        ctx._check(lib().salmon_gpu_pcnet_train(self._h, ctx._h, int(nsteps), int(batch), float(lr), int(seed),
                                                ctypes.byref(loss)), "pcnet_train")
        return loss.value

    def close(self):
        if self._h:
            lib().salmon_gpu_pcnet_finalize(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent set-up contexts run together (one launch per phase over all
    systems, the step replayed from a captured graph)."""

    def __init__(self, contexts, precision=FP64):
        self._h = ctypes.c_void_p()
        rc = lib().salmon_gpu_batch_create(int(precision), ctypes.byref(self._h))
        if rc != 0:
            raise SalmonGPUError(f"salmon_gpu_batch_create: {lib().salmon_gpu_errstr(rc).decode()}")
        self.contexts = list(contexts)
        for c in self.contexts:
            self._check(lib().salmon_gpu_batch_add(self._h, c._h), "batch_add")

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().salmon_gpu_batch_last_error(self._h).decode() if self._h else ""
            raise SalmonGPUError(f"{what}: {lib().salmon_gpu_errstr(rc).decode()}: {msg}")

    def set_kernel(self, kernel):
        """BATCH_AUTO, BATCH_PHASES (one launch per phase) or BATCH_SWEEP (one
        persistent launch per step)."""
        self._check(lib().salmon_gpu_batch_set_kernel(self._h, int(kernel)), "batch_set_kernel")

    def kernel(self):
        """The step kernel in use (BATCH_PHASES or BATCH_SWEEP); -1 before tune."""
        return lib().salmon_gpu_batch_kernel(self._h)

    def tune(self):
        self._check(lib().salmon_gpu_batch_tune(self._h), "batch_tune")
        return [lib().salmon_gpu_batch_zblock(self._h, i) for i in range(len(self.contexts))]

    def rt_run(self, ac, nsteps, dt):
        n = len(self.contexts)
        ac = _f64(np.asarray(ac)[: nsteps + 1])
        jt = np.zeros((n, nsteps, 3))
        et = np.zeros((n, nsteps))
        nt = np.zeros((n, nsteps))
        self._check(lib().salmon_gpu_batch_rt_run(self._h, _p(ac), int(nsteps), float(dt), _p(jt), _p(et), _p(nt)),
                    "batch_rt_run")
        return dict(J=jt, E=et, norm=nt)

    def rt_run_ms(self, ac, nsteps, dt):
        """nsteps steps with one A(t) per system: ac (nsys, nsteps+1, 3)."""
        n = len(self.contexts)
        ac = _f64(np.asarray(ac)[:, : nsteps + 1])
        if ac.shape != (n, nsteps + 1, 3):
            raise SalmonGPUError(f"ac must have shape (nsys, nsteps+1, 3) = {(n, nsteps + 1, 3)}")
        jt = np.zeros((n, nsteps, 3))
        et = np.zeros((n, nsteps))
        nt = np.zeros((n, nsteps))
        self._check(lib().salmon_gpu_batch_rt_run_ms(self._h, _p(ac), int(nsteps), float(dt), _p(jt), _p(et),
                                                     _p(nt)), "batch_rt_run_ms")
        return dict(J=jt, E=et, norm=nt)

    def close(self):
        if self._h:
            lib().salmon_gpu_batch_finalize(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
