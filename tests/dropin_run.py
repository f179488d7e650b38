"""Runs the reference's own Fortran (oracle/_ref harness) on a golden case,
either all-CPU (libsalmon_ref.so) or with the GPU drop-ins of
salmon_amd/fortran linked in place of the files they replace
(libsalmon_ref_gpu.so: hamiltonian, current_acc_impl, psi_rho_RT/GS,
Total_Energy_omp, module hpsi, CG_omp, Gram_Schmidt).  Invoked in a
subprocess by test_fortran_dropin.py (module globals: one system per
process).  Prints one JSON line.

  python dropin_run.py CASE cpu|gpu STEPS          the RT step loop
  python dropin_run.py CASE cpu-gs|gpu-gs NCG      Gram_Schmidt then CG_omp
  python dropin_run.py CASE cpu-obs|gpu-obs 0      hpsi, sp_energy, current, density, energy
  python dropin_run.py CASE cpu-md|gpu-md STEPS    the RT step loop, atom 1 moving from step 3 on
  python dropin_run.py cfg2 gpu-time|cpu-time STEPS  config-2 step time of the loop
"""
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from helpers import load_case  # noqa: E402
from oracle import ref  # noqa: E402


def _lib_path(which):
    return ref.LIB_GPU if which.startswith("gpu") else ref.LIB


def _counter(name, which):
    return ctypes.c_int.in_dll(ref.lib(), name).value if which.startswith("gpu") else -1


def _sample(z):
    return [z.real.ravel()[::97].tolist(), z.imag.ravel()[::97].tolist()]


def ground_state(case, which, ncg):
    """Gram_Schmidt then CG_omp(ncg) through the reference's harness."""
    s, d = load_case(case)
    R = ref.Ref(s, lib_path=_lib_path(which))
    R.gram_schmidt()
    zg = R.get()
    ev = R.cg(s.kAc0, ncg)
    z = R.get()
    print(json.dumps(dict(zg=_sample(zg), z=_sample(z), esp_var=ev.ravel().tolist(),
                          calls=_counter("arted_gpu_gs_calls", which))))


def observables(case, which):
    """One call of each observable and operator on the golden orbitals."""
    from oracle.cpu import kac_at
    s, d = load_case(case)
    R = ref.Ref(s, lib_path=_lib_path(which))
    ik = s.NK - 1
    h, t = R.hpsi(ik, d["hpsi_k"], s.zu[ik, s.NB - 1])
    esp = R.sp_energy(s.kAc0)
    k1 = kac_at(s, d["ac"][1])
    j = R.current(k1)
    rho = R.density()
    ek, en = R.energy(k1)
    print(json.dumps(dict(hpsi=_sample(h), tt=_sample(t), esp=esp.ravel().tolist(), J=j.tolist(),
                          rho=rho.tolist(), ekin=ek, enl=en, transfers=_counter("arted_gpu_orbital_transfers", which))))


def moving_atom(case, which, nsteps):
    """The RT step loop with the first atom moving every step from the third
    on, as in an Ehrenfest run: before each such step the harness takes the
    new position's sphere tables, uV and local potential (the state
    prep_ps_periodic('update_all') leaves), and the drop-in re-uploads them."""
    from salmon_amd import synth
    s, d = load_case(case)
    R = ref.Ref(s, lib_path=_lib_path(which))
    ac = d["ac"]
    coeff = np.where(s.iuV.reshape(s.NI, -1) > 0, 1.0, -1.0)
    J, E, mps = [], [], []
    for it in range(nsteps):
        if it >= 2:
            Rion = s.Rion.copy()
            Rion[0] += (it - 1) * np.array([0.19, -0.11, 0.07])
            t = synth.projector_tables(Rion, s.rc, (s.NLx, s.NLy, s.NLz), (s.aLx, s.aLy, s.aLz), 0.8, coeff)
            assert np.array_equal(t[5], s.a_tbl) and np.array_equal(t[7], s.iuV)
            R.move_atoms(Rion, t[0], t[1], t[2], t[3], t[4], t[6], s.vloc * (1.0 + 0.01 * it))
            mps.append(t[0].tolist())
        r = R.rt_run(ac[it:it + 2], 1)
        J.append(r["J"][0].tolist())
        E.append(float(r["E"][0]))
    z = R.get()
    print(json.dumps(dict(J=J, E=E, z=_sample(z), zabs=float(np.abs(z).sum()), mps=mps,
                          calls=_counter("arted_gpu_hamiltonian_calls", which),
                          transfers=_counter("arted_gpu_orbital_transfers", which))))


def timing(which, nsteps):
    """The reference's RT step loop at the config-2 size (16^3, 512 k x 32
    bands): ms per step after one warm-up step."""
    from salmon_amd import synth
    s, ac = synth.config2(nsteps=nsteps + 1, orbitals=False)
    s.zu = synth.random_orbitals(np.random.default_rng(7), s.NK, s.NB, s.NL, s.Hxyz, orthonormalise=False)
    R = ref.Ref(s, lib_path=_lib_path(which))
    R.rt_run(ac[:2], 1)
    t0 = time.perf_counter()
    r = R.rt_run(ac[1:nsteps + 2], nsteps)
    t = time.perf_counter() - t0
    print(json.dumps(dict(ms_per_step=t / nsteps * 1e3, steps=nsteps, J=r["J"].tolist(),
                          threads=R.num_threads(), transfers=_counter("arted_gpu_orbital_transfers", which))))


def main(case, which, n):
    if which.endswith("-gs"):
        return ground_state(case, which[:-3], n)
    if which.endswith("-obs"):
        return observables(case, which)
    if which.endswith("-time"):
        return timing(which, n)
    if which.endswith("-md"):
        return moving_atom(case, which, n)
    s, d = load_case(case)
    R = ref.Ref(s, lib_path=_lib_path(which))
    r = R.rt_run(d["ac"], n)
    z = R.get()
    print(json.dumps(dict(J=r["J"].tolist(), E=r["E"].tolist(), zabs=float(np.abs(z).sum()), z=_sample(z),
                          calls=_counter("arted_gpu_hamiltonian_calls", which),
                          transfers=_counter("arted_gpu_orbital_transfers", which))))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]))
