"""Drop-in check: the reference's own Fortran, with its hot-path files
replaced at link time by the GPU drop-ins of salmon_amd/fortran --
hamiltonian(zu, flag_current), the RT current through current_acc_impl,
psi_rho_RT/psi_rho_GS, Total_Energy_omp, module hpsi (hpsi_omp_KB_GS/RT/base),
CG_omp(iter_cg_max) and Gram_Schmidt -- reproduces the all-CPU reference run.
The orbitals stay on the GPU through the whole step loop: one upload at the
first hamiltonian call and one download when the host reads them back.
Needs oracle/_ref (built where the reference tree exists).

Tolerances: FP64, the GPU's summation order against the reference's OpenMP
one: 1e-11 relative on J(t), 1e-12 on E(t), orbitals, densities and
operator outputs.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(case, which, n):
    out = subprocess.run([sys.executable, os.path.join(HERE, "dropin_run.py"), case, which, str(n)],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-3000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def _need_ref():
    if not (os.path.exists(ref.LIB) and os.path.exists(ref.LIB_GPU)):
        pytest.skip("reference build (oracle/_ref) not present")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.gpu
@pytest.mark.parametrize("case,nsteps", [("small_k", 6), ("cfg1", 20), ("grid8", 6)])
def test_reference_loop_with_gpu_dropins(case, nsteps):
    _need_ref()
    print(f"[gpu] drop-in {case} {nsteps} steps", flush=True)
    cpu = _run(case, "cpu", nsteps)
    gpu = _run(case, "gpu", nsteps)
    assert gpu["calls"] == nsteps
    assert gpu["transfers"] == 2  # orbitals uploaded once, read back once
    print(f"[gpu] drop-in {case}: J {_rel(gpu['J'], cpu['J']):.2e} E {_rel(gpu['E'], cpu['E']):.2e} "
          f"z {_rel(gpu['z'], cpu['z']):.2e}", flush=True)
    assert _rel(gpu["J"], cpu["J"]) < 1e-11
    assert _rel(gpu["E"], cpu["E"]) < 1e-12
    assert _rel(gpu["z"], cpu["z"]) < 1e-12
    assert abs(gpu["zabs"] - cpu["zabs"]) / cpu["zabs"] < 1e-12


def _moved_systems(case, nsteps):
    """The systems of the moving-atom run of dropin_run.py: atom 1 displaced
    before every step from the third on, with its tables and local potential."""
    import dataclasses
    from helpers import load_case
    from salmon_amd import synth
    s, d = load_case(case)
    coeff = np.where(s.iuV.reshape(s.NI, -1) > 0, 1.0, -1.0)
    out = []
    for it in range(nsteps):
        if it < 2:
            out.append(s)
            continue
        Rion = s.Rion.copy()
        Rion[0] += (it - 1) * np.array([0.19, -0.11, 0.07])
        t = synth.projector_tables(Rion, s.rc, (s.NLx, s.NLy, s.NLz), (s.aLx, s.aLy, s.aLz), 0.8, coeff)
        out.append(dataclasses.replace(s, Rion=Rion, Mps=t[0], Jxyz=t[1], Jxx=t[2], Jyy=t[3], Jzz=t[4], uV=t[6],
                                       vloc=s.vloc * (1.0 + 0.01 * it)))
    return out, d


@pytest.mark.parametrize("case", ["small_k", "noncubic"])
def test_reference_moving_atom_matches_oracle(case):
    """Pins the harness's moved-atom state (ref_move_atoms: the tables
    prep_ps_periodic('update_all') leaves): the reference's step loop with atom
    1 moving every step from the third on equals the oracle run step by step
    on each step's system.  Atom 1's sphere grows past Nps in small_k (the
    arrays are reallocated) and shrinks in noncubic.  CPU only."""
    _need_ref()
    from oracle.cpu import Oracle
    nsteps = 6
    r = _run(case, "cpu-md", nsteps)
    systems, d = _moved_systems(case, nsteps)
    assert [list(x.Mps) for x in systems[2:]] == r["mps"]
    if case == "small_k":
        assert max(systems[-1].Mps) > systems[0].Nps
    zu = systems[0].zu
    J, E = [], []
    for it, s in enumerate(systems):
        o = Oracle(s)
        x = o.rt_run(d["ac"][it:it + 2], 1, zu=zu)
        zu = x["zu"]
        J.append(x["J"][0])
        E.append(x["E"][0])
    z = [zu.real.ravel()[::97], zu.imag.ravel()[::97]]
    print(f"moving atom {case}: J {_rel(r['J'], J):.2e} E {_rel(r['E'], E):.2e} z {_rel(r['z'], z):.2e}")
    assert _rel(r["J"], J) < 1e-11
    assert _rel(r["E"], E) < 1e-12
    assert _rel(r["z"], z) < 1e-12
    # the moved atom changes the trajectory
    plain = Oracle(systems[0]).rt_run(d["ac"], nsteps)
    assert _rel(r["J"], plain["J"]) > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small_k", "noncubic"])
def test_reference_loop_with_gpu_dropins_moving_atom(case):
    """An Ehrenfest-style run: the drop-in hamiltonian re-uploads the
    projector tables every step (use_ehrenfest_md='y') while atom 1 moves, the
    orbitals staying on the GPU; against the all-CPU reference loop."""
    _need_ref()
    nsteps = 6
    print(f"[gpu] drop-in moving atom {case} {nsteps} steps", flush=True)
    cpu = _run(case, "cpu-md", nsteps)
    gpu = _run(case, "gpu-md", nsteps)
    assert gpu["calls"] == nsteps
    assert gpu["transfers"] == 2
    assert gpu["mps"] == cpu["mps"]
    print(f"[gpu] drop-in moving atom {case}: J {_rel(gpu['J'], cpu['J']):.2e} E {_rel(gpu['E'], cpu['E']):.2e} "
          f"z {_rel(gpu['z'], cpu['z']):.2e}", flush=True)
    assert _rel(gpu["J"], cpu["J"]) < 1e-11
    assert _rel(gpu["E"], cpu["E"]) < 1e-12
    assert _rel(gpu["z"], cpu["z"]) < 1e-12
    assert abs(gpu["zabs"] - cpu["zabs"]) / cpu["zabs"] < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small_k", "noncubic", "grid8"])
def test_reference_observables_with_gpu_dropins(case):
    """hpsi_omp_KB_GS, sp_energy_omp (through the module hpsi drop-in),
    current_RT, psi_rho_RT and Total_Energy_omp one call each."""
    _need_ref()
    cpu = _run(case, "cpu-obs", 0)
    gpu = _run(case, "gpu-obs", 0)
    errs = {k: _rel(gpu[k], cpu[k]) for k in ("hpsi", "tt", "esp", "J", "rho", "ekin", "enl")}
    print(f"[gpu] drop-in observables {case}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert errs["J"] < 1e-11
    assert all(v < 1e-12 for k, v in errs.items() if k != "J")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small_k", "noncubic"])
def test_reference_gram_schmidt_and_cg_with_gpu_shims(case):
    _need_ref()
    print(f"[gpu] drop-in ground state {case}", flush=True)
    cpu = _run(case, "cpu-gs", 4)
    gpu = _run(case, "gpu-gs", 4)
    assert gpu["calls"] == 2
    for key in ("zg", "z"):
        a, b = np.array(cpu[key]), np.array(gpu[key])
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))
    a, b = np.array(cpu["esp_var"]), np.array(gpu["esp_var"])
    assert np.max(np.abs(a - b)) < 1e-11 * np.max(np.abs(a))


@pytest.mark.gpu
def test_dropin_step_time_config2():
    """The reference's step loop at the config-2 size with the GPU drop-ins:
    ms per step, printed beside the device-resident salmon_gpu_rt_run (FP64
    streamed path) of the same workload; one orbital upload in the whole run."""
    _need_ref()
    r = _run("cfg2", "gpu-time", 5)
    from salmon_amd import gpu, synth
    s, ac = synth.config2(nsteps=7, orbitals=False)
    s.zu = synth.random_orbitals(np.random.default_rng(7), s.NK, s.NB, s.NL, s.Hxyz, orthonormalise=False)
    with gpu.Context(s, gpu.FP64) as c:
        c.set_path(gpu.PATH_STREAMED)
        c.rt_run(ac[:2], 1)
        import time
        c.synchronize()
        t0 = time.perf_counter()
        c.rt_run(ac[1:7], 5)
        ms = (time.perf_counter() - t0) / 5 * 1e3
    print(f"[gpu] drop-in step loop at config 2: {r['ms_per_step']:.2f} ms/step ({r['threads']} host threads), "
          f"salmon_gpu_rt_run streamed FP64 {ms:.2f} ms/step, ratio {r['ms_per_step'] / ms:.2f}", flush=True)
    assert r["transfers"] == 1
    assert np.all(np.isfinite(r["J"]))
