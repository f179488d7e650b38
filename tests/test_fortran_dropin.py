"""Drop-in check: the reference's own Fortran, with hamiltonian(zu,
flag_current), CG_omp(iter_cg_max) and Gram_Schmidt replaced at link time by
the GPU shims (salmon_amd/fortran/arted_gpu_hamiltonian.f90, arted_gpu_gs.f90),
reproduces the all-CPU reference run.  Needs oracle/_ref (built where the
reference tree exists)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(case, which, nsteps):
    out = subprocess.run([sys.executable, os.path.join(HERE, "dropin_run.py"), case, which, str(nsteps)],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-3000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


@pytest.mark.gpu
@pytest.mark.parametrize("case,nsteps", [("small_k", 6), ("cfg1", 20)])
def test_reference_loop_with_gpu_hamiltonian(case, nsteps):
    if not (os.path.exists(ref.LIB) and os.path.exists(ref.LIB_GPU)):
        pytest.skip("reference build (oracle/_ref) not present")
    print(f"[gpu] drop-in {case} {nsteps} steps", flush=True)
    cpu = _run(case, "cpu", nsteps)
    gpu = _run(case, "gpu", nsteps)
    assert gpu["calls"] == nsteps
    J64, Jg = np.array(cpu["J"]), np.array(gpu["J"])
    assert np.max(np.abs(Jg - J64)) / np.max(np.abs(J64)) < 1e-11
    assert np.max(np.abs(np.array(gpu["E"]) - cpu["E"])) / np.max(np.abs(cpu["E"])) < 1e-12
    assert abs(gpu["zabs"] - cpu["zabs"]) / cpu["zabs"] < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small_k", "noncubic"])
def test_reference_gram_schmidt_and_cg_with_gpu_shims(case):
    if not (os.path.exists(ref.LIB) and os.path.exists(ref.LIB_GPU)):
        pytest.skip("reference build (oracle/_ref) not present")
    print(f"[gpu] drop-in ground state {case}", flush=True)
    cpu = _run(case, "cpu-gs", 4)
    gpu = _run(case, "gpu-gs", 4)
    assert gpu["calls"] == 2
    for key in ("zg", "z"):
        a, b = np.array(cpu[key]), np.array(gpu[key])
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))
    a, b = np.array(cpu["esp_var"]), np.array(gpu["esp_var"])
    assert np.max(np.abs(a - b)) < 1e-11 * np.max(np.abs(a))
