"""Drop-in check: the reference's own Fortran RT loop, with its
hamiltonian(zu, flag_current) replaced at link time by the GPU shim
(salmon_amd/fortran/arted_gpu_hamiltonian.f90), reproduces the all-CPU
reference run.  Needs oracle/_ref (built where the reference tree exists)."""
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
