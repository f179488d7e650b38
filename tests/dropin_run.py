"""Runs the reference's RT step loop (oracle/_ref harness) on a golden case,
with either the reference's own hamiltonian or the GPU drop-in linked in.
Invoked in a subprocess by test_fortran_dropin.py (module globals: one
system per process).  Prints one JSON line."""
import ctypes
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from helpers import load_case  # noqa: E402
from oracle import ref  # noqa: E402


def ground_state(case, which, ncg):
    """Gram_Schmidt then CG_omp(ncg) through the reference's harness."""
    s, d = load_case(case)
    path = ref.LIB_GPU if which == "gpu" else ref.LIB
    R = ref.Ref(s, lib_path=path)
    R.gram_schmidt()
    zg = R.get()
    ev = R.cg(s.kAc0, ncg)
    z = R.get()
    calls = ctypes.c_int.in_dll(ref.lib(), "arted_gpu_gs_calls").value if which == "gpu" else -1
    print(json.dumps(dict(zg=[zg.real.ravel()[::97].tolist(), zg.imag.ravel()[::97].tolist()],
                          z=[z.real.ravel()[::97].tolist(), z.imag.ravel()[::97].tolist()],
                          esp_var=ev.ravel().tolist(), calls=calls)))


def main(case, which, nsteps):
    if which.endswith("-gs"):
        return ground_state(case, which[:-3], nsteps)
    s, d = load_case(case)
    path = ref.LIB_GPU if which == "gpu" else ref.LIB
    R = ref.Ref(s, lib_path=path)
    r = R.rt_run(d["ac"], nsteps)
    z = R.get()
    calls = -1
    if which == "gpu":
        calls = ctypes.c_int.in_dll(ref.lib(), "arted_gpu_hamiltonian_calls").value
    print(json.dumps(dict(J=r["J"].tolist(), E=r["E"].tolist(), zabs=float(np.abs(z).sum()),
                          z0=[float(z.real.ravel()[7]), float(z.imag.ravel()[7])], calls=calls)))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]))
