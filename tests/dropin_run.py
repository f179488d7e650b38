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


def main(case, which, nsteps):
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
