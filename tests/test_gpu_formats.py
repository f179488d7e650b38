"""The on-disk formats around the GPU RT path: ground-state orbitals from
ARTED's gs_wfn_k files into a context, an RT run, and the rt.data /
rt_energy.data files read back as the reference testsuite reads them
(np.loadtxt, row = iteration), against the oracle's run of the same
orbitals.  Tolerance: the files' 16 significant digits (1e-13 relative)
on top of the RT path's FP64 parity (J 1e-10, E 1e-11)."""
import dataclasses

import numpy as np
import pytest

from helpers import load_case, rel
from oracle.cpu import Oracle, kac_at
from salmon_amd import formats, gpu, synth

pytestmark = pytest.mark.gpu


def test_gs_wfn_in_rt_data_out(tmp_path):
    s, _ = load_case("small_k")
    formats.write_gs_wfn_k(str(tmp_path), s.zu, s.occ)
    zu, occ = formats.read_gs_wfn_k(str(tmp_path), s.NL, s.NB, s.NK)
    s2 = dataclasses.replace(s, zu=zu, occ=occ)
    nsteps = 12
    # A on iterations -1..nsteps+1, as the reference's Ac arrays
    ac_full = np.vstack([np.zeros((1, 3)), synth.acos2_pulse(nsteps + 1, s.dt, 0.05, 0.3, 2.0)])
    ac = ac_full[1:]  # iterations 0..nsteps+1
    print(f"[gpu] formats small_k nsteps={nsteps}", flush=True)
    with gpu.Context(s2, gpu.FP64) as c:
        j0 = c.current(kac_at(s, ac[0]))
        e0 = sum(c.total_energy(kac_at(s, ac[0])))
        r = c.rt_run(ac, nsteps)
    javt = np.vstack([j0, r["J"]])
    eall = np.concatenate([[e0], r["E"]])
    formats.write_rt_data(str(tmp_path / "Si_rt.data"), s.dt, ac_full, ac_full, javt)
    formats.write_rt_energy_data(str(tmp_path / "Si_rt_energy.data"), s.dt, eall, e0, nstep_energy_calc=4)
    ref = Oracle(s).rt_run(ac, nsteps)
    o = Oracle(s)
    jr = np.vstack([o.current(kac_at(s, ac[0]), s.zu), ref["J"]])
    er = np.concatenate([[sum(o.energy(kac_at(s, ac[0]), s.zu))], ref["E"]])
    data = np.loadtxt(tmp_path / "Si_rt.data")
    de = np.loadtxt(tmp_path / "Si_rt_energy.data")
    print(f"[gpu] formats: J {rel(data[:, 13:16], jr):.2e} dE {rel(de[:, 2], er[::4] - er[0]):.2e}", flush=True)
    assert data.shape == (nsteps + 1, 16)
    assert rel(data[:, 13:16], jr) < 1e-10
    assert np.allclose(data[:, 7:10], ac[: nsteps + 1], rtol=1e-14, atol=0)  # Ac_tot x, y, z
    assert np.allclose(de[:, 1], er[::4], rtol=1e-11, atol=0)
    assert np.max(np.abs(de[:, 2] - (er[::4] - er[0]))) < 1e-11 * np.max(np.abs(er))
