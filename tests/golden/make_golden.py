"""Writes the golden fixtures of this directory from the reference's own
Fortran (oracle/_ref/libsalmon_ref.so, compiled from the SALMON sources by
oracle/ref/Makefile).  Run in the build container, where the reference tree
exists:

    python tests/golden/make_golden.py [case ...]

Each case runs in a subprocess because the reference keeps its state in
module globals (one system per process).  A fixture holds the inputs (the
full synthetic system) and the reference outputs; nothing of the reference's
source is stored.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {
    # 12^3 cell, 4 k points x 3 bands, diamond atoms, Acos2 pulse
    "small_k": dict(kind="make", N=(12, 12, 12), L=8.0, nk3=(2, 1, 2), NB=3, seed=3, atoms="diamond",
                    pulse=dict(nsteps=6, amplitude_A=0.05, omega=0.3, tpulse=2.0)),
    # non-cubic grid and cell, random atoms (overlapping spheres), 2 k x 2 bands
    "noncubic": dict(kind="make", N=(10, 12, 14), L=(7.0, 8.0, 9.0), nk3=(1, 2, 1), NB=2, seed=11, atoms="random",
                     natoms=5, pulse=dict(nsteps=5, amplitude_A=0.1, omega=0.5, tpulse=2.0)),
    # the 16^3 Gamma-point workload with 16 orbitals, 100 steps at A=(0.01,0,0)
    "cfg1": dict(kind="cfg1", seed=0),
    # 8^3 grid on the 10.26 bohr cell (the grid of the reference's ARTED RT
    # testsuite, testsuites/002_bulk_rt), 4 k points x 3 bands, diamond atoms
    "grid8": dict(kind="make", N=(8, 8, 8), L=10.26, nk3=(2, 1, 2), NB=3, seed=8, atoms="diamond",
                  pulse=dict(nsteps=6, amplitude_A=0.05, omega=0.3, tpulse=2.0)),
}


def system_arrays(s):
    return dict(NL3=np.array([s.NLx, s.NLy, s.NLz]), aL=np.array([s.aLx, s.aLy, s.aLz]), lapx=s.lapx, lapy=s.lapy,
                lapz=s.lapz, nabx=s.nabx, naby=s.naby, nabz=s.nabz, vloc=s.vloc, Rion=s.Rion, rc=np.array(s.rc),
                Mps=s.Mps, Jxyz=s.Jxyz, Jxx=s.Jxx, Jyy=s.Jyy, Jzz=s.Jzz, a_tbl=s.a_tbl, uV=s.uV, iuV=s.iuV,
                kAc0=s.kAc0, occ=s.occ, zu=s.zu, dt=np.array(s.dt))


def run_case(name):
    from salmon_amd import synth
    from oracle.ref import Ref
    from oracle.cpu import midpoint_kac, kac_at

    spec = CASES[name]
    if spec["kind"] == "cfg1":
        s, ac = synth.config1(seed=spec["seed"])
        nsteps = 100
    else:
        s = synth.make_system(N=spec["N"], L=spec["L"], nk3=spec["nk3"], NB=spec["NB"], seed=spec["seed"],
                              atoms=spec["atoms"], natoms=spec.get("natoms", 8), name=name)
        p = spec["pulse"]
        nsteps = p["nsteps"]
        ac = synth.acos2_pulse(nsteps, s.dt, p["amplitude_A"], p["omega"], p["tpulse"])
    R = Ref(s)
    out = system_arrays(s)
    out["ac"] = ac
    out["ref_fd"] = R.fd()
    if spec["kind"] != "cfg1":
        rng = np.random.default_rng(100)
        ik = s.NK - 1
        k = s.kAc0[:, ik] + rng.uniform(-0.2, 0.2, 3)
        psi = s.zu[ik, s.NB - 1]
        h, t = R.hpsi(ik, k, psi)
        out.update(hpsi_k=k, hpsi_ik=np.array(ik), hpsi_out=h, hpsi_tt=t)
        kmid = midpoint_kac(s, ac, 0)
        R.put(s.zu)
        R.hamiltonian(kmid)
        z1 = R.get()
        out["taylor_kac"] = kmid
        out["taylor_zu"] = z1
        k1 = kac_at(s, ac[1])
        out["obs_kac"] = k1
        out["current"] = R.current(k1)
        out["rho"] = R.density()
        out["energy"] = np.array(R.energy(k1))
        R.put(s.zu)
        out["esp"] = R.sp_energy(s.kAc0)
        R.put(s.zu)
        R.gram_schmidt()
        out["gs_zu"] = R.get()
        R.put(s.zu)
        out["cg_esp_var"] = R.cg(s.kAc0, 4)
        out["cg_zu"] = R.get()
    R.put(s.zu)
    rr = R.rt_run(ac, nsteps)
    out["rt_J"] = rr["J"]
    out["rt_E"] = rr["E"]
    zf = R.get()
    if spec["kind"] == "cfg1":
        out["rt_band_norm"] = np.sum(np.abs(zf) ** 2, axis=2) * s.Hxyz
        out["rt_zu_head"] = zf[:, :2]
    else:
        out["rt_zu"] = zf
    out["ref_threads"] = np.array(R.num_threads())
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--case":
        run_case(sys.argv[2])
        return
    from oracle import ref
    if not ref.build():
        raise SystemExit("reference build not available")
    names = sys.argv[1:] or list(CASES)  # python make_golden.py [case ...]
    for name in names:
        print("case", name, flush=True)
        subprocess.check_call([sys.executable, os.path.abspath(__file__), "--case", name])


if __name__ == "__main__":
    main()
