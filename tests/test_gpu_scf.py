"""Self-consistent local potential on the GPU: the Hartree potential
(ARTED/common/Hartree.f90) and the PZ81 LDA (src/xc/builtin_pz.f90) against
the oracle (itself pinned by the plane-wave closed form and by the
reference's own exc_cor_pz), and the RT loop that updates Vloc after every
step against the oracle doing the same in the reference's order
(dt_evolve.f90:106-216 with no_update_func = 'n', then current and energy).

Tolerances: FP64 1e-12 relative on V_H and V_xc (summation order of the
DFT passes) and on the orbitals after the steps; J 1e-10, E 1e-11 (as the
frozen-potential RT tests); complex64 J within 1e-5 of FP64.
"""
import dataclasses

import numpy as np
import pytest

from helpers import load_case, rel
from oracle import cpu
from oracle.cpu import Oracle, kac_at, midpoint_kac
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["small_k", "noncubic", "cfg1"])
def test_hartree_xc_matches_oracle(name):
    s, _ = load_case(name)
    rng = np.random.default_rng(5)
    rho = 0.02 + rng.random(s.NL) * 0.3
    with gpu.Context(s, gpu.FP64) as c:
        vh, vx = c.hartree_xc(rho, gpu.XC_PZ)
        vh0, vx0 = c.hartree_xc(rho, gpu.XC_NONE)
    print(f"[gpu] hartree_xc {name}: vh {rel(vh, cpu.hartree(s, rho)):.2e} vxc {rel(vx, cpu.pz(rho)[0]):.2e}", flush=True)
    assert rel(vh, cpu.hartree(s, rho)) < 1e-12
    assert rel(vx, cpu.pz(rho)[0]) < 1e-13
    assert np.array_equal(vh0, vh) and not np.any(vx0)


def _scf_oracle(s, vpsl, ac, nsteps, etrs=False):
    """The RT loop with the potential refreshed after every step; ETRS as
    dt_evolve.f90:218-348 (history seeded with V(0), control_sc.f90:82)."""
    z = s.zu.copy()

    def potential(z):
        rho = Oracle(s).density(z)
        return cpu.hartree(s, rho) + vpsl + cpu.pz(rho)[0]

    def with_v(v):
        return Oracle(dataclasses.replace(s, vloc=v))

    v = potential(z)
    old1, old2 = v.copy(), v.copy()
    J, E = [], []
    for n in range(nsteps):
        if etrs:
            z = with_v(v).hamiltonian(kac_at(s, ac[n]), z, 0.5 * s.dt)
            v, old1, old2 = 3 * v - 3 * old1 + old2, v, old1
            z = with_v(v).hamiltonian(kac_at(s, ac[n + 1]), z, 0.5 * s.dt)
        else:
            z = with_v(v).hamiltonian(midpoint_kac(s, ac, n), z, s.dt)
        v = potential(z)
        o = with_v(v)
        k1 = kac_at(s, ac[n + 1])
        J.append(o.current(k1, z))
        E.append(sum(o.energy(k1, z)))
    return z, np.array(J), np.array(E)


@pytest.mark.parametrize("prop", ["middlepoint", "etrs"])
@pytest.mark.parametrize("name,path", [("cfg1", gpu.PATH_AUTO), ("small_k", gpu.PATH_AUTO),
                                       ("cfg1", gpu.PATH_STREAMED)])
def test_scf_rt_matches_oracle(name, path, prop):
    """RT steps with Vloc = V_H + Vpsl + V_xc refreshed after each step, on
    the fused (16^3) and the streamed paths, FP64, both propagators (ETRS
    over four steps so that the extrapolation sees a full history)."""
    s, _ = load_case(name)
    vpsl = s.vloc.copy()
    nsteps = 4 if prop == "etrs" else 3
    ac = synth.acos2_pulse(nsteps + 1, s.dt, 0.05, 0.3, 2.0)
    zr, Jr, Er = _scf_oracle(s, vpsl, ac, nsteps, etrs=prop == "etrs")
    with gpu.Context(s, gpu.FP64) as c:
        c.set_path(path)
        c.set_propagator(prop)
        c.set_scf(vpsl, gpu.XC_PZ)
        r = c.rt_run(ac, nsteps)
        z = c.get_orbitals()
    print(f"[gpu] scf {name} path={path} {prop}: zu {rel(z, zr):.2e} J {rel(r['J'], Jr):.2e} E {rel(r['E'], Er):.2e}",
          flush=True)
    assert rel(z, zr) < 1e-12
    assert rel(r["J"], Jr) < 1e-10
    assert rel(r["E"], Er) < 1e-11


@pytest.mark.parametrize("prop", ["middlepoint", "etrs"])
def test_scf_mixed_close_to_fp64(prop):
    s, _ = load_case("cfg1")
    vpsl = s.vloc.copy()
    nsteps = 20
    ac = synth.acos2_pulse(nsteps + 1, s.dt, 0.05, 0.3, 2.0)
    out = {}
    for prec in (gpu.FP64, gpu.MIXED):
        with gpu.Context(s, prec) as c:
            c.set_propagator(prop)
            c.set_scf(vpsl, gpu.XC_PZ)
            out[prec] = c.rt_run(ac, nsteps)
    with gpu.Context(s, gpu.FP64) as c:
        c.set_propagator(prop)
        frozen = c.rt_run(ac, nsteps)
    J64 = out[gpu.FP64]["J"]
    jerr = np.max(np.abs(out[gpu.MIXED]["J"] - J64)) / np.max(np.abs(J64))
    dfro = np.max(np.abs(frozen["J"] - J64)) / np.max(np.abs(J64))
    print(f"[gpu] scf {prop} mixed J err {jerr:.2e}; scf vs frozen {dfro:.2e}", flush=True)
    assert jerr < 1e-5
    assert dfro > 1e-6  # the potential does move
    # with re-orthonormalisation every 7 steps (the observables of those
    # steps are taken before it) both precisions agree the same way
    re = {}
    for prec in (gpu.FP64, gpu.MIXED):
        with gpu.Context(s, prec) as c:
            c.set_propagator(prop)
            c.set_scf(vpsl, gpu.XC_PZ)
            re[prec] = c.rt_run(ac, nsteps, reortho_every=7)
    jerr_re = np.max(np.abs(re[gpu.MIXED]["J"] - re[gpu.FP64]["J"])) / np.max(np.abs(re[gpu.FP64]["J"]))
    print(f"[gpu] scf {prop} reortho 7: mixed J err {jerr_re:.2e}", flush=True)
    assert jerr_re < 1e-5
    assert rel(re[gpu.FP64]["J"][:7], J64[:7]) < 1e-13  # identical up to the first re-orthonormalisation


def test_scf_errors_and_switch_off():
    s, _ = load_case("small_k")
    with gpu.Context(s, gpu.FP64) as c:
        with pytest.raises(gpu.SalmonGPUError, match="xc"):
            c.set_scf(s.vloc, 5)
        with pytest.raises(gpu.SalmonGPUError, match="xc"):
            c.hartree_xc(np.ones(s.NL), -1)
        # switched off again: the frozen-potential loop with the last Vloc,
        # here V(0) = V_H[rho0] + Vpsl + V_xc[rho0]
        c.set_scf(s.vloc, gpu.XC_PZ)
        c.set_scf(None)
        ac = synth.acos2_pulse(3, s.dt, 0.05, 0.3, 2.0)
        r = c.rt_run(ac, 2)
    rho0 = Oracle(s).density(s.zu)
    v0 = cpu.hartree(s, rho0) + s.vloc + cpu.pz(rho0)[0]
    ref = Oracle(dataclasses.replace(s, vloc=v0)).rt_run(ac, 2)
    assert rel(r["J"], ref["J"]) < 1e-10
