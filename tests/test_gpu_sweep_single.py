"""rt_run of a single context through the persistent sweep kernel (the path
SALMON_PATH_AUTO takes off the fused 16^3 shape): against the streamed
kernels and the fused step, with re-orthonormalisation, after new projector
tables, and where the sweep does not apply (small grids, ETRS), where AUTO
falls back and an explicit SALMON_PATH_SWEEP fails.

Tolerance: FP64 1e-12 relative (J, E, norm, density, orbitals), the sweep's
summation order against the streamed kernels' up to FMA contraction."""
import dataclasses

import numpy as np
import pytest

from helpers import load_case, rel
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _run(s, path, ac, nsteps, prec=gpu.FP64, **kw):
    with gpu.Context(s, prec) as c:
        c.set_path(path)
        p = c.path()
        r = c.rt_run(ac, nsteps, **kw)
        r["zu"] = c.get_orbitals()
        return p, r


def _close(a, b, keys=("J", "E", "norm", "rho", "zu"), tol=TOL):
    for k in keys:
        e = rel(a[k], b[k])
        assert e < tol, (k, e)


@pytest.mark.parametrize("name", ["small_k", "noncubic", "grid8"])
def test_sweep_matches_streamed(name):
    s, d = load_case(name)
    n = d["ac"].shape[0] - 1
    print(f"[gpu] single-context sweep {name} {n} steps", flush=True)
    p1, r1 = _run(s, gpu.PATH_AUTO, d["ac"], n)
    p2, r2 = _run(s, gpu.PATH_STREAMED, d["ac"], n)
    assert p1 == gpu.PATH_SWEEP and p2 == gpu.PATH_STREAMED
    _close(r1, r2)


def test_sweep_with_reorthonormalisation():
    s, d = load_case("small_k")
    n = 6
    print("[gpu] single-context sweep, Gram-Schmidt every 2 steps", flush=True)
    _, r1 = _run(s, gpu.PATH_SWEEP, d["ac"], n, reortho_every=2)
    _, r2 = _run(s, gpu.PATH_STREAMED, d["ac"], n, reortho_every=2)
    _close(r1, r2)


def test_sweep_on_the_fused_shape_matches_fused():
    s, d = load_case("cfg1")
    print("[gpu] single-context sweep on 16^3 against the fused step", flush=True)
    p1, r1 = _run(s, gpu.PATH_SWEEP, d["ac"], 8)
    p2, r2 = _run(s, gpu.PATH_AUTO, d["ac"], 8)
    assert p1 == gpu.PATH_SWEEP and p2 == gpu.PATH_FUSED
    _close(r1, r2)


def test_sweep_after_new_projector_tables():
    """Runs, new tables (an atom moved), runs again: the sweep rebuilds its
    descriptors and follows the streamed kernels doing the same."""
    s, d = load_case("noncubic")
    coeff = np.where(s.iuV.reshape(s.NI, -1) > 0, 1.0, -1.0)
    R = s.Rion.copy()
    R[0] += [0.31, -0.17, 0.11]
    t = synth.projector_tables(R, s.rc, (s.NLx, s.NLy, s.NLz), (s.aLx, s.aLy, s.aLz), 0.8, coeff)
    s2 = dataclasses.replace(s, Rion=R, Mps=t[0], Jxyz=t[1], Jxx=t[2], Jyy=t[3], Jzz=t[4], uV=t[6])
    print("[gpu] single-context sweep across a table change", flush=True)
    out = {}
    for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
        with gpu.Context(s, gpu.FP64) as c:
            c.set_path(path)
            a = c.rt_run(d["ac"][:3], 2)
            c.set_projectors(s2)
            b = c.rt_run(d["ac"][2:], 3)
            out[path] = dict(J=np.concatenate([a["J"], b["J"]]), E=np.concatenate([a["E"], b["E"]]),
                             norm=np.concatenate([a["norm"], b["norm"]]), rho=b["rho"], zu=c.get_orbitals())
            assert c.path() == (gpu.PATH_SWEEP if path == gpu.PATH_AUTO else gpu.PATH_STREAMED)
    _close(out[gpu.PATH_AUTO], out[gpu.PATH_STREAMED])


def test_sweep_mixed_close_to_fp64():
    s, d = load_case("small_k")
    n = d["ac"].shape[0] - 1
    _, r64 = _run(s, gpu.PATH_AUTO, d["ac"], n)
    p, r32 = _run(s, gpu.PATH_AUTO, d["ac"], n, prec=gpu.MIXED)
    assert p == gpu.PATH_SWEEP
    _close(r32, r64, keys=("J", "E", "norm"), tol=1e-5)


def test_where_the_sweep_does_not_apply():
    """6^3 (under 8 points per axis): AUTO takes the streamed kernels, an
    explicit SALMON_PATH_SWEEP fails with a message."""
    small = synth.make_system(N=(6, 6, 6), L=6.0, nk3=(1, 1, 1), NB=2, seed=3)
    ac = synth.acos2_pulse(2, 0.08, 0.05, 0.057, 826.8)
    with gpu.Context(small, gpu.FP64) as c:
        assert c.path() == gpu.PATH_STREAMED
        c.set_path(gpu.PATH_SWEEP)
        with pytest.raises(gpu.SalmonGPUError, match="persistent sweep"):
            c.rt_run(ac, 2, dt=0.08)


@pytest.mark.parametrize("prop", [gpu.PROP_MIDPOINT, gpu.PROP_ETRS])
def test_scf_through_the_sweep(prop):
    """The self-consistent potential off 16^3: per step the Taylor-only sweep
    launch(es) (ETRS: with the extrapolated potential between the halves),
    the density, the Hartree + PZ potential from it and the observables-only
    launch with that potential, against the streamed kernels in rt_run_scf's
    order."""
    s, d = load_case("noncubic")
    n = d["ac"].shape[0] - 1
    out = {}
    for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
        with gpu.Context(s, gpu.FP64) as c:
            c.set_propagator(prop)
            c.set_scf(s.vloc, gpu.XC_PZ)
            c.set_path(path)
            assert c.path() == (gpu.PATH_SWEEP if path == gpu.PATH_AUTO else gpu.PATH_STREAMED)
            r = c.rt_run(d["ac"], n, reortho_every=2)
            r["zu"] = c.get_orbitals()
            out[path] = r
    print(f"[gpu] SCF sweep prop={prop}: J {rel(out[gpu.PATH_AUTO]['J'], out[gpu.PATH_STREAMED]['J']):.2e} "
          f"rho {rel(out[gpu.PATH_AUTO]['rho'], out[gpu.PATH_STREAMED]['rho']):.2e}", flush=True)
    _close(out[gpu.PATH_AUTO], out[gpu.PATH_STREAMED])


@pytest.mark.parametrize("name", ["small_k", "noncubic"])
def test_etrs_through_the_sweep(name):
    """ETRS off 16^3: per step two Taylor-only sweep launches over dt/2 (at
    kAc0 + A_n and kAc0 + A_{n+1}) and the observables-only launch, against
    the streamed ETRS, also with Gram-Schmidt every 2 steps."""
    s, d = load_case(name)
    n = d["ac"].shape[0] - 1
    for reortho in (0, 2):
        out = {}
        for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
            with gpu.Context(s, gpu.FP64) as c:
                c.set_propagator(gpu.PROP_ETRS)
                c.set_path(path)
                assert c.path() == (gpu.PATH_SWEEP if path == gpu.PATH_AUTO else gpu.PATH_STREAMED)
                r = c.rt_run(d["ac"], n, reortho_every=reortho)
                r["zu"] = c.get_orbitals()
                out[path] = r
        print(f"[gpu] ETRS sweep {name} reortho={reortho}: J {rel(out[gpu.PATH_AUTO]['J'], out[gpu.PATH_STREAMED]['J']):.2e}",
              flush=True)
        _close(out[gpu.PATH_AUTO], out[gpu.PATH_STREAMED])


@pytest.mark.parametrize("name", ["small_k", "noncubic", "grid8"])
def test_hamiltonian_entry_through_the_sweep(name):
    """The hamiltonian entry (the drop-in's Taylor step at the caller's
    kAc(NK,3)) off 16^3 runs the persistent sweep without observables: three
    calls at per-k wave vectors against the streamed kernels and the oracle,
    then an rt_run continuing from those orbitals."""
    from oracle.cpu import Oracle
    s, d = load_case(name)
    rng = np.random.default_rng(9)
    ks = [s.kAc0 + rng.uniform(-0.05, 0.05, size=s.kAc0.shape) for _ in range(3)]
    out = {}
    for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
        with gpu.Context(s, gpu.FP64) as c:
            c.set_path(path)
            for k in ks:
                c.hamiltonian(k, s.dt)
            z = c.get_orbitals()
            r = c.rt_run(d["ac"][:3], 2)
            r["zu"] = c.get_orbitals()
            out[path] = (z, r)
    o = Oracle(s)
    z = s.zu
    for k in ks:
        z = o.hamiltonian(k, z, s.dt)
    print(f"[gpu] hamiltonian entry {name}: sweep vs streamed {rel(out[gpu.PATH_AUTO][0], out[gpu.PATH_STREAMED][0]):.2e}"
          f" vs oracle {rel(out[gpu.PATH_AUTO][0], z):.2e}", flush=True)
    assert rel(out[gpu.PATH_AUTO][0], z) < TOL
    assert rel(out[gpu.PATH_AUTO][0], out[gpu.PATH_STREAMED][0]) < TOL
    _close(out[gpu.PATH_AUTO][1], out[gpu.PATH_STREAMED][1])


@pytest.mark.parametrize("name", ["small_k", "noncubic", "grid8"])
def test_observables_entries_through_the_sweep(name):
    """current, energy and norm off 16^3 take one observables-only sweep
    launch: against k_observe (the streamed path) at per-k wave vectors,
    before and after a Taylor step."""
    s, d = load_case(name)
    rng = np.random.default_rng(13)
    ks = [s.kAc0 + rng.uniform(-0.05, 0.05, size=s.kAc0.shape) for _ in range(2)]
    out = {}
    for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
        vals = []
        with gpu.Context(s, gpu.FP64) as c:
            c.set_path(path)
            for k in ks:
                vals.append((c.current(k), c.total_energy(k), c.norm()))
                c.hamiltonian(k, s.dt)
            vals.append((c.current(ks[0]), c.total_energy(ks[0]), c.norm()))
        out[path] = vals
    for a, b in zip(out[gpu.PATH_AUTO], out[gpu.PATH_STREAMED]):
        assert rel(a[0], b[0]) < 1e-11
        assert rel(a[1], b[1]) < TOL
        assert rel(a[2], b[2]) < TOL


def test_switching_paths_mid_run():
    """AUTO (sweep), then STREAMED (the sweep's work buffers freed), then AUTO
    again (rebuilt): the same trajectory as the streamed kernels throughout."""
    s, d = load_case("small_k")
    seq = [(gpu.PATH_AUTO, 0, 2), (gpu.PATH_STREAMED, 2, 4), (gpu.PATH_AUTO, 4, 6)]
    out = {}
    for mode in ("switch", "streamed"):
        J = []
        with gpu.Context(s, gpu.FP64) as c:
            for path, a, b in seq:
                c.set_path(path if mode == "switch" else gpu.PATH_STREAMED)
                J.append(c.rt_run(d["ac"][a:b + 1], b - a)["J"])
            out[mode] = dict(J=np.concatenate(J), zu=c.get_orbitals())
    _close(out["switch"], out["streamed"], keys=("J", "zu"))


@pytest.mark.parametrize("scf,prop", [(False, gpu.PROP_ETRS), (True, gpu.PROP_MIDPOINT), (True, gpu.PROP_ETRS)])
def test_mixed_step_by_step_sweep_close_to_fp64(scf, prop):
    """complex64 ETRS and self-consistent runs on the sweep (step by step)
    within 1e-5 of FP64 on the same path (J, E, norm)."""
    s, d = load_case("noncubic")
    n = d["ac"].shape[0] - 1
    out = {}
    for prec in (gpu.FP64, gpu.MIXED):
        with gpu.Context(s, prec) as c:
            c.set_propagator(prop)
            if scf:
                c.set_scf(s.vloc, gpu.XC_PZ)
            assert c.path() == gpu.PATH_SWEEP
            out[prec] = c.rt_run(d["ac"], n)
    print(f"[gpu] mixed sweep scf={scf} prop={prop}: J {rel(out[gpu.MIXED]['J'], out[gpu.FP64]['J']):.2e}", flush=True)
    _close(out[gpu.MIXED], out[gpu.FP64], keys=("J", "E", "norm"), tol=1e-5)
