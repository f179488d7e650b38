"""Batched independent systems on the GPU: every system of a batch (different
grids, bands, k points and atoms; the persistent sweep kernel and the
one-launch-per-phase kernel with autotuned z blocking, the step replayed from
a captured graph) against the same system run alone through the
single-context streamed path.

Tolerance: FP64 1e-12 relative on J(t), E(t) and the norms (the batched
stencil keeps the single-context summation order; only FMA contraction may
differ); complex64 1e-5 (max_t |dJ| / max_t |J|).
"""
import numpy as np
import pytest

from helpers import rel
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu

SYSTEMS = [0, 1, 2, 3, 5, 9]  # shapes 12^3, 20x20x28, 32^3, 16^3, 16^3, ...


def _systems():
    return [synth.config4_system(i, scale=0.12) for i in SYSTEMS]


def _ac(n):
    return synth.acos2_pulse(n, 0.08, 0.05, 0.057, 826.8)


@pytest.mark.parametrize("kernel", [gpu.BATCH_SWEEP, gpu.BATCH_PHASES])
@pytest.mark.parametrize("prec", [gpu.FP64, gpu.MIXED])
def test_batch_matches_single_contexts(prec, kernel):
    nsteps = 6
    ac = _ac(nsteps)
    systems = _systems()
    print(f"[gpu] batch prec={prec} kernel={kernel}: "
          + ", ".join(f"{s.NLx}x{s.NLy}x{s.NLz}/{s.NK}k/{s.NB}b" for s in systems), flush=True)
    singles = []
    for s in systems:
        with gpu.Context(s, prec) as c:
            c.set_path(gpu.PATH_STREAMED)
            singles.append(c.rt_run(ac, nsteps, dt=0.08))
    ctxs = [gpu.Context(s, prec) for s in systems]
    with gpu.Batch(ctxs, prec) as b:
        b.set_kernel(kernel)
        zb = b.tune()
        assert b.kernel() == kernel
        print(f"[gpu] batch z blocking {zb}", flush=True)
        assert all(s.NLz % z == 0 for s, z in zip(systems, zb))
        r1 = b.rt_run(ac[: 3 + 1], 3, 0.08)
        r2 = b.rt_run(ac[3:], 3, 0.08)  # continues from the resident orbitals
    J = np.concatenate([r1["J"], r2["J"]], axis=1)
    E = np.concatenate([r1["E"], r2["E"]], axis=1)
    N = np.concatenate([r1["norm"], r2["norm"]], axis=1)
    for i, r in enumerate(singles):
        tol = 1e-12 if prec == gpu.FP64 else 1e-5
        ej = np.max(np.abs(J[i] - r["J"])) / np.max(np.abs(r["J"]))
        print(f"[gpu] batch system {i}: J {ej:.2e} E {rel(E[i], r['E']):.2e} norm {rel(N[i], r['norm']):.2e}",
              flush=True)
        assert ej < tol
        assert rel(E[i], r["E"]) < tol
        assert rel(N[i], r["norm"]) < tol
    for c in ctxs:
        c.close()


def test_batch_orbitals_match_single_context():
    """The batch advances the contexts' resident orbitals exactly as a single
    context does."""
    nsteps = 2
    ac = _ac(nsteps)
    s = synth.config4_system(1, scale=0.12)
    with gpu.Context(s, gpu.FP64) as c:
        c.set_path(gpu.PATH_STREAMED)
        c.rt_run(ac, nsteps, dt=0.08)
        z1 = c.get_orbitals()
    with gpu.Context(s, gpu.FP64) as c:
        with gpu.Batch([c], gpu.FP64) as b:
            b.rt_run(ac, nsteps, 0.08)
        z2 = c.get_orbitals()
    assert rel(z2, z1) < 1e-13


def test_batch_sweep_needs_eight_points_per_axis():
    """A 6^3 grid is outside the persistent sweep: asking for the sweep fails
    with a message, the automatic choice runs it with the phase kernel."""
    s = synth.make_system(N=(6, 6, 6), L=6.0, nk3=(1, 1, 1), NB=2, seed=3)
    ac = _ac(2)
    with gpu.Context(s, gpu.FP64) as c:
        c.set_path(gpu.PATH_STREAMED)
        ref = c.rt_run(ac, 2, dt=0.08)
    with gpu.Context(s, gpu.FP64) as c:
        with gpu.Batch([c], gpu.FP64) as b:
            b.set_kernel(gpu.BATCH_SWEEP)
            with pytest.raises(gpu.SalmonGPUError, match="persistent sweep"):
                b.tune()
            b.set_kernel(gpu.BATCH_AUTO)
            r = b.rt_run(ac, 2, 0.08)
            assert b.kernel() == gpu.BATCH_PHASES
    assert rel(r["J"][0], ref["J"]) < 1e-12 and rel(r["E"][0], ref["E"]) < 1e-12


@pytest.mark.parametrize("prec", [gpu.FP64, gpu.MIXED])
def test_batch_sweep_odd_grids_against_single_contexts(prec):
    """The persistent sweep on grids outside the 64-system workload: odd and
    unequal axes (columns of 1, 2 and 3 z points, teams of 1 and 2 waves,
    idle threads in a team), random atoms in non-cubic cells; two runs in a
    row (each closed by its observables-only launch, the second continuing
    from the first's orbitals) against the single-context streamed path."""
    systems = [synth.make_system(N=(9, 10, 11), L=(6.0, 6.5, 7.0), nk3=(2, 1, 1), NB=3, seed=21, atoms="random",
                                 natoms=3),
               synth.make_system(N=(12, 8, 10), L=(7.5, 5.0, 6.3), nk3=(1, 2, 1), NB=2, seed=22, atoms="random",
                                 natoms=2),
               synth.make_system(N=(8, 8, 8), L=5.2, nk3=(1, 1, 2), NB=4, seed=23, atoms="random", natoms=1)]
    nsteps = 4
    ac = _ac(nsteps)
    singles = []
    for s in systems:
        with gpu.Context(s, prec) as c:
            c.set_path(gpu.PATH_STREAMED)
            singles.append(c.rt_run(ac, nsteps, dt=0.08))
    ctxs = [gpu.Context(s, prec) for s in systems]
    try:
        with gpu.Batch(ctxs, prec) as b:
            b.set_kernel(gpu.BATCH_SWEEP)
            r1 = b.rt_run(ac[:3], 2, 0.08)
            assert b.kernel() == gpu.BATCH_SWEEP
            r2 = b.rt_run(ac[2:], 2, 0.08)
    finally:
        for c in ctxs:
            c.close()
    tol = 1e-12 if prec == gpu.FP64 else 1e-5
    for i, r in enumerate(singles):
        J = np.concatenate([r1["J"][i], r2["J"][i]])
        E = np.concatenate([r1["E"][i], r2["E"][i]])
        N = np.concatenate([r1["norm"][i], r2["norm"][i]])
        s = systems[i]
        ej = np.max(np.abs(J - r["J"])) / np.max(np.abs(r["J"]))
        print(f"[gpu] sweep {s.NLx}x{s.NLy}x{s.NLz} prec={prec}: J {ej:.2e} E {rel(E, r['E']):.2e} "
              f"norm {rel(N, r['norm']):.2e}", flush=True)
        assert ej < tol and rel(E, r["E"]) < tol and rel(N, r["norm"]) < tol


def test_batch_rejects_mixed_precisions():
    s = synth.config4_system(0, scale=0.1)
    with gpu.Context(s, gpu.FP64) as c64, gpu.Context(s, gpu.MIXED) as c32:
        with pytest.raises(gpu.SalmonGPUError, match="precision"):
            gpu.Batch([c64, c32], gpu.FP64)
