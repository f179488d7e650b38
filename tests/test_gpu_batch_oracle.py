"""Config-4 batch against the C oracle.

One system of each grid shape of the 64-system workload (12^3, 20x20x28,
32^3, 16^3, 24^3; bands and k points scaled by 0.12) runs through the batched
step -- the persistent sweep kernel and the one-launch-per-phase kernel, the
step replayed from a captured graph -- and each system's J(t), E(t), norm and
final orbitals are compared with the oracle's RT loop (oracle/salmon_oracle.c
or_rt_run: dt_evolve.f90:106-216 then control_sc.f90:203-222).  The
single-context streamed path is held to the oracle on the same shapes.

Then the full 64-system workload (10 GB of FP64 orbitals) runs two batched
steps: J and E finite for every system, no norm grows (a Taylor-4 step damps
every component, |R(i x)|^2 = 1 - x^6/72 + x^8/576 <= 1 inside its region),
and one k point of one system of each shape equals the oracle propagating that
k point alone (with a frozen potential every k point evolves on its own).

Tolerances: FP64 1e-12 relative (max |d| / max |ref|) on orbitals, J, E and
norm (the batched stencil sums in the oracle's order up to FMA contraction);
complex64 against the FP64 oracle 1e-5 on J, E and norm (the J(t) bar of the
precision form) and 1e-4 on orbitals.
"""
import dataclasses
import time

import numpy as np
import pytest

from helpers import rel
from oracle.cpu import Oracle
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu

# one system of each grid shape: 12^3, 20x20x28, 32^3, 16^3, 24^3
SHAPE_SYSTEMS = [0, 1, 2, 3, 9]
DT = 0.08


def _ac(n):
    return synth.acos2_pulse(n, DT, 0.05, 0.057, 826.8)


def _shape(s):
    return f"{s.NLx}x{s.NLy}x{s.NLz}"


def _jerr(j, jr):
    return float(np.max(np.abs(j - jr)) / np.max(np.abs(jr)))


@pytest.mark.parametrize("kernel", [gpu.BATCH_SWEEP, gpu.BATCH_PHASES])
@pytest.mark.parametrize("prec", [gpu.FP64, gpu.MIXED])
def test_batch_every_shape_against_oracle(prec, kernel):
    nsteps = 3
    ac = _ac(nsteps)
    systems = [synth.config4_system(i, scale=0.12) for i in SHAPE_SYSTEMS]
    print(f"[gpu] batch vs oracle prec={prec} kernel={kernel}: "
          + ", ".join(f"{_shape(s)}/{s.NK}k/{s.NB}b/{s.NI}at" for s in systems), flush=True)
    ctxs = [gpu.Context(s, prec) for s in systems]
    try:
        with gpu.Batch(ctxs, prec) as b:
            b.set_kernel(kernel)
            r = b.rt_run(ac, nsteps, DT)
            assert b.kernel() == kernel
        zs = [c.get_orbitals() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    assert {_shape(s) for s in systems} == {"12x12x12", "16x16x16", "20x20x28", "24x24x24", "32x32x32"}
    tol, tolz = (1e-12, 1e-12) if prec == gpu.FP64 else (1e-5, 1e-4)
    for i, s in enumerate(systems):
        o = Oracle(s).rt_run(ac, nsteps)
        ej, ee, en, ez = _jerr(r["J"][i], o["J"]), rel(r["E"][i], o["E"]), rel(r["norm"][i], o["norm"]), rel(zs[i], o["zu"])
        print(f"[gpu] batch {_shape(s)}: J {ej:.2e} E {ee:.2e} norm {en:.2e} zu {ez:.2e}", flush=True)
        assert ej < tol and ee < tol and en < tol and ez < tolz


@pytest.mark.parametrize("idx", [1, 2, 9])  # 20x20x28, 32^3, 24^3
def test_single_context_other_shapes_against_oracle(idx):
    """A single context on the batch's shapes (every grid but 16^3), through
    the persistent sweep (what SALMON_PATH_AUTO takes there) and through the
    streamed kernels: FP64 against the oracle, complex64 against the FP64
    oracle; J, E, norm, the last step's density and the orbitals."""
    nsteps = 3
    ac = _ac(nsteps)
    s = synth.config4_system(idx, scale=0.12)
    o = Oracle(s).rt_run(ac, nsteps, want_rho=True)
    for prec, tol, tolz in ((gpu.FP64, 1e-12, 1e-12), (gpu.MIXED, 1e-5, 1e-4)):
        for path in (gpu.PATH_AUTO, gpu.PATH_STREAMED):
            with gpu.Context(s, prec) as c:
                c.set_path(path)
                assert c.path() == (gpu.PATH_SWEEP if path == gpu.PATH_AUTO else gpu.PATH_STREAMED)
                r = c.rt_run(ac, nsteps, dt=DT)
                z = c.get_orbitals()
            ej, ee, en, ez = _jerr(r["J"], o["J"]), rel(r["E"], o["E"]), rel(r["norm"], o["norm"]), rel(z, o["zu"])
            er = rel(r["rho"], o["rho"])
            print(f"[gpu] single context {_shape(s)} prec={prec} path={path}: J {ej:.2e} "
                  f"E {ee:.2e} norm {en:.2e} rho {er:.2e} zu {ez:.2e}", flush=True)
            assert ej < tol and ee < tol and en < tol and ez < tolz and er < tolz


# one system of each shape, each with a modest orbital count, whose last k point is compared
FULL_SAMPLES = {27: "12x12x12", 12: "20x20x28", 48: "32x32x32", 28: "16x16x16", 10: "24x24x24"}


def _cheap_orbitals(s, seed):
    """Orbitals for the full-size run without 10 GB of random draws: eight
    random vectors per system, each orbital one of them times a random phase
    and a band-dependent scale."""
    rng = np.random.default_rng(seed)
    base = rng.standard_normal((8, s.NL)) + 1j * rng.standard_normal((8, s.NL))
    phase = np.exp(2j * np.pi * rng.uniform(size=(s.NK, s.NB))) * (1.0 + 0.1 * np.arange(s.NB) / s.NB)[None, :]
    pick = np.add.outer(np.arange(s.NK), 3 * np.arange(s.NB)) % 8
    return base[pick] * phase[:, :, None]


def test_full_config4_workload_two_steps():
    nsteps = 2
    ac = _ac(nsteps)
    t0 = time.time()
    ctxs, systems, start, norm0 = [], [], {}, []
    try:
        for i in range(64):
            s = synth.config4_system(i, orbitals=False)
            s.zu = _cheap_orbitals(s, 30_000 + i)
            if i in FULL_SAMPLES:
                assert _shape(s) == FULL_SAMPLES[i]
                start[i] = s.zu[-1:].copy()
            c = gpu.Context(s, gpu.FP64)
            norm0.append(c.norm())
            s.zu = None  # the device copy is the state from here on
            ctxs.append(c)
            systems.append(s)
        gb = sum(s.NK * s.NB * s.NL for s in systems) * 16 / 1e9
        print(f"[gpu] full config 4: 64 systems, {gb:.1f} GB of FP64 orbitals set up in {time.time() - t0:.0f} s",
              flush=True)
        assert gb > 9.9
        with gpu.Batch(ctxs, gpu.FP64) as b:
            t1 = time.time()
            r = b.rt_run(ac, nsteps, DT)
            print(f"[gpu] full config 4: tune + {nsteps} steps in {time.time() - t1:.1f} s", flush=True)
        assert np.all(np.isfinite(r["J"])) and np.all(np.isfinite(r["E"])) and np.all(np.isfinite(r["norm"]))
        n = np.asarray(r["norm"])
        assert np.all(n[:, 0] <= np.asarray(norm0) * (1 + 1e-12)) and np.all(n[:, 1] <= n[:, 0] * (1 + 1e-12))
        for i, z0 in start.items():
            s = systems[i]
            zk = ctxs[i].get_orbitals()[-1]
            sub = dataclasses.replace(s, kAc0=np.ascontiguousarray(s.kAc0[:, -1:]),
                                      occ=np.ascontiguousarray(s.occ[-1:]), zu=z0)
            o = Oracle(sub).rt_run(ac, nsteps)
            ez = rel(zk, o["zu"][0])
            print(f"[gpu] full config 4 system {i} ({_shape(s)}, {s.NK}k x {s.NB}b): last k point vs oracle {ez:.2e}",
                  flush=True)
            assert ez < 1e-12
    finally:
        for c in ctxs:
            c.close()
