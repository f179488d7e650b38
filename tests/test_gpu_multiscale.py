"""Macro-points on the GPU: the batch with one A(t) per system
(salmon_gpu_batch_rt_run_ms) against single contexts, and the coupled 1D
Maxwell-TDDFT loop (salmon_amd/multiscale.py) with the batch as the micro
step against the same loop with the oracle stepping every macro-point.
Tolerances: FP64 batch vs single contexts 1e-12 (the batch's own bar),
coupled currents 1e-10 and fields 1e-12 relative; complex64 1e-5."""
import numpy as np
import pytest

from helpers import load_case, rel
from oracle.cpu import Oracle
from salmon_amd import gpu, multiscale as ms, synth

pytestmark = pytest.mark.gpu


def _pulses(s, nsteps):
    return np.stack([synth.acos2_pulse(nsteps + 1, s.dt, amp, 0.3, 2.0, pol=pol)
                     for amp, pol in ((0.02, (1.0, 0.0, 0.0)), (0.05, (0.0, 0.6, 0.8)), (0.1, (0.0, 0.0, 1.0)))])


@pytest.mark.parametrize("prec,tol", [(gpu.FP64, 1e-12), (gpu.MIXED, 1e-5)])
def test_batch_with_one_pulse_per_system_matches_single_contexts(prec, tol):
    s, _ = load_case("small_k")
    nsteps = 5
    ac = _pulses(s, nsteps)
    print(f"[gpu] batch_ms small_k x3 prec={prec}", flush=True)
    single = []
    for i in range(3):
        with gpu.Context(s, gpu.FP64) as c:
            single.append(c.rt_run(ac[i], nsteps))
    ctxs = [gpu.Context(s, prec) for _ in range(3)]
    try:
        with gpu.Batch(ctxs, prec) as b:
            out = b.rt_run_ms(ac, nsteps, s.dt)
            out2 = b.rt_run(ac[1], 2, s.dt)  # the shared-table form after it still works
    finally:
        for c in ctxs:
            c.close()
    for i in range(3):
        print(f"[gpu]   sys {i}: J {rel(out['J'][i], single[i]['J']):.2e} E {rel(out['E'][i], single[i]['E']):.2e}",
              flush=True)
        assert rel(out["J"][i], single[i]["J"]) < tol
        assert rel(out["E"][i], single[i]["E"]) < tol
        assert rel(out["norm"][i], single[i]["norm"]) < tol
    assert rel(out["J"][0], out["J"][2]) > 1e-3  # different pulses, different currents
    assert np.all(np.isfinite(out2["J"]))


def test_coupled_maxwell_loop_matches_oracle():
    s, _ = load_case("small_k")
    nmacro, nsteps = 3, 6
    dt = s.dt
    x = np.arange(-3, 7)

    def fresh():
        m = ms.Maxwell1D(-3, 6, 2.0 * ms.C_LIGHT * dt, dt, macropoints=[1, 2, 3])
        m.set_fields(np.stack([0 * x, 0.03 * np.sin(0.4 * x), 0.05 * np.cos(0.4 * x)], 1),
                     np.stack([0 * x, 0.03 * np.sin(0.4 * x - 0.2), 0.05 * np.cos(0.4 * x - 0.2)], 1))
        return m

    # oracle: every macro-point stepped by the C oracle
    o = Oracle(s)
    zs = [s.zu.copy() for _ in range(nmacro)]

    def oracle_micro(acs):
        J = np.zeros((nmacro, 3))
        for i in range(nmacro):
            zs[i] = o.hamiltonian(s.kAc0 + 0.5 * (acs[i, 0] + acs[i, 1])[:, None], zs[i], dt)
            J[i] = o.current(s.kAc0 + acs[i, 1][:, None], zs[i])
        return J

    mr = fresh()
    ref = ms.run(mr, nsteps, oracle_micro)
    print(f"[gpu] multiscale small_k nmacro={nmacro} nsteps={nsteps}", flush=True)
    ctxs = [gpu.Context(s, gpu.FP64) for _ in range(nmacro)]
    try:
        with gpu.Batch(ctxs, gpu.FP64) as b:
            m = fresh()
            out = ms.run(m, nsteps, ms.batch_micro(b, dt))
    finally:
        for c in ctxs:
            c.close()
    mz = fresh()
    ms.run(mz, nsteps, lambda acs: np.zeros((nmacro, 3)))
    coupling = np.max(np.abs(m.ac_new - mz.ac_new)) / np.max(np.abs(mz.ac_new))
    print(f"[gpu] multiscale: J {rel(out['jm'], ref['jm']):.2e} Ac {rel(m.ac_new, mr.ac_new):.2e} "
          f"coupling {coupling:.2e}", flush=True)
    assert rel(out["jm"], ref["jm"]) < 1e-10
    assert rel(m.ac_new, mr.ac_new) < 1e-12
    assert rel(out["ac"], ref["ac"]) < 1e-12
    assert coupling > 1e-7  # the currents do act back on the field
