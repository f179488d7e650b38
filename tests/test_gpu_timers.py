"""Region timers (salmon_gpu_timer_*, the reference's timer_begin/timer_end
regions of misc/timer.f90): call counts per region on the streamed and fused
RT paths and the ground state, nesting of the times, LOG_ALL as the sum of
the outermost regions, reset, and results unchanged by timing."""
import time

import numpy as np
import pytest

from helpers import load_case
from salmon_amd import gpu

pytestmark = pytest.mark.gpu


def _run(s, path, prec, ac, nsteps, timed):
    with gpu.Context(s, prec) as c:
        c.set_path(path)
        if timed:
            c.timer_enable(True)
        r = c.rt_run(ac, nsteps)
        return r, (c.timers() if timed else {})


def test_streamed_regions_count_and_nest():
    s, d = load_case("small_k")
    nsteps = 3
    print(f"[gpu] timers streamed small_k {nsteps} steps", flush=True)
    r0, _ = _run(s, gpu.PATH_STREAMED, gpu.FP64, d["ac"], nsteps, False)
    r1, t = _run(s, gpu.PATH_STREAMED, gpu.FP64, d["ac"], nsteps, True)
    print(f"[gpu] timers: {t}", flush=True)
    # timing changes nothing in the results
    assert np.array_equal(r0["J"], r1["J"]) and np.array_equal(r0["E"], r1["E"])
    assert t["DYNAMICS"][1] == 1
    assert t["DT_EVOLVE"][1] == nsteps
    assert t["PSI_RHO"][1] == nsteps
    assert t["CURRENT"][1] == nsteps
    assert t["HPSI"][1] >= 4 * nsteps and t["HPSI"][1] % (4 * nsteps) == 0
    # nesting: inner regions within their outer one
    assert 0 < t["HPSI"][0] <= t["DT_EVOLVE"][0] * 1.001
    assert t["DT_EVOLVE"][0] + t["PSI_RHO"][0] + t["CURRENT"][0] <= t["DYNAMICS"][0] * 1.001
    assert t["ALL"] == t["DYNAMICS"]


def test_fused_regions_and_reset():
    s, d = load_case("cfg1")
    nsteps = 10
    print(f"[gpu] timers fused cfg1 {nsteps} steps", flush=True)
    with gpu.Context(s, gpu.MIXED) as c:
        assert c.path() == gpu.PATH_FUSED
        c.rt_run(d["ac"], 2)  # warm
        c.timer_enable(True)
        c.synchronize()
        w0 = time.perf_counter()
        c.rt_run(d["ac"], nsteps)
        wall = time.perf_counter() - w0
        t = c.timers()
        print(f"[gpu] timers: {t}, wall {wall * 1e3:.2f} ms", flush=True)
        assert t["DYNAMICS"][1] == 1
        # one launch per step, and the observables-only launch closing the run
        assert t["DT_EVOLVE"][1] == nsteps and t["CURRENT"][1] == 1
        assert "HPSI" not in t and "PSI_RHO" not in t
        assert t["DT_EVOLVE"][0] + t["CURRENT"][0] <= t["DYNAMICS"][0] * 1.001 <= wall * 1.01
        c.timer_reset()
        assert c.timers() == {}
        c.timer_enable(False)
        c.rt_run(d["ac"], 2)
        assert c.timers() == {}


def test_ground_state_regions():
    s, _ = load_case("small_k")
    print("[gpu] timers ground state small_k", flush=True)
    with gpu.Context(s, gpu.FP64) as c:
        c.timer_enable(True)
        c.gram_schmidt()
        c.cg(s.kAc0, 4)
        c.sp_energy(s.kAc0)
        t = c.timers()
    print(f"[gpu] timers: {t}", flush=True)
    assert t["GRAM_SCHMIDT"][1] >= 1
    assert t["CG"][1] == 1
    assert t["SP_ENERGY"][1] == 1
    assert t["CG"][0] > 0
