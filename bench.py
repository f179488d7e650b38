"""Throughput of ARTED's real-time step on one MI355X.

Workload: the 16^3 cell (L = 10.26 bohr) with 32 bands at each of 512 k points
(Monkhorst-Pack 8^3), frozen synthetic potential with 8 atoms x 4 KB
projectors, Acos2 pulse (A0 = 0.05, w = 0.057 Ha, 20 fs), dt = 0.08, Taylor
order 4.  One step = Taylor-4 propagation of every (k, band) orbital at the
midpoint k+A, the density, and the current and energy at k+A(t+dt).

Metric: Taylor steps * k * band per second (whole job), for the mixed path
(complex64 stencil/KB/Taylor, FP64 sums) with the FP64 path measured on the
same orbitals; the J(t) relative error max|J32 - J64| / max|J64| over the run
is reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


class HipEvents:
    """hipEvent timing on a given stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def copy_calibration(self, nbytes):
        """One device-to-device copy of nbytes (the runtime's copy kernel reads
        and writes nbytes): the known-size reference for the FETCH_SIZE and
        WRITE_SIZE counters."""
        L = self.L
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert L.hipMalloc(ctypes.byref(a), nbytes) == 0 and L.hipMalloc(ctypes.byref(b), nbytes) == 0
        assert L.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
        assert L.hipDeviceSynchronize() == 0
        L.hipFree(a)
        L.hipFree(b)

    def elapsed_ms(self, a, b):
        assert self.L.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)


def measured_traffic():
    """HBM bytes per launch of the fused step kernels, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/), scaled by the same
    counters read on a 1 GiB copy; {} when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "fused_step_16c_512k32b_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        d = json.load(f)
    return {k: d[k]["hbm_bytes_per_launch"] for k in ("mixed", "fp64") if k in d}


def workload(seed=0, nsteps=2000):
    from salmon_amd import synth
    s, ac = synth.config2(seed=seed, nsteps=nsteps, orbitals=False)
    rng = np.random.default_rng(seed + 1)
    s.zu = rng.standard_normal((s.NK, s.NB, s.NL)) + 1j * rng.standard_normal((s.NK, s.NB, s.NL))
    return s, ac


def ref_flop_per_orbital_step(s):
    """The reference's own FLOP model of one RT step of one orbital
    (ARTED/modules/performance_analyzer.f90:242-301): four H applications of
    158 flop per grid point for the stencil and 4 + 21 sum_lma Mps for the
    projectors, and 8 flop per point per Taylor term for the update."""
    mps = np.asarray(s.Mps)
    sum_mps = int(np.sum(mps[np.asarray(s.a_tbl) - 1]))
    return 4 * 158 * s.NL + 4 * (4 + 21 * sum_mps) + 4 * 8 * s.NL


# seconds of steps before each timed run, so that the GPU clocks are steady
WARM_S = 0.3


def run_gpu(s, ac, prec, warmup, steps, ev, scf=False, etrs=False):
    from salmon_amd import gpu
    with gpu.Context(s, prec) as c:
        if etrs:  # enforced time-reversal symmetry: two half steps per step
            c.set_propagator("etrs")
        if scf:  # Vloc = V_H + Vpsl + V_xc (PZ81) refreshed after every step
            c.set_scf(s.vloc, gpu.XC_PZ)
        # steady GPU clocks before timing: about 0.3 s of steps, after which the
        # orbitals (and the self-consistent potential) start over, so every
        # precision times the same trajectory
        t_warm = time.perf_counter()
        while time.perf_counter() - t_warm < WARM_S:
            c.rt_run(ac[: max(warmup, 1) + 1], max(warmup, 1))
            c.synchronize()
        c.put_orbitals(s.zu)
        if scf:
            c.set_scf(s.vloc, gpu.XC_PZ)
        c.rt_run(ac[: warmup + 1], warmup)
        c.synchronize()
        st = c.stream()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, st)
        t0 = time.perf_counter()
        r = c.rt_run(ac[warmup: warmup + steps + 1], steps)
        ev.record(e1, st)
        c.synchronize()
        t1 = time.perf_counter()
        ms = ev.elapsed_ms(e0, e1)
    return dict(ms=ms, wall=t1 - t0, J=r["J"], E=r["E"], norm=r["norm"])


def other_grid(ev, steps=10):
    """rt_run off the fused 16^3 shape: a 32^3 cell (L = 15 bohr, 8 atoms),
    64 k x 32 bands, frozen potential, middlepoint, through the persistent
    sweep (what the default path takes there) and through the streamed
    kernels, FP64 and mixed; HIP events around the timed steps, and the same
    12 S NL bytes per orbital-step against 8 TB/s."""
    from salmon_amd import gpu, synth
    s = synth.make_system(N=(32, 32, 32), L=15.0, nk3=(4, 4, 4), NB=32, seed=1, orbitals=False)
    s.zu = synth.random_orbitals(np.random.default_rng(3), s.NK, s.NB, s.NL, s.Hxyz, orthonormalise=False)
    ac = synth.acos2_pulse(steps + 3, 0.08, 0.05, 0.057, 826.8)
    out = {"workload": f"32^3 cell, 64 k x 32 bands (NKB={s.NK * s.NB}), frozen potential, middlepoint, "
                       f"{steps} steps after 2"}
    for prec, tag in ((gpu.FP64, "fp64"), (gpu.MIXED, "mixed")):
        for path, name in ((gpu.PATH_AUTO, "sweep"), (gpu.PATH_STREAMED, "streamed")):
            with gpu.Context(s, prec) as c:
                c.set_path(path)
                c.rt_run(ac[:3], 2)
                c.synchronize()
                st = c.stream()
                e0, e1 = ev.event(), ev.event()
                ev.record(e0, st)
                c.rt_run(ac[2:steps + 3], steps)
                ev.record(e1, st)
                c.synchronize()
                ms = ev.elapsed_ms(e0, e1) / steps
            B = 12 * (16 if prec == gpu.FP64 else 8) * s.NL * s.NK * s.NB
            out[f"{tag}_{name}_ms_per_step"] = ms
            out[f"{tag}_{name}_frac"] = B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    return out


CPU_SAMPLE_SCRIPT = r"""
import sys, time, json, os
sys.path.insert(0, sys.argv[1])
import numpy as np
from salmon_amd import synth
nk, nsteps, kind, which = int(sys.argv[2]), int(sys.argv[3]), sys.argv[4], sys.argv[5]
if which == "cfg1":
    s, ac = synth.config1()
    nsteps = min(nsteps, ac.shape[0] - 2)
else:
    s, ac = synth.config2(nsteps=nsteps + 1, orbitals=False)
    s.kAc0 = np.ascontiguousarray(s.kAc0[:, :nk]); s.occ = np.ascontiguousarray(s.occ[:nk])
    s.zu = synth.random_orbitals(np.random.default_rng(7), nk, s.NB, s.NL, s.Hxyz)
if kind == "reference":
    from oracle.ref import Ref
    R = Ref(s)
    threads = R.num_threads()
    R.rt_run(ac[:2], 1)
    t0 = time.perf_counter(); R.rt_run(ac[1:nsteps + 2], nsteps); t = time.perf_counter() - t0
else:
    from oracle.cpu import Oracle, lib
    O = Oracle(s)
    threads = lib().or_num_threads()
    O.rt_run(ac[:2], 1)
    t0 = time.perf_counter(); O.rt_run(ac[1:nsteps + 2], nsteps); t = time.perf_counter() - t0
print(json.dumps(dict(seconds=t, threads=threads, nkb=s.NK * s.NB, steps=nsteps)))
"""


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may run on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return dict(model=model, nproc=os.cpu_count() or 1, allowed=allowed)


def _cpu_run(kind, which, nk, nsteps, env):
    out = subprocess.run([sys.executable, "-c", CPU_SAMPLE_SCRIPT, ROOT, str(nk), str(nsteps), kind, which],
                         capture_output=True, text=True, env=env, timeout=600)
    if out.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + out.stderr[-2000:])
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(nk=64, nsteps=4):
    """The reference's Fortran (oracle/_ref, built with flang -O3 -fopenmp and
    the reference's default stencil flags) on a bounded sample of the same
    workload, and on the whole config-1 run (16^3, Gamma, 16 orbitals, 100
    steps), on the host cores, in subprocesses (its module globals hold one
    system per process).  Threads: OMP_NUM_THREADS where the environment sets
    it (the GPU box gives each GPU's jobs a 16-CPU share and sets it to 16),
    else every CPU this process may run on."""
    from oracle import ref
    kind = "reference" if ref.available() else "port"
    host = host_cpu()
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(host["allowed"]))
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    d = _cpu_run(kind, "cfg2", nk, nsteps, env)
    rate = d["nkb"] * d["steps"] / d["seconds"]
    res = dict(value=rate, unit="k*band*steps/s", cores=int(d["threads"]), kind=kind,
               sample=f"{nk} of 512 k points x 32 bands on the 16^3 cell, {nsteps} steps after 1 warm-up "
                      f"(Taylor-4 + density + current + energy per step), {d['seconds']:.2f} s",
               host=dict(host, omp_num_threads=env["OMP_NUM_THREADS"]))
    try:
        d1 = _cpu_run(kind, "cfg1", 1, 99, env)
        res["config1"] = dict(value=d1["nkb"] * d1["steps"] / d1["seconds"], unit="k*band*steps/s",
                              steps=d1["steps"], seconds=d1["seconds"], threads=int(d1["threads"]))
    except Exception as e:  # reported, not fatal
        res["config1"] = {"error": str(e)[:300]}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-nk", type=int, default=128)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy (reference size for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the k-point sharding over several GPUs is not built; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from salmon_amd import gpu
    ev = HipEvents()
    if args.copy_calibration:
        ev.copy_calibration(1 << 30)
    s, ac = workload(nsteps=args.warmup + args.steps + 1)
    # orthonormalise once in FP64 on the GPU; both precisions start from it
    with gpu.Context(s, gpu.FP64) as c:
        c.gram_schmidt()
        s.zu = c.get_orbitals()
    nkb, steps = s.NK * s.NB, args.steps
    r64 = run_gpu(s, ac, gpu.FP64, args.warmup, steps, ev)
    r32 = run_gpu(s, ac, gpu.MIXED, args.warmup, steps, ev)
    jerr = float(np.max(np.abs(r32["J"] - r64["J"])) / np.max(np.abs(r64["J"])))
    eerr = float(np.max(np.abs(r32["E"] - r64["E"])) / np.max(np.abs(r64["E"])))
    rate = lambda r: nkb * steps / (r["ms"] * 1e-3)
    v32, v64 = rate(r32), rate(r64)
    ms32, ms64 = r32["ms"] / steps, r64["ms"] / steps
    # algorithmic bytes per orbital-step: 12*S*NL (11 S NL for four streamed
    # H applications + 1 S NL for one observables pass); one step covers all
    # NKB orbitals
    B32 = 12 * 8 * s.NL * nkb
    B64 = 12 * 16 * s.NL * nkb
    ach = B32 / (ms32 * 1e-3) / 1e9
    traffic = measured_traffic()
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic.get("mixed"),
            "kernel": "k_fused_step<float> (one launch per RT step, which also evaluates the previous step's "
                      "observables, plus one observables-only launch closing the run; duration = timed region / "
                      "steps, which also holds the small row and density reduction kernels)",
            "algorithmic_bytes_per_launch": B32,
            "fp64": {"kernel": "k_fused_step<double>", "achieved": B64 / (ms64 * 1e-3) / 1e9,
                     "frac": B64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": B64,
                     "traffic": traffic.get("fp64")}}
    # the same workload with the self-consistent potential (Hartree + LDA
    # after every step), a shorter run
    nscf = min(steps, 25)
    s64 = run_gpu(s, ac, gpu.FP64, args.warmup, nscf, ev, scf=True)
    s32 = run_gpu(s, ac, gpu.MIXED, args.warmup, nscf, ev, scf=True)
    scf = {"steps": nscf, "ms_per_step": s32["ms"] / nscf, "value": nkb * nscf / (s32["ms"] * 1e-3),
           "fp64": {"ms_per_step": s64["ms"] / nscf, "value": nkb * nscf / (s64["ms"] * 1e-3)},
           "J_rel_err": float(np.max(np.abs(s32["J"] - s64["J"])) / np.max(np.abs(s64["J"])))}
    # the same workload with the ETRS propagator (frozen potential)
    e64 = run_gpu(s, ac, gpu.FP64, args.warmup, nscf, ev, etrs=True)
    e32 = run_gpu(s, ac, gpu.MIXED, args.warmup, nscf, ev, etrs=True)
    etrs = {"steps": nscf, "ms_per_step": e32["ms"] / nscf, "value": nkb * nscf / (e32["ms"] * 1e-3),
            "fp64": {"ms_per_step": e64["ms"] / nscf, "value": nkb * nscf / (e64["ms"] * 1e-3)},
            "J_rel_err": float(np.max(np.abs(e32["J"] - e64["J"])) / np.max(np.abs(e64["J"])))}
    try:
        og = other_grid(ev)
    except Exception as e:  # reported, not fatal
        og = {"error": str(e)[:300]}
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_nk, args.cpu_steps)
        except Exception as e:  # reported, not fatal
            cpu = {"error": str(e)[:300]}
    fps = ref_flop_per_orbital_step(s) * nkb
    flop_model = {"gflop_per_step": fps / 1e9, "mixed_tflops": fps / (ms32 * 1e-3) / 1e12,
                  "fp64_tflops": fps / (ms64 * 1e-3) / 1e12,
                  "source": "ARTED/modules/performance_analyzer.f90:242-301 (stencil 158 flop/point, "
                            "KB 4 + 21 sum Mps, update 8 flop/point/term; four H applications per step)"}
    if isinstance(cpu, dict) and "value" in cpu:
        flop_model["cpu_baseline_gflops"] = cpu["value"] * ref_flop_per_orbital_step(s) / 1e9
    line = {
        "metric": "Taylor steps*k*band/s (16^3, 512k x 32) FP64 vs mixed @ J(t) rel err<=1e-5",
        "value": v32, "unit": "k*band*steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (FP64 sums)", "data": "synthetic",
        "config": {"workload": "16^3 cell, 512 k x 32 bands (NKB=16384), 8 atoms x 4 KB projectors, Acos2 pulse, "
                               "dt=0.08, Taylor-4, J+E+rho every step",
                   "NL": s.NL, "NK": s.NK, "NB": s.NB, "precision": "mixed"},
        "fp64": {"value": v64, "ms_per_step": ms64},
        "mixed_vs_fp64_speedup": v32 / v64,
        "J_rel_err": jerr, "E_rel_err": eerr,
        "roofline": roof,
        "scf": scf,
        "etrs": etrs,
        "other_grid": og,
        "flop_model": flop_model,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
