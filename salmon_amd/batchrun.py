"""The 64-system batched workload: independent seeded systems with grids from
{12^3, 16^3, 20x20x28, 24^3, 32^3}, bands U{8..64}, k points U{1..64}, atoms
U{1..16} (synth.config4_system), propagated together with the batched step
(the persistent sweep kernel, one launch per step, or with --kernel phases one
launch per phase over all systems with z blocking tuned per grid shape; the
step replayed from a captured HIP graph), FP64 and complex64.  The same
systems propagated one context after another give the unbatched rate beside it.

  python -m salmon_amd.batchrun --steps 10 [--scale 1.0] [--kernel auto|sweep|phases] [--out result.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import gpu, synth


def _log(msg):
    print(msg, flush=True)


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


KERNELS = {"auto": gpu.BATCH_AUTO, "sweep": gpu.BATCH_SWEEP, "phases": gpu.BATCH_PHASES}


def run(nsys=64, steps=10, scale=1.0, precisions=(gpu.FP64, gpu.MIXED), sequential=True, log=_log, kernel="auto"):
    t0 = time.time()
    systems = [synth.config4_system(i, scale=scale) for i in range(nsys)]
    orbs = sum(s.NK * s.NB for s in systems)
    pts = sum(s.NK * s.NB * s.NL for s in systems)
    log(f"[batch] {nsys} systems, {orbs} orbitals, {pts * 16 / 1e9:.1f} GB in FP64 (built in {time.time() - t0:.0f} s)")
    ac = synth.acos2_pulse(steps + 1, 0.08, 0.05, 0.057, 826.8)
    out = dict(systems=nsys, orbitals=int(orbs), grid_points_x_orbitals=int(pts), steps=steps,
               shapes=sorted({f"{s.NLx}x{s.NLy}x{s.NLz}" for s in systems}))
    for prec in precisions:
        name = "fp64" if prec == gpu.FP64 else "mixed"
        ctxs = [gpu.Context(s, prec) for s in systems]
        with gpu.Batch(ctxs, prec) as b:
            b.set_kernel(KERNELS[kernel])
            tt = time.time()
            zb = b.tune()
            tune_s = time.time() - tt
            kern = "sweep" if b.kernel() == gpu.BATCH_SWEEP else "phases"
            b.rt_run(ac[:2], 1, 0.08)  # warm-up step (captures the graph)
            tt = time.time()
            r = b.rt_run(ac[1:], steps, 0.08)
            dt = time.time() - tt
        res = dict(kernel=kern, ms_per_step=dt / steps * 1e3, orbital_steps_per_s=orbs * steps / dt,
                   point_orbital_steps_per_s=pts * steps / dt, tune_s=tune_s,
                   zblock={f"{s.NLx}x{s.NLy}x{s.NLz}": int(z) for s, z in zip(systems, zb)},
                   finite=bool(np.all(np.isfinite(r["J"])) and np.all(np.isfinite(r["E"]))))
        # HBM roofline with the algorithmic bytes of the streamed step, 12 S NL
        # per orbital-step (four H applications and the observables pass)
        S = 16 if prec == gpu.FP64 else 8
        res["algorithmic_bytes_per_step"] = 12 * S * pts
        res["roofline_gbs"] = res["algorithmic_bytes_per_step"] / (res["ms_per_step"] * 1e-3) / 1e9
        res["roofline_frac"] = res["roofline_gbs"] / HBM_PEAK_GBS
        zdesc = f"z blocking {res['zblock']}" if kern == "phases" else "persistent sweep"
        log(f"[batch] {name}: {res['ms_per_step']:.2f} ms/step, {res['orbital_steps_per_s']:.3e} orbital-steps/s "
            f"(tune {tune_s:.1f} s, {zdesc}); {res['roofline_gbs']:.0f} GB/s of algorithmic "
            f"bytes = {res['roofline_frac']:.3f} of {HBM_PEAK_GBS:.0f} GB/s")
        if sequential:
            # the same kernels one context at a time (each context frees its
            # work buffers before the next one runs)
            ds = 0.0
            for c in ctxs:
                c.set_path(gpu.PATH_STREAMED)
                c.rt_run(ac[:2], 1, dt=0.08)
                tt = time.time()
                c.rt_run(ac[1:], steps, dt=0.08)
                ds += time.time() - tt
                c.close()
            res["sequential_ms_per_step"] = ds / steps * 1e3
            log(f"[batch] {name}: one context after another {res['sequential_ms_per_step']:.2f} ms/step")
        for c in ctxs:
            c.close()
        out[name] = res
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--kernel", choices=sorted(KERNELS), default="auto")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    res = run(a.systems, a.steps, a.scale, sequential=not a.no_sequential, kernel=a.kernel)
    _log(json.dumps(res))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
