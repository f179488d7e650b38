"""The trained learned preconditioner on the config-5 cell (32^3, L = 15 bohr,
Gamma, 64 bands, 27 random-position atoms), with the shipped weights
salmon_amd/data/pcnet_cfg5.npy, on held-out seeds (training used seeds
1000..1199).

Per seed: the kinetic-preconditioned sweep loop (ground_state.f90:136-140,
4 CG iterations per band per sweep) continued to a 1e-12 Ha change per sweep
gives the reference band energies; the learned solve starts from the same
orbitals.  Counted: the first sweep after which every band energy is within
1e-8 Ha of the reference.  Asserted: both solves reach 1e-8 Ha, the learned
solve's final band energies agree with the reference to 1e-8 Ha, and over the
seeds the learned solve needs no more sweeps than the kinetic one (the gain
is printed).  The 50-seed evaluation is salmon_amd/data/pcnet_cfg5_eval.json.
"""
import os

import numpy as np
import pytest

from salmon_amd import gpu, pctrain, synth

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "salmon_amd", "data",
                       "pcnet_cfg5.npy")
SEEDS = [0, 1, 3]
ACC = 1e-8
MAX_SWEEPS = 120


def test_shipped_network_on_held_out_config5_seeds():
    w = np.load(WEIGHTS, allow_pickle=False)
    assert w.size == gpu.PCNET_WEIGHTS
    net = gpu.PcNet(weights=w)
    tot_k = tot_l = 0
    try:
        for seed in SEEDS:
            s = synth.config5(seed=seed)
            print(f"[gpu] config 5 seed {seed}: 32^3, {s.NB} bands, {s.NI} atoms", flush=True)
            with gpu.Context(s, gpu.FP64) as c:
                rk = c.ground_state(s.kAc0, MAX_SWEEPS, iter_cg_max=pctrain.NCG, precond=gpu.PC_KINETIC, tol=1e-12)
                ref = rk["esp"]
                c.put_orbitals(s.zu)
                c.set_pcnet(net)
                rl = c.ground_state(s.kAc0, MAX_SWEEPS, iter_cg_max=pctrain.NCG, precond=gpu.PC_LEARNED, tol=1e-12)
            nk = pctrain.sweeps_to(rk["history"], ref, ACC)
            nl = pctrain.sweeps_to(rl["history"], ref, ACC)
            err = float(np.max(np.abs(rl["esp"] - ref)))
            print(f"[gpu] config 5 seed {seed}: sweeps to {ACC:g} Ha kinetic {nk} learned {nl}; "
                  f"learned final error {err:.1e} Ha", flush=True)
            assert nk is not None and nl is not None
            assert err < ACC
            tot_k += nk
            tot_l += nl
    finally:
        net.close()
    print(f"[gpu] config 5: learned preconditioner needs {tot_k / tot_l:.3f}x fewer sweeps than kinetic "
          f"over seeds {SEEDS} ({tot_l} vs {tot_k})", flush=True)
    assert tot_l <= tot_k
