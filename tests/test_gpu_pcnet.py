"""Learned CG preconditioner on the GPU: the training kernels' loss and
gradient against the NumPy oracle (oracle/pcnet.py, itself pinned by finite
differences), training lowering the loss, and the learned ground-state solve
reaching the same eigenvalues as the kinetic one.

Tolerance: the network runs in FP32 (weights, activations, FP32 sums inside a
1024-point chunk, FP64 across chunks); the oracle runs in FP64 with the same
weights, so the loss agrees to ~1e-6 and the gradient to ~1e-4 relative.
"""
import numpy as np
import pytest

from oracle import pcnet
from oracle.cpu import Oracle
from salmon_amd import gpu, synth

pytestmark = pytest.mark.gpu


def _system():
    return synth.make_system(N=(12, 12, 12), L=5.6, nk3=(1, 1, 1), NB=4, seed=11, atoms="random", natoms=2,
                             name="pc12")


@pytest.fixture(scope="module")
def recorded():
    s = _system()
    print(f"[gpu] pcnet fixture {s.name} NL={s.NL} NB={s.NB}", flush=True)
    c = gpu.Context(s, gpu.FP64)
    c.ground_state(s.kAc0, 200, iter_cg_max=4, precond=gpu.PC_KINETIC, tol=1e-12)
    c.pc_set_targets(s.kAc0)
    Z = c.get_orbitals()[0]
    lam = c.sp_energy(s.kAc0)[0]
    c.put_orbitals(s.zu)
    c.gram_schmidt()
    c.pc_record(48, 1)
    c.cg(s.kAc0, 4, gpu.PC_KINETIC)
    c.cg(s.kAc0, 4, gpu.PC_KINETIC)
    stored, valid = c.pc_samples()
    print(f"[gpu] recorded {stored} samples, {valid} valid", flush=True)
    yield s, c, Z, lam
    c.close()


def test_recording_and_targets(recorded):
    s, c, Z, lam = recorded
    stored, valid = c.pc_samples()
    assert stored == 32 and valid >= 16
    smp = c.pc_get_sample(0)
    assert smp["band"] == 0
    # the recorded theta is <x|H|x> of the recorded iterate
    o = Oracle(s)
    x = smp["x"]
    th = np.vdot(x, o.hpsi(s.kAc0[:, 0], x)).real * s.Hxyz
    assert abs(th - smp["theta"]) < 1e-10 * max(1.0, abs(th))
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(Z.conj() @ Z.T * s.Hxyz - np.eye(s.NB))) < 1e-10


def test_gradient_matches_oracle(recorded):
    s, c, Z, lam = recorded
    o = Oracle(s)
    net = gpu.PcNet(seed=3)
    w = net.weights().astype(np.float64)
    idx = [i for i in range(c.pc_samples()[0]) if c.pc_get_sample(i)["valid"]][::5][:4]
    loss, grad = net.gradient(c, idx)
    shape = (s.NLx, s.NLy, s.NLz)
    hpsi = lambda v: o.hpsi(s.kAc0[:, 0], v)
    rl, rg = 0.0, np.zeros(pcnet.NWEIGHTS)
    for i in idx:
        smp = c.pc_get_sample(i)
        b = smp["band"]
        f, gw = pcnet.sample_loss_grad(w, hpsi, s.Hxyz, shape, smp["g"], smp["x"], Z[:b], smp["theta"], lam[b])
        rl += f / len(idx)
        rg += gw / len(idx)
    print(f"[gpu] pcnet loss {loss:.8f} oracle {rl:.8f}; grad rel {np.linalg.norm(grad - rg) / np.linalg.norm(rg):.2e}",
          flush=True)
    assert abs(loss - rl) < 1e-5 * max(1.0, abs(rl))
    assert np.linalg.norm(grad - rg) < 2e-4 * np.linalg.norm(rg)
    for lo, hi in ((0, pcnet.NW1), (pcnet.NW1, pcnet.NW1 + pcnet.NW2), (pcnet.NW1 + pcnet.NW2, pcnet.NWEIGHTS)):
        assert np.linalg.norm(grad[lo:hi] - rg[lo:hi]) < 1e-3 * np.linalg.norm(rg[lo:hi])
    net.close()


def test_training_lowers_the_loss(recorded):
    s, c, Z, lam = recorded
    net = gpu.PcNet(seed=4)
    idx = [i for i in range(c.pc_samples()[0]) if c.pc_get_sample(i)["valid"]]
    before, _ = net.gradient(c, idx)
    net.train(c, 150, batch=4, lr=2e-3, seed=1)
    after, _ = net.gradient(c, idx)
    print(f"[gpu] pcnet training: mean loss {before:.4f} -> {after:.4f}", flush=True)
    assert after < before - 0.05
    net.close()


def test_learned_ground_state_reaches_kinetic_eigenvalues(recorded):
    """SALMON_PC_LEARNED (with its fallback to the kinetic preconditioner)
    reaches the kinetic solve's eigenvalues; a learned CG call is
    deterministic."""
    s, c, Z, lam = recorded
    net = gpu.PcNet(seed=4)
    net.train(c, 150, batch=4, lr=2e-3, seed=1)
    with gpu.Context(s, gpu.FP64) as c2:
        c2.set_pcnet(net)
        r = c2.ground_state(s.kAc0, 300, iter_cg_max=4, precond=gpu.PC_LEARNED, tol=1e-12)
        print(f"[gpu] learned ground state: {r['sweeps']} sweeps, err {np.max(np.abs(r['esp'][0] - lam)):.2e}",
              flush=True)
        assert np.max(np.abs(r["esp"][0] - lam)) < 1e-8
    runs = []
    for _ in range(2):
        with gpu.Context(s, gpu.FP64) as c3:
            c3.set_pcnet(net)
            ev, _ = c3.cg(s.kAc0, 3, gpu.PC_LEARNED)
            runs.append((ev, c3.get_orbitals()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    net.close()


def test_learned_requires_network():
    s = _system()
    with gpu.Context(s, gpu.FP64) as c:
        with pytest.raises(gpu.SalmonGPUError, match="network"):
            c.cg(s.kAc0, 2, gpu.PC_LEARNED)


def test_mfma_inference_matches_oracle(recorded):
    """The BF16 matrix-core inference (as used inside CG) against the FP64
    oracle network with the same weights: bf16 weights and activations over
    three layers leave ~1e-2 relative error."""
    s, c, Z, lam = recorded
    net = gpu.PcNet(seed=5)
    w = net.weights().astype(np.float64)
    shape = (s.NLx, s.NLy, s.NLz)
    for i in (0, 7, 19):
        g = c.pc_get_sample(i)["g"]
        d = net.apply(c, g)
        ref = pcnet.direction(w, g, shape) * np.linalg.norm(g) / np.sqrt(s.NL)
        err = np.linalg.norm(d - ref) / np.linalg.norm(ref)
        print(f"[gpu] mfma inference sample {i}: rel err {err:.2e}", flush=True)
        assert err < 3e-2
    net.close()
