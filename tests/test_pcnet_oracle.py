"""Pins of the learned-preconditioner oracle (oracle/pcnet.py): brute-force
finite differences of the training gradient, the convolution against a
direct sum, and the symmetries the direction is built to have.  CPU only."""
import numpy as np
import pytest

from oracle import pcnet
from oracle.cpu import Oracle
from salmon_amd import synth


@pytest.fixture(scope="module")
def small():
    s = synth.make_system(N=(6, 6, 5), L=(3.2, 3.2, 2.8), nk3=(1, 1, 1), NB=3, seed=4, atoms="random", natoms=1,
                          rc=1.2, sigma=0.5, name="pc6")
    return s, Oracle(s)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((4, 5, 3, 2))
    W = rng.standard_normal((27, 2, 3))
    out = pcnet.conv(W, a)
    ix, iy, iz, co = 1, 4, 2, 2
    ref = 0.0
    for t, (dx, dy, dz) in enumerate(pcnet.taps()):
        ref += a[(ix + dx) % 4, (iy + dy) % 5, (iz + dz) % 3] @ W[t][:, co]
    assert abs(out[ix, iy, iz, co] - ref) < 1e-12


def test_direction_is_phase_and_scale_covariant(small):
    s, _ = small
    rng = np.random.default_rng(1)
    w = pcnet.init_weights(2)
    g = rng.standard_normal(s.NL) + 1j * rng.standard_normal(s.NL)
    shape = (s.NLx, s.NLy, s.NLz)
    d = pcnet.direction(w, g, shape)
    c = 2.5 * np.exp(0.3j)
    d2 = pcnet.direction(w, c * g, shape)
    assert np.max(np.abs(d2 - np.exp(0.3j) * d)) < 1e-12 * np.max(np.abs(d))


def test_training_gradient_matches_finite_differences(small):
    """d loss / d w of the Rayleigh-Ritz objective against central
    differences along random weight directions (all three layers)."""
    s, o = small
    rng = np.random.default_rng(3)
    shape = (s.NLx, s.NLy, s.NLz)
    h = s.Hxyz
    k = s.kAc0[:, 0]
    Z = s.zu[0]
    hpsi = lambda v: o.hpsi(k, v)
    x = Z[1]
    hx = hpsi(x)
    theta = np.vdot(x, hx).real * h
    g = hx - theta * x
    g = g - Z[0] * (np.vdot(Z[0], g) * h)
    lam = theta - 200.0
    w = pcnet.init_weights(5)
    f0, grad = pcnet.sample_loss_grad(w, hpsi, h, shape, g, x, Z[:1], theta, lam)
    assert -1.0 < f0 < 0.0
    # small steps keep the ReLU pattern fixed; one random direction per layer
    eps = 1e-6
    bounds = [0, pcnet.NW1, pcnet.NW1 + pcnet.NW2, pcnet.NWEIGHTS]
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        v = np.zeros(pcnet.NWEIGHTS)
        v[lo:hi] = rng.standard_normal(hi - lo)
        fp, _ = pcnet.sample_loss_grad(w + eps * v, hpsi, h, shape, g, x, Z[:1], theta, lam)
        fm, _ = pcnet.sample_loss_grad(w - eps * v, hpsi, h, shape, g, x, Z[:1], theta, lam)
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - grad @ v) < 1e-6 * abs(fd) + 1e-9
