"""NumPy oracle of the learned CG preconditioner (test infrastructure only; it
is the checker the HIP kernels are compared against and has no reference
counterpart -- ARTED's CG_ompk applies no preconditioner,
ARTED/GS/CG.f90:85-92).

Network: three 3x3x3 periodic convolutions without biases, channels
2 -> 16 -> 16 -> 2, ReLU after the first two.  Weights W[t][ci][co], tap
t = (dx+1)*9 + (dy+1)*3 + (dz+1) reads the point (ix+dx, iy+dy, iz+dz).

Preconditioned direction of a residual g (NL complex):
  phi = arg(sum g^2) / 2, u = g e^{-i phi} sqrt(NL)/|g|   (re, im channels)
  d   = (y_re + i y_im) e^{i phi},  y = net(u)
so d(c g) = c d(g) for every complex c != 0 (up to the sign of e^{i phi}).

Training objective of one sample (g, x, band ib, theta = <x|H|x>, target
eigenvalue lam) on a system with lower bands Z_j (j < ib):
  q = d - sum_{z in {Z_j, x}} z <z|d> h,  p = q / sqrt(h |q|^2)
  c = Re<p|Hp> h,  b = <g|p> h,  R = sqrt((theta-c)^2 + 4|b|^2)
  ev = (theta + c - R)/2,  gain = (theta - ev)/(theta - lam)
  loss = -mean over samples of min(gain, 1)
i.e. the fraction of the band's eigenvalue error removed by one 2x2
Rayleigh-Ritz step in span{x, p} (CG.f90:107-113).
"""
from __future__ import annotations

import numpy as np

NW1, NW2, NW3 = 27 * 2 * 16, 27 * 16 * 16, 27 * 16 * 2
NWEIGHTS = NW1 + NW2 + NW3


def split(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[:NW1].reshape(27, 2, 16), w[NW1:NW1 + NW2].reshape(27, 16, 16),
            w[NW1 + NW2:].reshape(27, 16, 2))


def taps():
    return [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def _shift(a, d):
    """a(ix+dx, iy+dy, iz+dz) on the periodic grid; a has shape (NLx, NLy, NLz, C)."""
    return np.roll(a, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2))


def conv(W, a):
    out = np.zeros(a.shape[:3] + (W.shape[2],))
    for t, d in enumerate(taps()):
        out += _shift(a, d) @ W[t]
    return out


def conv_back(W, a_in, g_out):
    """(grad wrt a_in, grad wrt W) of out = conv(W, a_in)."""
    gin = np.zeros_like(a_in)
    gW = np.zeros_like(W)
    for t, d in enumerate(taps()):
        sa = _shift(a_in, d)
        gW[t] = np.einsum("xyzi,xyzo->io", sa, g_out)
        gin += _shift(g_out @ W[t].T, (-d[0], -d[1], -d[2]))
    return gin, gW


def forward(w, u):
    """u (NLx, NLy, NLz, 2) -> (y, a1, a2)."""
    W1, W2, W3 = split(w)
    a1 = np.maximum(conv(W1, u), 0.0)
    a2 = np.maximum(conv(W2, a1), 0.0)
    return conv(W3, a2), a1, a2


def backward(w, u, a1, a2, gy):
    W1, W2, W3 = split(w)
    ga2, gW3 = conv_back(W3, a2, gy)
    ga2 = ga2 * (a2 > 0)
    ga1, gW2 = conv_back(W2, a1, ga2)
    ga1 = ga1 * (a1 > 0)
    _, gW1 = conv_back(W1, u, ga1)
    return np.concatenate([gW1.ravel(), gW2.ravel(), gW3.ravel()])


def canon(g):
    s2 = np.sum(g * g)
    ph = np.exp(0.5j * np.angle(s2))
    return g / ph * np.sqrt(g.size) / np.linalg.norm(g), ph


def direction(w, g, shape):
    gc, ph = canon(g)
    u = np.stack([gc.real, gc.imag], -1).reshape(shape + (2,))
    y, _, _ = forward(w, u)
    return (y[..., 0] + 1j * y[..., 1]).ravel() * ph


def sample_loss_grad(w, hpsi, h, shape, g, x, Z, theta, lam):
    """Loss contribution -min(gain,1) of one sample and its weight gradient.
    hpsi(v) applies H; Z: lower bands (n, NL) (may be empty)."""
    gc, ph = canon(g)
    u = np.stack([gc.real, gc.imag], -1).reshape(shape + (2,))
    y, a1, a2 = forward(w, u)
    d = (y[..., 0] + 1j * y[..., 1]).ravel() * ph
    V = list(Z) + [x]
    q = d - sum(z * (np.vdot(z, d) * h) for z in V)
    nu = np.sqrt(h * np.vdot(q, q).real)
    p = q / nu
    hp = hpsi(p)
    c = np.vdot(p, hp).real * h
    b = np.vdot(g, p) * h
    R = np.sqrt((theta - c) ** 2 + 4 * abs(b) ** 2)
    ev = 0.5 * (theta + c - R)
    gain = (theta - ev) / (theta - lam)
    if gain >= 1.0:
        return -1.0, np.zeros(NWEIGHTS)
    dl_dev = 1.0 / (theta - lam)                       # d(-gain)/d ev
    dev_dc = 0.5 * (1.0 - (c - theta) / R)
    dev_db2 = -1.0 / R
    alpha = dl_dev * dev_dc * 2 * h
    beta = dl_dev * dev_db2 * 2 * h * b
    gp = alpha * hp + beta * g
    r = np.vdot(p, gp).real
    gq = (gp - p * h * r) / nu
    gd = gq - sum(z * (np.vdot(z, gq) * h) for z in V)
    gyc = (gd / ph).reshape(shape)
    gy = np.stack([gyc.real, gyc.imag], -1)
    return -gain, backward(w, u, a1, a2, gy)


def init_weights(seed):
    """He-scaled normal weights (fan-in 27*cin)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.standard_normal(NW1) * np.sqrt(2.0 / 54),
                           rng.standard_normal(NW2) * np.sqrt(2.0 / 432),
                           rng.standard_normal(NW3) * np.sqrt(1.0 / 432)])
