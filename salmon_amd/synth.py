"""Seeded synthetic inputs for the ARTED real-time electron-dynamics path.

Everything here is set-up, done once on the host before the time loop, as in
the reference.  The construction follows the reference's preparation code:

* finite-difference weights ............ ARTED/preparation/fd_coef.f90:20-83
* grid index i = ix*NLy*NLz+iy*NLz+iz .. ARTED/preparation/init.f90:28-37
* uniform (Monkhorst-Pack) k grid ...... ARTED/preparation/init.f90:283-300 (Sym=1)
* occupations occ = 2/NKxyz*wk ......... ARTED/preparation/init.f90:92-96
* (k,band) flattening ikb = ik*NB+ib ... ARTED/preparation/init.f90:101-109
* projector sphere tables Mps/Jxyz/Jxx . atom/pp/prep_pp.f90:182-347, ARTED/preparation/prep_ps.f90:120
* signed KB strengths iuV = +-1 ......... ARTED/preparation/prep_ps.f90:196-208
* Acos2 vector-potential pulse ......... ARTED/RT/init_Ac.f90:61-83
* Gram-Schmidt of the initial orbitals . ARTED/GS/Gram_Schmidt.f90:32-55

The physics inputs that the reference reads from pseudopotential files are
replaced by a synthetic model (a frozen local potential made of plane-wave
modes and Gaussian s+p projectors); the layout of every table is the
reference's.

Array conventions (the C-ABI boundary is Fortran column-major): a numpy
array of C shape (a, b) holds the Fortran array X(b, a).  So ``zu`` of shape
(NK, NB, NL) is ``zu_t(NL, NB, NK)``, ``kAc0`` of shape (3, NK) is
``kAc0(NK, 3)``, ``Jxyz`` of shape (NI, Nps) is ``Jxyz(Nps, NI)`` and ``uV`` of
shape (Nlma, Nps) is ``uV(Nps, Nlma)``.
"""
from __future__ import annotations

import dataclasses
import math
from fractions import Fraction

import numpy as np

ND = 4  # stencil half-width, Global_Variables: Nd = 4


def fd_weights(nd: int = ND):
    """Central-difference weights of order 2*nd, exactly as fd_coef.f90:26-73.

    Returns (lap[0..nd], nab[0..nd]) unscaled (nab[0] = 0).  Computed in exact
    rationals following the same loops, then rounded once to float64.
    """
    s = Fraction(0)
    for i in range(-nd, nd + 1):
        if i == 0:
            continue
        for j in range(-nd, nd + 1):
            if j == i or j == 0:
                continue
            s += Fraction(1, i * j)
    lap = [s]
    nab = [Fraction(0)]
    for j in range(1, nd + 1):
        t = Fraction(1)
        for i in range(-nd, nd + 1):
            if i == j:
                continue
            t *= (j - i)
        ss = Fraction(0)
        for k in range(-nd, nd + 1):
            if k == j or k == 0:
                continue
            s0 = Fraction(1)
            for i in range(-nd, nd + 1):
                if i == k or i == j or i == 0:
                    continue
                s0 *= -i
            ss += s0
        lap.append(2 * ss / t)
        sn = Fraction(1)
        for i in range(-nd, nd + 1):
            if i == j or i == 0:
                continue
            sn *= -i
        nab.append(sn / t)
    return np.array([float(x) for x in lap]), np.array([float(x) for x in nab])


@dataclasses.dataclass
class System:
    """One periodic cell with its orbitals, in the reference's layout."""
    NLx: int
    NLy: int
    NLz: int
    aLx: float
    aLy: float
    aLz: float
    lapx: np.ndarray  # (5,) lap(0:4)/Hx**2
    lapy: np.ndarray
    lapz: np.ndarray
    nabx: np.ndarray  # (5,) nab(0:4)/Hx, nab(0) = 0
    naby: np.ndarray
    nabz: np.ndarray
    vloc: np.ndarray  # (NL,) float64
    # projectors
    Rion: np.ndarray  # (NI, 3)
    rc: float
    Mps: np.ndarray   # (NI,) int32
    Jxyz: np.ndarray  # (NI, Nps) int32, 1-based unpadded grid index
    Jxx: np.ndarray   # (NI, Nps) int32 image offsets -2..2
    Jyy: np.ndarray
    Jzz: np.ndarray
    a_tbl: np.ndarray  # (Nlma,) int32, 1-based atom of each projector
    uV: np.ndarray     # (Nlma, Nps) float64
    iuV: np.ndarray    # (Nlma,) int32 (+1/-1)
    # k points and orbitals
    kAc0: np.ndarray   # (3, NK)
    occ: np.ndarray    # (NK, NB)
    zu: np.ndarray     # (NK, NB, NL) complex128
    dt: float = 0.08
    name: str = ""

    @property
    def NL(self):
        return self.NLx * self.NLy * self.NLz

    @property
    def Hx(self):
        return self.aLx / self.NLx

    @property
    def Hy(self):
        return self.aLy / self.NLy

    @property
    def Hz(self):
        return self.aLz / self.NLz

    @property
    def Hxyz(self):
        return self.Hx * self.Hy * self.Hz

    @property
    def aLxyz(self):
        return self.aLx * self.aLy * self.aLz

    @property
    def NI(self):
        return int(self.Mps.shape[0])

    @property
    def Nps(self):
        return int(self.Jxyz.shape[1])

    @property
    def Nlma(self):
        return int(self.a_tbl.shape[0])

    @property
    def NK(self):
        return int(self.kAc0.shape[1])

    @property
    def NB(self):
        return int(self.occ.shape[1])

    def copy(self):
        return dataclasses.replace(self, zu=self.zu.copy())


def grid_coords(NLx, NLy, NLz):
    """Lx, Ly, Lz integer coordinates of each grid index (init.f90:28-37)."""
    ix, iy, iz = np.meshgrid(np.arange(NLx), np.arange(NLy), np.arange(NLz), indexing="ij")
    return ix.ravel(), iy.ravel(), iz.ravel()


def fd_coefficients(H):
    lap, nab = fd_weights()
    return lap / H ** 2, nab / H


def uniform_kgrid(nk3, aL):
    """Sym=1 uniform grid, init.f90:289-299.  Returns kAc0 as (3, NK)."""
    nkx, nky, nkz = nk3
    bL = [2.0 * math.pi / a for a in aL]
    ks = []
    for ix in range(1, nkx + 1):
        for iy in range(1, nky + 1):
            for iz in range(1, nkz + 1):
                ks.append((-bL[0] / 2 + (ix - 0.5) * bL[0] / nkx,
                           -bL[1] / 2 + (iy - 0.5) * bL[1] / nky,
                           -bL[2] / 2 + (iz - 0.5) * bL[2] / nkz))
    return np.ascontiguousarray(np.array(ks, dtype=np.float64).T)


def plane_wave_potential(rng, NL3, aL, nmodes=20, gmax=3.0, amp=0.5):
    """Real, inversion-symmetric local potential V(r) = sum_m a_m cos(G_m.r).

    G_m = 2*pi/L * n_m with integer n_m, 0 < |n_m| <= gmax, one of each +-n
    pair, a_m ~ U[-amp, amp] Ha.
    """
    g = int(math.floor(gmax))
    cands = []
    for a in range(-g, g + 1):
        for b in range(-g, g + 1):
            for c in range(-g, g + 1):
                n = (a, b, c)
                if n == (0, 0, 0) or a * a + b * b + c * c > gmax * gmax:
                    continue
                if (-a, -b, -c) in cands:
                    continue
                cands.append(n)
    pick = rng.choice(len(cands), size=nmodes, replace=False)
    modes = np.array([cands[p] for p in pick], dtype=np.float64)
    coef = rng.uniform(-amp, amp, size=nmodes)
    ix, iy, iz = grid_coords(*NL3)
    H = [aL[d] / NL3[d] for d in range(3)]
    x, y, z = ix * H[0], iy * H[1], iz * H[2]
    v = np.zeros(ix.shape[0])
    for m in range(nmodes):
        ph = 2 * math.pi * (modes[m, 0] * x / aL[0] + modes[m, 1] * y / aL[1] + modes[m, 2] * z / aL[2])
        v += coef[m] * np.cos(ph)
    return v, modes, coef


def diamond_positions(aL):
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
    frac = np.concatenate([frac, frac + 0.25])
    return frac * np.asarray(aL)[None, :]


def projector_tables(Rion, rc, NL3, aL, sigma, coeff):
    """Sphere tables and Gaussian s+p KB projectors.

    The point order is that of calc_jxyz (prep_pp.f90:314-343): for each atom,
    image offsets ix,iy,iz in -2..2 (outer), then grid points in index order;
    a point is kept when r < rc.  Jxyz is the 1-based unpadded index
    (prep_ps.f90:120).  Projector lma of atom a is one of g, x g, y g, z g with
    g = exp(-r^2/(2 sigma^2)), normalised so that Hxyz * sum_j p_j^2 = 1;
    uV = sqrt(|c|) p and iuV = sign(c) (the reference folds sqrt|E_l| into uV and
    keeps the sign in iuV, input_pp.f90:75-88, prep_ps.f90:204).
    """
    NLx, NLy, NLz = NL3
    H = [aL[d] / NL3[d] for d in range(3)]
    Hxyz = H[0] * H[1] * H[2]
    ix, iy, iz = grid_coords(*NL3)
    NI = Rion.shape[0]
    pts = []
    for a in range(NI):
        rows = []
        for jx in range(-2, 3):
            for jy in range(-2, 3):
                for jz in range(-2, 3):
                    x = ix * H[0] - (Rion[a, 0] + jx * aL[0])
                    y = iy * H[1] - (Rion[a, 1] + jy * aL[1])
                    z = iz * H[2] - (Rion[a, 2] + jz * aL[2])
                    r = np.sqrt(x * x + y * y + z * z)
                    sel = np.nonzero(r < rc)[0]
                    for i in sel:
                        rows.append((i + 1, jx, jy, jz, x[i], y[i], z[i]))
        pts.append(rows)
    Mps = np.array([len(p) for p in pts], dtype=np.int32)
    Nps = int(Mps.max())
    Jxyz = np.zeros((NI, Nps), np.int32)
    Jxx = np.zeros((NI, Nps), np.int32)
    Jyy = np.zeros((NI, Nps), np.int32)
    Jzz = np.zeros((NI, Nps), np.int32)
    nproj = 4
    Nlma = NI * nproj
    uV = np.zeros((Nlma, Nps))
    iuV = np.zeros(Nlma, np.int32)
    a_tbl = np.zeros(Nlma, np.int32)
    for a in range(NI):
        rows = np.array(pts[a])
        m = rows.shape[0]
        Jxyz[a, :m] = rows[:, 0].astype(np.int32)
        Jxx[a, :m] = rows[:, 1].astype(np.int32)
        Jyy[a, :m] = rows[:, 2].astype(np.int32)
        Jzz[a, :m] = rows[:, 3].astype(np.int32)
        x, y, z = rows[:, 4], rows[:, 5], rows[:, 6]
        g = np.exp(-(x * x + y * y + z * z) / (2 * sigma * sigma))
        for p, f in enumerate((g, x * g, y * g, z * g)):
            f = f / math.sqrt(Hxyz * float(np.sum(f * f)))
            lma = a * nproj + p
            c = float(coeff[a, p])
            uV[lma, :m] = math.sqrt(abs(c)) * f
            iuV[lma] = 1 if c >= 0 else -1
            a_tbl[lma] = a + 1
    return Mps, Jxyz, Jxx, Jyy, Jzz, a_tbl, uV, iuV


def gram_schmidt(zu, Hxyz):
    """Modified Gram-Schmidt with the Hxyz-weighted product, per k.

    Gram_Schmidt_ompk (GS/Gram_Schmidt.f90:41-51): for band ib, subtract its
    overlap with each lower band in turn, then normalise.  Vectorised over k.
    """
    zu = np.array(zu, dtype=np.complex128, copy=True)
    NK, NB, NL = zu.shape
    for ib in range(NB):
        for ibt in range(ib):
            zov = np.einsum("kn,kn->k", zu[:, ibt].conj(), zu[:, ib]) * Hxyz
            zu[:, ib] -= zu[:, ibt] * zov[:, None]
        s = np.sum(np.abs(zu[:, ib]) ** 2, axis=1) * Hxyz
        zu[:, ib] /= np.sqrt(s)[:, None]
    return zu


def random_orbitals(rng, NK, NB, NL, Hxyz, orthonormalise=True):
    zu = rng.standard_normal((NK, NB, NL)) + 1j * rng.standard_normal((NK, NB, NL))
    if orthonormalise:
        zu = gram_schmidt(zu, Hxyz)
    return zu


def acos2_pulse(nsteps, dt, amplitude_A, omega, tpulse, pol=(1.0, 0.0, 0.0), phi_cep=0.0, t_delay=0.0):
    """Ac_ext(t_n), n = 0..nsteps, for ae_shape1='Acos2' (init_Ac.f90:74-83).

    amplitude_A is A0 = f0/omega (so f0 = A0*omega).  Returns (nsteps+1, 3).
    """
    f0 = amplitude_A * omega
    ep_re = np.asarray(pol, dtype=np.float64)
    ep_im = np.zeros(3)
    ac = np.zeros((nsteps + 1, 3))
    for it in range(nsteps + 1):
        tt = it * dt - 0.5 * tpulse - t_delay
        if abs(tt) < 0.5 * tpulse:
            env = math.cos(math.pi * tt / tpulse) ** 2
            ph = complex(math.cos(omega * tt + phi_cep * 2 * math.pi), math.sin(omega * tt + phi_cep * 2 * math.pi))
            ac[it, :] = -f0 / omega * env * np.imag((ep_re + 1j * ep_im) * ph)
    return ac


def constant_field(nsteps, A):
    return np.tile(np.asarray(A, dtype=np.float64)[None, :], (nsteps + 1, 1))


def spectral_bound(sys: System, kmax=None):
    """Upper bound of the spectral radius of H (for dt*rho(H) < 2*sqrt(2))."""
    th = np.linspace(0, 2 * math.pi, 2001)
    if kmax is None:
        kmax = np.max(np.abs(sys.kAc0), axis=1) if sys.NK else np.zeros(3)
    tot = 0.0
    for d, (lap, nab) in enumerate(((sys.lapx, sys.nabx), (sys.lapy, sys.naby), (sys.lapz, sys.nabz))):
        kin = -0.5 * (lap[0] + 2 * sum(lap[n] * np.cos(n * th) for n in range(1, 5)))
        grad = 2 * sum(nab[n] * np.sin(n * th) for n in range(1, 5))
        tot += float(np.max(kin + abs(kmax[d]) * np.abs(grad))) + 0.5 * float(kmax[d]) ** 2
    vmax = float(np.max(np.abs(sys.vloc)))
    # each projector has Hxyz*sum(uV^2) = |c|; an atom's block of V_NL is bounded
    # by the sum of its |c|, and a point sees the atoms whose spheres cover it.
    strength = (sys.uV ** 2).sum(axis=1) * sys.Hxyz
    per_atom = np.zeros(sys.NI)
    np.add.at(per_atom, sys.a_tbl - 1, strength)
    cover = np.zeros(sys.NL)
    for a in range(sys.NI):
        cover[np.unique(sys.Jxyz[a, :sys.Mps[a]] - 1)] += per_atom[a]
    return tot + vmax + float(cover.max())


# ---------------------------------------------------------------------------
# The five workloads
# ---------------------------------------------------------------------------

def make_system(N=(16, 16, 16), L=10.26, nk3=(1, 1, 1), NB=16, seed=0, atoms="diamond", natoms=8,
                sigma=0.8, rc=2.0, nmodes=20, amp=0.5, dt=0.08, orbitals=True, name=""):
    """Build one periodic cell with the reference's tables (see module docstring)."""
    rng = np.random.default_rng(seed)
    aL = (float(L), float(L), float(L)) if np.isscalar(L) else tuple(float(x) for x in L)
    NL3 = tuple(int(n) for n in N)
    H = [aL[d] / NL3[d] for d in range(3)]
    lapx, nabx = fd_coefficients(H[0])
    lapy, naby = fd_coefficients(H[1])
    lapz, nabz = fd_coefficients(H[2])
    vloc, _, _ = plane_wave_potential(rng, NL3, aL, nmodes=nmodes, amp=amp)
    if atoms == "diamond":
        Rion = diamond_positions(aL)
    else:
        Rion = rng.uniform(0.0, 1.0, size=(natoms, 3)) * np.asarray(aL)[None, :]
    coeff = rng.uniform(-2.0, 2.0, size=(Rion.shape[0], 4))
    Mps, Jxyz, Jxx, Jyy, Jzz, a_tbl, uV, iuV = projector_tables(Rion, rc, NL3, aL, sigma, coeff)
    kAc0 = uniform_kgrid(nk3, aL)
    NK = kAc0.shape[1]
    occ = np.full((NK, NB), 2.0 / NK)
    NL = NL3[0] * NL3[1] * NL3[2]
    Hxyz = H[0] * H[1] * H[2]
    if orbitals:
        zu = random_orbitals(rng, NK, NB, NL, Hxyz)
    else:
        zu = np.zeros((NK, NB, NL), np.complex128)
    return System(NLx=NL3[0], NLy=NL3[1], NLz=NL3[2], aLx=aL[0], aLy=aL[1], aLz=aL[2],
                  lapx=lapx, lapy=lapy, lapz=lapz, nabx=nabx, naby=naby, nabz=nabz, vloc=vloc,
                  Rion=Rion, rc=rc, Mps=Mps, Jxyz=Jxyz, Jxx=Jxx, Jyy=Jyy, Jzz=Jzz, a_tbl=a_tbl,
                  uV=uV, iuV=iuV, kAc0=kAc0, occ=occ, zu=zu, dt=dt, name=name)


def config1(seed=0):
    """16^3, L=10.26, Gamma only, 16 orbitals, A=(0.01,0,0), dt=0.08, 100 steps."""
    s = make_system(N=(16, 16, 16), L=10.26, nk3=(1, 1, 1), NB=16, seed=seed, name="cfg1")
    return s, constant_field(100, (0.01, 0.0, 0.0))


CFG2_PULSE = dict(amplitude_A=0.05, omega=0.057, tpulse=20.0 / 0.02418884254)


def config2(seed=0, nsteps=2000, nk3=(8, 8, 8), NB=32, orbitals=True):
    """16^3 cell, 32 bands x MP 8^3 (512 k), Acos2 pulse A0=0.05, w=0.057, 20 fs."""
    s = make_system(N=(16, 16, 16), L=10.26, nk3=nk3, NB=NB, seed=seed, orbitals=orbitals, name="cfg2")
    return s, acos2_pulse(nsteps, s.dt, **CFG2_PULSE)


def config5(seed=0, NB=64):
    """32^3, L=15, Gamma, 64 bands, 27 random-position atoms (ground-state CG)."""
    return make_system(N=(32, 32, 32), L=15.0, nk3=(1, 1, 1), NB=NB, seed=seed, atoms="random",
                       natoms=27, name="cfg5")


CFG4_SHAPES = [(12, 12, 12), (16, 16, 16), (20, 20, 28), (24, 24, 24), (32, 32, 32)]
CFG4_H = 10.26 / 16  # grid spacing of every system (keeps dt * rho(H) inside the Taylor-4 region)


def config4_system(i, scale=1.0, orbitals=True):
    """System i of the 64-system batch: grid from CFG4_SHAPES, bands
    U{8..64}, k points U{1..64} (the first NK points of a 4x4x4
    Monkhorst-Pack grid), atoms U{1..16} at random positions; seed i.
    scale < 1 shrinks bands and k points (for quick checks)."""
    rng = np.random.default_rng(10_000 + i)
    N = CFG4_SHAPES[int(rng.integers(len(CFG4_SHAPES)))]
    NB = max(1, int(round(int(rng.integers(8, 65)) * scale)))
    NK = max(1, int(round(int(rng.integers(1, 65)) * scale)))
    natoms = int(rng.integers(1, 17))
    L = tuple(n * CFG4_H for n in N)
    s = make_system(N=N, L=L, nk3=(4, 4, 4), NB=NB, seed=i, atoms="random", natoms=natoms, orbitals=False,
                    name=f"cfg4_{i}")
    s.kAc0 = np.ascontiguousarray(s.kAc0[:, :NK])
    s.occ = np.full((NK, NB), 2.0 / NK)
    if orbitals:
        s.zu = random_orbitals(np.random.default_rng(20_000 + i), NK, NB, s.NL, s.Hxyz, orthonormalise=False)
    else:
        s.zu = np.zeros((NK, NB, s.NL), np.complex128)
    return s
