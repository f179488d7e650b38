"""Pins of the CPU oracle: the reference's own Fortran (golden fixtures), closed
forms, invariants and brute force.  CPU only."""
import math

import numpy as np
import pytest

from helpers import load_case, rel
from oracle.cpu import Oracle, kac_at, midpoint_kac
from salmon_amd import synth

CASES = ["small_k", "noncubic", "grid8"]


# ---------------------------------------------------------------- reference
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fortran(name):
    """The C restatement reproduces the reference's Fortran on the same inputs."""
    s, d = load_case(name)
    o = Oracle(s)
    ik = int(d["hpsi_ik"])
    h, t = o.hpsi(d["hpsi_k"], s.zu[ik, s.NB - 1], with_tt=True)
    assert rel(h, d["hpsi_out"]) < 1e-13
    assert rel(t, d["hpsi_tt"]) < 1e-13
    z1 = o.hamiltonian(d["taylor_kac"], s.zu, s.dt)
    assert rel(z1, d["taylor_zu"]) < 1e-13
    k1 = d["obs_kac"]
    assert rel(o.current(k1, z1), d["current"]) < 1e-12
    assert rel(o.density(z1), d["rho"]) < 1e-13
    ek, en, _ = o.energy(k1, z1)
    assert rel([ek, en], d["energy"]) < 1e-13
    assert rel(o.sp_energy(s.kAc0, s.zu), d["esp"]) < 1e-13
    assert rel(o.gram_schmidt(s.zu), d["gs_zu"]) < 1e-13
    zc, ev, _ = o.cg(s.kAc0, s.zu, 4)
    assert rel(ev, d["cg_esp_var"]) < 1e-10
    assert rel(zc, d["cg_zu"]) < 1e-11
    nsteps = d["rt_J"].shape[0]
    r = o.rt_run(d["ac"], nsteps)
    assert rel(r["J"], d["rt_J"]) < 1e-12
    assert rel(r["E"], d["rt_E"]) < 1e-13
    assert rel(r["zu"], d["rt_zu"]) < 1e-13


def test_oracle_matches_reference_cfg1_trajectory():
    """100 Taylor steps of the 16^3 Gamma workload: J(t), E(t), band norms."""
    s, d = load_case("cfg1")
    o = Oracle(s)
    r = o.rt_run(d["ac"], 100)
    assert rel(r["J"], d["rt_J"]) < 1e-11
    assert rel(r["E"], d["rt_E"]) < 1e-12
    bn = np.sum(np.abs(r["zu"]) ** 2, axis=2) * s.Hxyz
    assert rel(bn, d["rt_band_norm"]) < 1e-12
    assert rel(r["zu"][:, :2], d["rt_zu_head"]) < 1e-11


def test_fd_weights_closed_form():
    lap, nab = synth.fd_weights()
    assert np.array_equal(lap, [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    assert np.array_equal(nab, [0.0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])
    s, d = load_case("noncubic")  # reference fd_coef on a non-cubic cell
    ref = d["ref_fd"]
    for i, name in enumerate(("lapx", "lapy", "lapz", "nabx", "naby", "nabz")):
        assert np.allclose(ref[i], getattr(s, name), rtol=4e-16, atol=0)


# ---------------------------------------------------------------- closed forms
def _no_kb(s):
    t = s.copy()
    t.uV = np.zeros_like(s.uV)
    return t


@pytest.mark.parametrize("N,L", [((16, 16, 16), 10.26), ((10, 12, 14), (7.0, 8.0, 9.0))])
def test_hpsi_plane_wave_eigen_identity(N, L):
    """For psi = exp(iG.r) on the grid, H psi / psi is the FD symbol + |k|^2/2 + V."""
    s = _no_kb(synth.make_system(N=N, L=L, NB=1, seed=1, orbitals=False))
    o = Oracle(s)
    ix, iy, iz = synth.grid_coords(*N)
    m = (1, -2, 3)
    k = np.array([0.13, -0.21, 0.05])
    psi = np.exp(2j * math.pi * (m[0] * ix / N[0] + m[1] * iy / N[1] + m[2] * iz / N[2]))
    ratio = o.hpsi(k, psi) / psi
    expect = 0.5 * k @ k + s.vloc
    for d, (lap, nab) in enumerate(((s.lapx, s.nabx), (s.lapy, s.naby), (s.lapz, s.nabz))):
        th = 2 * math.pi * m[d] / N[d]
        expect = expect - 0.5 * (lap[0] + 2 * sum(lap[n] * math.cos(n * th) for n in range(1, 5)))
        expect = expect + k[d] * 2 * sum(nab[n] * math.sin(n * th) for n in range(1, 5))
    assert np.max(np.abs(ratio - expect)) < 1e-12


def test_hpsi_hermitian():
    s, _ = load_case("noncubic")
    o = Oracle(s)
    k = s.kAc0[:, 1] + 0.07
    a, b = s.zu[1, 0], s.zu[1, 1]
    lhs = np.vdot(b, o.hpsi(k, a))
    rhs = np.conj(np.vdot(a, o.hpsi(k, b)))
    assert abs(lhs - rhs) < 1e-13 * abs(lhs) + 1e-13


def test_kb_projector_eigen_and_gauge():
    """V_NL u_l = c_l u_l for a diamond atom's own projector at k=0 (s and p
    are orthogonal on a symmetric sphere), and V_NL(k=G) = e^{-iGr} V_NL(0) e^{iGr}."""
    s = synth.make_system(N=(16, 16, 16), L=10.26, NB=1, seed=5, orbitals=False)
    o, o0 = Oracle(s), Oracle(_no_kb(s))
    k0 = np.zeros(3)
    for lma in (0, 1, 6):
        a = s.a_tbl[lma] - 1
        u = np.zeros(s.NL, np.complex128)
        u[s.Jxyz[a, : s.Mps[a]] - 1] = s.uV[lma, : s.Mps[a]]
        vnl = o.hpsi(k0, u) - o0.hpsi(k0, u)
        c = s.iuV[lma] * s.Hxyz * np.sum(s.uV[lma] ** 2)
        assert np.max(np.abs(vnl - c * u)) < 1e-12 * abs(c) * np.max(np.abs(u))
    # reciprocal-lattice gauge covariance of the nonlocal part
    ix, iy, iz = synth.grid_coords(16, 16, 16)
    G = 2 * math.pi / 10.26 * np.array([1.0, -2.0, 1.0])
    e = np.exp(1j * (G[0] * ix * s.Hx + G[1] * iy * s.Hy + G[2] * iz * s.Hz))
    psi = np.random.default_rng(0).standard_normal(s.NL) + 0j
    lhs = o.hpsi(G, psi) - o0.hpsi(G, psi)
    rhs = np.conj(e) * (o.hpsi(k0, e * psi) - o0.hpsi(k0, e * psi))
    assert np.max(np.abs(lhs - rhs)) < 1e-12 * np.max(np.abs(rhs))


def _dense_h(o, k, NL):
    H = np.zeros((NL, NL), np.complex128)
    for i in range(NL):
        e = np.zeros(NL, np.complex128)
        e[i] = 1.0
        H[:, i] = o.hpsi(k, e)
    return H


@pytest.fixture(scope="module")
def tiny():
    s = synth.make_system(N=(9, 9, 9), L=6.0, nk3=(1, 1, 1), NB=4, seed=7, atoms="random", natoms=2, rc=1.6,
                          sigma=0.6, dt=0.05)
    o = Oracle(s)
    k = s.kAc0[:, 0] + np.array([0.1, -0.05, 0.02])
    return s, o, k, _dense_h(o, k, s.NL)


def test_taylor_equals_dense_polynomial(tiny):
    s, o, k, H = tiny
    kac = k[:, None]
    z1 = o.hamiltonian(kac, s.zu, s.dt)
    M = -1j * s.dt * H
    P = np.eye(s.NL) + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24
    assert rel(z1[0], (P @ s.zu[0].T).T) < 1e-13


def test_taylor_eigenstate_norm_decay(tiny):
    """|R(x)|^2 = 1 - x^6/72 + x^8/576 per step for an eigenstate (x = dt*eps)."""
    s, o, k, H = tiny
    w, v = np.linalg.eigh(H)
    for idx in (0, s.NL // 2, s.NL - 1):
        t = s.copy()
        t.zu = np.tile(v[:, idx][None, None, :], (1, s.NB, 1))
        z = t.zu
        for _ in range(10):
            z = o.hamiltonian(k[:, None], z, s.dt)
        x = s.dt * w[idx]
        expect = (1 - x ** 6 / 72 + x ** 8 / 576) ** 10
        got = np.sum(np.abs(z[0, 0]) ** 2)
        assert abs(got - expect) < 1e-12


def test_hermitian_dense(tiny):
    _, _, _, H = tiny
    assert np.max(np.abs(H - H.conj().T)) < 1e-13 * np.max(np.abs(H))


def test_current_hellmann_feynman():
    """J_d = (1/Omega) sum occ d<psi|H(k)|psi>Hxyz/dk_d at fixed psi (all three terms)."""
    s, _ = load_case("noncubic")
    o = Oracle(s)
    kac = kac_at(s, [0.03, -0.02, 0.05])
    J = o.current(kac, s.zu)

    def etot(kk):
        esp = o.sp_energy(kk, s.zu)
        return float(np.sum(esp * s.occ))

    h = 1e-5
    for d in range(3):
        dk = np.zeros((3, 1))
        dk[d] = h
        fd = (etot(kac + dk) - etot(kac - dk)) / (2 * h) / s.aLxyz
        assert abs(fd - J[d]) < 1e-8 * max(1.0, abs(J[d])) + 1e-9


def test_density_sum_rule():
    s, _ = load_case("small_k")
    o = Oracle(s)
    rho = o.density(s.zu)
    assert np.all(rho >= 0)
    assert abs(rho.sum() * s.Hxyz - o.norm(s.zu)) < 1e-12 * o.norm(s.zu)
    # single orbital: rho = occ |psi|^2
    t = s.copy()
    t.occ = np.zeros_like(s.occ)
    t.occ[1, 2] = 0.7
    rho1 = Oracle(t).density(s.zu)
    assert np.max(np.abs(rho1 - 0.7 * np.abs(s.zu[1, 2]) ** 2)) < 1e-14


def test_energy_matches_expectation_of_h():
    """Ekin + Enl + Eloc = sum occ <psi|H|psi> Hxyz for normalised orbitals."""
    s, _ = load_case("noncubic")
    o = Oracle(s)
    kac = kac_at(s, [0.02, 0.01, -0.03])
    ek, en, el = o.energy(kac, s.zu)
    esp = o.sp_energy(kac, s.zu)
    assert abs((ek + en + el) - np.sum(esp * s.occ)) < 1e-12 * abs(ek)


def test_gram_schmidt_orthonormal_same_span():
    s, _ = load_case("small_k")
    o = Oracle(s)
    raw = np.random.default_rng(1).standard_normal(s.zu.shape) + 1j * np.random.default_rng(2).standard_normal(s.zu.shape)
    q = o.gram_schmidt(raw)
    for ik in range(s.NK):
        Q = q[ik].T
        assert np.max(np.abs(Q.conj().T @ Q * s.Hxyz - np.eye(s.NB))) < 1e-13
        Qn, _ = np.linalg.qr(raw[ik].T)
        Pq = Q @ Q.conj().T * s.Hxyz
        assert np.max(np.abs(Pq - Qn @ Qn.conj().T)) < 1e-12
        first = raw[ik, 0] / np.sqrt(np.sum(np.abs(raw[ik, 0]) ** 2) * s.Hxyz)
        assert np.max(np.abs(q[ik, 0] - first)) < 1e-14


def test_cg_converges_to_dense_eigenvalues(tiny):
    s, o, k, H = tiny
    """The GS sweep of ground_state.f90:136-140 (Gram-Schmidt, subspace
    diagonalisation, Gram-Schmidt, CG) reaches the lowest eigenvalues of the
    dense H."""
    w = np.linalg.eigvalsh(H)[: s.NB]
    sweeps = {}
    for precond in (0, 1):
        z = s.zu
        for sw in range(80):
            z = o.gram_schmidt(z)
            Z = z[0].T
            hs = Z.conj().T @ H @ Z * s.Hxyz
            _, u = np.linalg.eigh(hs)
            z = (Z @ u).T[None]
            z = o.gram_schmidt(z)
            z, ev, _ = o.cg(k[:, None], z, 8, precond=precond)
            esp = o.sp_energy(k[:, None], z)[0]
            if np.max(np.abs(np.sort(esp) - w)) < 1e-8:
                break
        assert np.max(np.abs(np.sort(esp) - w)) < 1e-8
        sweeps[precond] = sw + 1
    # the kinetic preconditioner converges in fewer sweeps than plain CG
    assert sweeps[1] < sweeps[0]


def test_acos2_pulse_closed_form():
    A0, w, T = 0.05, 0.057, 826.8
    dt = 0.08
    ac = synth.acos2_pulse(2000, dt, A0, w, T)
    t = np.arange(2001) * dt
    tt = t - T / 2
    expect = -A0 * np.cos(math.pi * tt / T) ** 2 * np.sin(w * tt)
    expect[np.abs(tt) >= T / 2] = 0.0
    assert np.max(np.abs(ac[:, 0] - expect)) < 1e-15
    assert np.all(ac[:, 1:] == 0) and ac[0, 0] == 0.0


@pytest.mark.parametrize("maker", [lambda: synth.config1()[0], lambda: synth.config2(nk3=(8, 8, 8), NB=1,
                                                                                      orbitals=False)[0]])
def test_stability_margin(maker):
    """dt * rho(H) stays inside the Taylor-4 stability region x < 2 sqrt 2."""
    s = maker()
    assert s.dt * synth.spectral_bound(s) < 2 * math.sqrt(2)


def test_kinetic_preconditioner_closed_form():
    """K e^{iG.r} = k(x) e^{iG.r}, x = t(G)/Tx with t the finite-difference
    kinetic symbol and k(x) = f/(f+16x^4), f = 27+18x+12x^2+8x^3 (Teter-Payne-
    Allan form) on a non-cubic, non-power-of-two grid; K is linear and
    Hermitian positive definite."""
    s = synth.make_system(N=(10, 12, 14), L=(5.0, 6.3, 7.1), nk3=(1, 1, 1), NB=1, seed=2, name="kin")
    o = Oracle(s)
    ix, iy, iz = synth.grid_coords(s.NLx, s.NLy, s.NLz)
    lap = [s.lapx, s.lapy, s.lapz]
    N = [s.NLx, s.NLy, s.NLz]
    tx = 1.7
    for n in [(0, 0, 0), (1, 0, 0), (2, -3, 5), (5, 6, 7), (-4, 1, -6)]:
        pw = np.exp(2j * np.pi * (n[0] * ix / N[0] + n[1] * iy / N[1] + n[2] * iz / N[2]))
        t = 0.0
        for a in range(3):
            th = 2 * np.pi * n[a] / N[a]
            t += -0.5 * (lap[a][0] + 2 * sum(lap[a][m] * np.cos(m * th) for m in range(1, 5)))
        x = t / tx
        f = 27 + 18 * x + 12 * x ** 2 + 8 * x ** 3
        assert np.max(np.abs(o.kinetic_precondition(tx, pw) - f / (f + 16 * x ** 4) * pw)) < 1e-13
    rng = np.random.default_rng(0)
    a = rng.standard_normal(s.NL) + 1j * rng.standard_normal(s.NL)
    b = rng.standard_normal(s.NL) + 1j * rng.standard_normal(s.NL)
    ka, kb = o.kinetic_precondition(tx, a), o.kinetic_precondition(tx, b)
    assert rel(o.kinetic_precondition(tx, 2 * a - 3j * b), 2 * ka - 3j * kb) < 1e-13
    assert abs(np.vdot(b, ka) - np.conj(np.vdot(a, kb))) < 1e-12 * abs(np.vdot(b, ka))
    assert np.vdot(a, ka).real > 0


def test_config4_generator():
    """The 64-system batch: every system draws its grid from the five shapes,
    bands in 8..64, k points in 1..64 and atoms in 1..16, deterministically;
    the grid spacing keeps dt * rho(H) inside the Taylor-4 region."""
    seen = set()
    for i in range(64):
        s = synth.config4_system(i, orbitals=False)
        seen.add((s.NLx, s.NLy, s.NLz))
        assert (s.NLx, s.NLy, s.NLz) in synth.CFG4_SHAPES
        assert 8 <= s.NB <= 64 and 1 <= s.NK <= 64 and 1 <= s.NI <= 16
        assert s.occ.shape == (s.NK, s.NB) and abs(s.occ.sum() - 2 * s.NB) < 1e-12
    assert seen == set(synth.CFG4_SHAPES)
    a, b = synth.config4_system(7, orbitals=False), synth.config4_system(7, orbitals=False)
    assert np.array_equal(a.vloc, b.vloc) and a.NB == b.NB and a.NK == b.NK
    s = synth.config4_system(2, orbitals=False)
    assert s.dt * synth.spectral_bound(s) < 2 * math.sqrt(2)


def test_pz_matches_reference_and_its_derivative():
    """The oracle's PZ81 LDA against the reference's own exc_cor_pz (built in
    oracle/_ref, called with rho_s = rho/2 as salmon_xc.f90:499-511 does), and
    v_xc = d(rho e_xc)/drho by central differences (both sides of rs = 1)."""
    from oracle import cpu
    rho = np.concatenate([np.geomspace(1e-6, 0.2, 40), np.geomspace(0.25, 30.0, 40)])
    v, e = cpu.pz(rho)
    h = 1e-6 * rho
    vp, ep = cpu.pz(rho + h)
    vm, em = cpu.pz(rho - h)
    fd = ((rho + h) * ep - (rho - h) * em) / (2 * h)
    assert np.max(np.abs(fd - v) / np.abs(v)) < 1e-7
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build (oracle/_ref) not present")
    import ctypes
    L = ctypes.CDLL(ref.LIB)
    f = L._QMbuiltin_pzPexc_cor_pz
    n = ctypes.c_int(rho.size)
    rs = np.ascontiguousarray(rho * 0.5)
    exc, eexc, vexc = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    f(ctypes.byref(n), p(rs), p(exc), p(eexc), p(vexc))
    assert np.max(np.abs(vexc - v)) < 1e-14 * np.max(np.abs(v))
    assert np.max(np.abs(exc - e)) < 1e-14 * np.max(np.abs(e))


def test_hartree_closed_form():
    """V_H of a plane-wave density cos(G.r) is 4 pi/|G|^2 cos(G.r); a constant
    density (G = 0 only) gives zero; linear in rho."""
    from oracle import cpu
    s = synth.make_system(N=(12, 10, 14), L=(6.0, 5.5, 7.2), nk3=(1, 1, 1), NB=1, seed=2, name="hart",
                          orbitals=False)
    ix, iy, iz = np.meshgrid(np.arange(s.NLx), np.arange(s.NLy), np.arange(s.NLz), indexing="ij")
    for n in ((1, 0, 0), (2, -3, 1), (0, 5, -7)):
        G = 2 * np.pi * np.array(n) / np.array([s.aLx, s.aLy, s.aLz])
        ph = (G[0] * ix * s.aLx / s.NLx + G[1] * iy * s.aLy / s.NLy + G[2] * iz * s.aLz / s.NLz).ravel()
        vh = cpu.hartree(s, np.cos(ph))
        assert np.max(np.abs(vh - 4 * np.pi / (G @ G) * np.cos(ph))) < 1e-12 * 4 * np.pi / (G @ G)
    assert np.max(np.abs(cpu.hartree(s, np.full(s.NL, 0.7)))) < 1e-13
    rng = np.random.default_rng(1)
    a, b = rng.random(s.NL), rng.random(s.NL)
    assert rel(cpu.hartree(s, 2 * a - b), 2 * cpu.hartree(s, a) - cpu.hartree(s, b)) < 1e-12
