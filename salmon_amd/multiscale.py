"""Multiscale Maxwell-TDDFT along one axis: the macroscopic vector potential
Ac(x, t) on a 1D grid, propagated by the wave equation with the matter
current of the macro-points as source; each macro-point is a crystal
(a context of this package) driven by its own Ac(t).

The loop of one iteration follows ARTED/control/control_ms.f90:205-440
(FDTDdim = '1D', no ions, trans_mat = identity):
  proceed    Ac_old <- Ac, Ac <- Ac_new, Ac_new <- 0, Jm <- Jm_new, Jm_new <- 0
             (control_ms_raman.f90:807-866)
  evolve     Ac_new = 2 Ac - Ac_old - 4 pi dt^2 Jm - (c dt)^2 rot rot Ac on the
             interior points (FDTD/FDTD.f90:373-408)
  assign     Ac_m, Ac_new_m of each macro-point (control_ms_raman.f90:869-880)
  micro      every macro-point: one RT step from Ac_m to Ac_new_m (midpoint
             k, dt_evolve.f90:85) and its current at kAc0 + Ac_new_m
             (control_ms.f90:302-323) -- here one batched launch per phase
             over all macro-points (salmon_gpu_batch_rt_run_ms)
  map        Jm_new at each macro-point's grid point (control_ms.f90:426-440)
The ghost points beyond the interior are not evolved: after the periodic
copy of the initial fields (FDTD.f90:297-318) they hold zero from the
second iteration on, as in the reference.
"""
from __future__ import annotations

import numpy as np

C_LIGHT = 137.03953250  # speed of light in a.u. (ARTED/modules/global_variables.f90:28)


class Maxwell1D:
    """Macro grid ix = nx1..nx2 (interior) with ``novlp`` ghost points on
    either side; ``macropoints`` are the grid indices of the macro-points."""

    def __init__(self, nx1, nx2, hx, dt, macropoints, novlp=1):
        if nx2 < nx1 or hx <= 0 or dt <= 0:
            raise ValueError("need nx2 >= nx1, hx > 0, dt > 0")
        self.nx1, self.nx2, self.hx, self.dt, self.novlp = int(nx1), int(nx2), float(hx), float(dt), int(novlp)
        self.mx1 = self.nx1 - self.novlp
        n = self.nx2 - self.nx1 + 1 + 2 * self.novlp
        self.mp = np.asarray(macropoints, dtype=np.int64)
        if self.mp.size and (self.mp.min() < self.nx1 or self.mp.max() > self.nx2):
            raise ValueError("macro-points outside the interior nx1..nx2")
        if len(set(self.mp.tolist())) != self.mp.size:
            raise ValueError("two macro-points on one grid point")
        self.ac_old = np.zeros((n, 3))
        self.ac = np.zeros((n, 3))
        self.ac_new = np.zeros((n, 3))
        self.jm = np.zeros((n, 3))      # used on the interior only
        self.jm_new = np.zeros((n, 3))
        self.iteration = 0

    # grid index -> array row
    def row(self, ix):
        return np.asarray(ix) - self.mx1

    @property
    def interior(self):
        return slice(self.novlp, self.novlp + self.nx2 - self.nx1 + 1)

    @property
    def x_index(self):
        return np.arange(self.mx1, self.mx1 + self.ac.shape[0])

    def _periodic_ghosts(self):
        g, i = self.novlp, self.interior
        for a in (self.ac, self.ac_new):
            core = a[i]
            a[:g] = core[-g:]
            a[-g:] = core[:g]

    def set_fields(self, ac0, ac1):
        """Ac at t = 0 and t = dt on the interior points, (nx2-nx1+1, 3) each
        (the init_acfield form of FDTD.f90:286-292), ghosts copied periodically."""
        self.ac[self.interior] = ac0
        self.ac_new[self.interior] = ac1
        self._periodic_ghosts()

    def init_acos2(self, f0, omega, pulse_tw, epdir_re=(0.0, 0.0, 1.0), epdir_im=(0.0, 0.0, 0.0), phi_cep=0.0,
                   npower=2):
        """The incident Acos^n pulse of FDTD.f90:50-113 (first pulse,
        Xstart = 5 HX): |x| < tw c/2 with x = (ix-1) HX + Xstart + tw c/2,
        transverse components y and z, at t = 0 and t = dt."""
        c = C_LIGHT
        xstart = 5.0 * self.hx
        ep = np.asarray(epdir_re, dtype=np.float64) + 1j * np.asarray(epdir_im, dtype=np.float64)
        ac0 = np.zeros((self.nx2 - self.nx1 + 1, 3))
        ac1 = np.zeros_like(ac0)
        for r, ix in enumerate(range(self.nx1, self.nx2 + 1)):
            x = (ix - 1) * self.hx + xstart + 0.5 * pulse_tw * c
            for out, xx in ((ac0, x), (ac1, x - self.dt * c)):
                if abs(xx) < 0.5 * pulse_tw * c:
                    env = f0 / omega * np.cos(np.pi * xx / (pulse_tw * c)) ** npower
                    ph = np.exp(1j * (omega * xx / c + phi_cep * 2.0 * np.pi))
                    for d in (2, 1):
                        out[r, d] = env * (ep[d] * ph).imag
        self.set_fields(ac0, ac1)

    # --- one iteration ------------------------------------------------------
    def proceed(self):
        self.ac_old, self.ac, self.ac_new = self.ac, self.ac_new, np.zeros_like(self.ac)
        self.jm, self.jm_new = self.jm_new, np.zeros_like(self.jm)

    def evolve_ac(self):
        i = self.interior
        a = self.ac
        lo, hi = i.start, i.stop
        rr = np.zeros((hi - lo, 3))
        rr[:, 1:] = -(a[lo + 1:hi + 1, 1:] - 2.0 * a[lo:hi, 1:] + a[lo - 1:hi - 1, 1:]) * (1.0 / self.hx**2)
        self.ac_new[i] = (2 * a[i] - self.ac_old[i] - self.jm[i] * 4.0 * np.pi * self.dt**2
                          - rr * (C_LIGHT * self.dt) ** 2)

    def macro_fields(self):
        """(nmacro, 2, 3): Ac_m and Ac_new_m of every macro-point."""
        r = self.row(self.mp)
        return np.stack([self.ac[r], self.ac_new[r]], axis=1)

    def put_currents(self, jm_m):
        self.jm_new[self.row(self.mp)] = np.asarray(jm_m, dtype=np.float64).reshape(-1, 3)

    def step(self, micro):
        """One iteration; micro(ac (nmacro, 2, 3)) -> currents (nmacro, 3)."""
        self.iteration += 1
        self.proceed()
        self.evolve_ac()
        acs = self.macro_fields()
        jm = micro(acs)
        self.put_currents(jm)
        return acs, np.asarray(jm, dtype=np.float64)


def run(maxwell, nsteps, micro):
    """nsteps iterations; returns dict(ac (nsteps, nmacro, 2, 3), jm
    (nsteps, nmacro, 3)) of the macro-points."""
    acs, jms = [], []
    for _ in range(nsteps):
        a, j = maxwell.step(micro)
        acs.append(a)
        jms.append(j)
    return dict(ac=np.array(acs), jm=np.array(jms))


def batch_micro(batch, dt):
    """The micro step of every macro-point as one batched GPU step."""
    def micro(acs):
        return batch.rt_run_ms(acs, 1, dt)["J"][:, 0, :]
    return micro
