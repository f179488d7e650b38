"""The 1D macroscopic field of the multiscale loop (salmon_amd/multiscale.py,
after ARTED/FDTD/FDTD.f90:50-113, 373-408 and control_ms.f90), pinned by
what the mathematics fixes: at the magic time step c dt = HX the leapfrog
wave equation carries a pulse one cell per step exactly, so after n
iterations the field is the initial pulse shifted by n cells (a wrong sign,
factor or time level in the update, or an incident pulse set up moving the
wrong way, fails it); and a uniform field with a uniform current gains
exactly -4 pi dt^2 J per step."""
import numpy as np
import pytest

from salmon_amd import multiscale as ms


def test_vacuum_pulse_moves_one_cell_per_step_at_the_magic_time_step():
    dt = 0.08
    hx = ms.C_LIGHT * dt
    m = ms.Maxwell1D(-60, 40, hx, dt, macropoints=[])
    tw = 24 * dt  # the pulse spans 24 cells
    m.init_acos2(f0=0.01, omega=2 * np.pi / (6 * dt), pulse_tw=tw, epdir_re=(0.0, 0.6, 0.8), phi_cep=0.1)
    a0 = m.ac[m.interior].copy()
    assert np.count_nonzero(a0[:, 2]) > 15 and not np.any(a0[:, 0])
    nsteps = 12
    ms.run(m, nsteps, lambda acs: np.zeros((0, 3)))
    a = m.ac[m.interior]
    shifted = np.zeros_like(a0)
    shifted[nsteps:] = a0[:-nsteps]
    assert np.max(np.abs(a - shifted)) < 1e-13 * np.max(np.abs(a0))
    # and Ac_new holds the next time level
    shifted1 = np.zeros_like(a0)
    shifted1[nsteps + 1:] = a0[:-nsteps - 1]
    assert np.max(np.abs(m.ac_new[m.interior] - shifted1)) < 1e-13 * np.max(np.abs(a0))


def test_uniform_current_source():
    dt, hx = 0.05, 40.0
    m = ms.Maxwell1D(1, 9, hx, dt, macropoints=[])
    a = np.tile([0.0, 0.02, -0.03], (9, 1))
    m.set_fields(a, a)
    j0 = np.array([0.0, 1e-3, 2e-3])
    m.jm_new[m.interior] = j0
    ms.run(m, 1, lambda acs: np.zeros((0, 3)))
    assert np.allclose(m.ac_new[m.interior], a - 4 * np.pi * dt**2 * j0, rtol=0, atol=1e-16)


def test_macro_points_read_and_feed_the_field():
    dt, hx = 0.05, 30.0
    m = ms.Maxwell1D(-3, 5, hx, dt, macropoints=[1, 3])
    x = np.arange(-3, 6)
    m.set_fields(np.stack([0 * x, 0.01 * np.sin(x), 0.02 * np.cos(x)], 1),
                 np.stack([0 * x, 0.01 * np.sin(x - 0.5), 0.02 * np.cos(x - 0.5)], 1))
    seen = []

    def micro(acs):
        seen.append(acs.copy())
        return np.array([[0.0, 1e-3, 0.0], [0.0, 0.0, -1e-3]])

    ms.run(m, 2, micro)
    # first iteration: Ac_m is the field set for t = dt at the macro-points
    assert np.allclose(seen[0][:, 0, 1], 0.01 * np.sin(np.array([1, 3]) - 0.5))
    # the currents of iteration 1 are the source of iteration 2's update
    assert np.allclose(m.jm[m.row([1, 3])], [[0.0, 1e-3, 0.0], [0.0, 0.0, -1e-3]])
    with pytest.raises(ValueError, match="outside"):
        ms.Maxwell1D(0, 4, hx, dt, macropoints=[7])
