"""Pins the composition of the oracle's RK3 step (Cross.run: the order of the
phase maps, the offset update, the RK3 constants, the twist of each substep,
the viscous terms) with the rapid-distortion closed form of a Kelvin mode.

* ``rdt_velocity`` (rdtfunc, operators.f90:1516-1562) is itself checked
  against a high-accuracy integration of the linear Kelvin-mode equations
  du/dt = -S v e_x + 2 S kx v k/|k|^2 - nu |k|^2 u,  k(t) = (kx, k0y - S kx t, kz).
* A run from the mode converges to it at the order of the time scheme:
  3 with explicit viscous terms, 2 with the implicit (Crank-Nicolson-type)
  stages.  Runs end exactly at t = 0.5 through the etime clipping.

A wrong moving-frame factor, offset update, twist or RK3 constant leaves an
O(dt) or O(1) error (checked on planted mistakes while building this test:
all gave errors >= 1e-3 at CFL 0.125 against < 1e-8 here).
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle.cross import Cross
from oracle.grid import Grid
from oracle.rdt import KelvinMode, rdt_velocity

T_END = 0.5


def _kelvin_ode(u0, v0, w0, k0, nu, t):
    k0x, k0y, k0z = k0

    def rhs(tt, y):
        u = y[0:3] + 1j * y[3:6]
        k = np.array([k0x, k0y - k0x * tt, k0z])
        kk = k @ k
        du = 2.0 * k0x * u[1] * k / kk - nu * kk * u
        du[0] -= u[1]
        return np.concatenate([du.real, du.imag])

    y0 = np.array([u0, v0, w0])
    sol = solve_ivp(rhs, (0.0, t), np.concatenate([y0.real, y0.imag]), method="DOP853",
                    rtol=1e-13, atol=1e-16)
    y = sol.y[:, -1]
    return y[0:3] + 1j * y[3:6]


@pytest.mark.parametrize("k0,t", [((math.pi, 2 * math.pi, 2 * math.pi), 0.7),
                                  ((2 * math.pi, -4 * math.pi, 6 * math.pi), 1.3)])
def test_rdt_closed_form_solves_the_kelvin_equations(k0, t):
    nu = 1.0 / 300.0
    v0, u0 = 0.1 + 0.03j, 0.02 - 0.05j
    w0 = -(k0[0] * u0 + k0[1] * v0) / k0[2]
    got, kt = rdt_velocity(u0, v0, w0, *k0, nu, t)
    ref = _kelvin_ode(u0, v0, w0, k0, nu, t)
    assert np.linalg.norm(np.array(got) - ref) < 1e-10 * np.linalg.norm(ref)
    assert kt[1] == pytest.approx(k0[1] - k0[0] * t)
    # the closed form stays solenoidal at the tilted wave vector
    assert abs(np.dot(kt, got)) < 1e-13 * np.linalg.norm(ref) * np.linalg.norm(kt)


def kelvin_run(CFL, explicit, grid=(16, 33, 16), Re=300.0):
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=CFL)
    mode = KelvinMode(g, amp=0.1)
    c = Cross(g, explicit=explicit)
    c.set_state(*mode.state())
    c.etime = T_END
    c.run(100000)
    assert c.time == pytest.approx(T_END, abs=1e-14)
    assert c.xofft == pytest.approx(g.s * g.Ly * T_END, rel=1e-13)
    return mode.error(c.vor, c.phi, c.time)


@pytest.mark.parametrize("explicit,order", [(1, 3), (0, 2)])
def test_kelvin_mode_converges_at_scheme_order(explicit, order):
    e1 = kelvin_run(0.25, explicit)
    e2 = kelvin_run(0.125, explicit)
    ratio = e1[0] / e2[0]
    assert 0.75 * 2 ** order < ratio < 1.3 * 2 ** order
    assert e2[1] < 2 * e2[0] + 1e-10
    assert e2[0] < (1e-8 if explicit else 1e-6)
    assert e2[2] < 1e-7           # nothing leaks out of the mode


def kelvin_expected_record(mode, t, s=1.0):
    """UNER entries (cross.f90:2067-2119) of the closed-form Kelvin mode at
    time t: per-plane means of the real field 2 Re(a exp(i theta)), whose
    square averages to 2|a|^2.  The pseudo-dissipation sum_ij |d_j u_i|^2
    of the mode is 2|k|^2|u|^2, plus S^2 from the mean shear; the
    production -S <uv> + S^2/Re."""
    g = mode.g
    (u1, u2, u3), k = rdt_velocity(mode.u0, mode.v0, mode.w0, mode.k0x, mode.k0y, mode.k0z,
                                   1.0 / g.Re, t)
    u = np.array([u1, u2, u3])
    k = np.array(k)
    om = 1j * np.cross(k, u)
    c = lambda a, b: 2.0 * (a * np.conj(b)).real
    uner = np.concatenate([np.sqrt(2.0) * np.abs(u), [c(u[0], u[1]), c(u[0], u[2]), c(u[1], u[2])],
                           np.sqrt(2.0) * np.abs(om), [s * s + 2.0 * (k @ k) * np.sum(np.abs(u) ** 2)]])
    uner2 = np.array([-s * c(u[0], u[1]) + s * s / g.Re, c(om[0], om[1])])
    return uner, uner2


def check_kelvin_history(mode, history, stats, spec=None):
    """UNER records (time, Deltat, uner(10), CFL, CFLv, umax(1:3), xoffb,
    xofft, umax(0), uner2(4)), per-plane statistics stats(23, my) and the
    spectra sp(8, kxn, mz) of a Kelvin-mode run against the closed form at
    each record's time."""
    g = mode.g
    assert len(history) >= 80
    acc = np.zeros(12)
    sp = np.zeros(8)
    for rec in history:
        t = rec[0]
        uner, uner2 = kelvin_expected_record(mode, t)
        (u1, u2, u3), k = rdt_velocity(mode.u0, mode.v0, mode.w0, mode.k0x, mode.k0y, mode.k0z,
                                       1.0 / g.Re, t)
        u = np.array([u1, u2, u3])
        om = 1j * np.cross(np.array(k), u)
        c = lambda a, b: 2.0 * (a * np.conj(b)).real * g.my     # kx > 0: twice, all planes
        sp += [c(u[0], u[0]), c(u[1], u[1]), c(u[2], u[2]), c(u[0], u[1]),
               c(om[0], om[0]), c(om[1], om[1]), c(om[2], om[2]), c(om[0], om[1])]
        got = rec[2:12]
        scale = np.array([1, 1, 1, 0, 0, 0, 1, 1, 1, 1]) * np.abs(uner) + \
            np.array([0, 0, 0, 1, 1, 1, 0, 0, 0, 0]) * uner[0] ** 2
        assert np.all(np.abs(got - uner) <= 1e-7 * scale), (t, got, uner)
        assert np.all(np.abs(rec[20:22] - uner2) <= 1e-7 * np.abs(uner2))
        assert rec[12] == g.CFL
        assert rec[17] == pytest.approx(-g.s * g.Ly * t, abs=1e-14)
        assert rec[18] == pytest.approx(g.s * g.Ly * t, abs=1e-14)
        acc[0:3] += np.abs(uner[0:3]) ** 2        # stats 3..5: sums of <u_i^2>
        acc[3:6] += uner[3:6]                     # stats 6..8: <u_i u_j>
        acc[6:9] += uner[6:9] ** 2                # stats 12..14
        acc[9] += uner[9]                         # stats 15: dissipation
        acc[10:12] += uner2                       # stats 19, 20
    rows = [3, 4, 5, 6, 7, 8, 12, 13, 14, 15, 19, 20]
    for r, a in zip(rows, acc):
        assert np.allclose(stats[r], a, rtol=1e-7, atol=1e-9 * abs(acc[0]))
    # the means of a single mode vanish; its triple products too, up to the
    # O(1e-7) harmonics that the compact y-derivative error lets the
    # nonlinear terms generate
    for r in (0, 1, 2, 9, 10, 11):
        assert np.max(np.abs(stats[r])) < 1e-12
    cube = len(history) * g.mgalx * g.mgalz * (np.sqrt(2.0) * abs(mode.v0)) ** 3
    for r in (16, 17, 18):
        assert np.max(np.abs(stats[r])) < 1e-6 * cube
    if spec is not None:
        got = spec[:, mode.i, mode.k]
        scale = np.array([1, 1, 1, 0, 1, 1, 1, 0]) * np.abs(sp) + \
            np.array([0, 0, 0, 1, 0, 0, 0, 1]) * max(sp[0], sp[1])
        scale[3] = max(sp[0], sp[1])
        scale[7] = max(sp[4], sp[5])
        assert np.all(np.abs(got - sp) <= 1e-7 * scale), (got, sp)
        rest = spec.copy()
        rest[:, mode.i, mode.k] = 0.0
        assert np.max(np.abs(rest[[0, 1, 2, 4, 5, 6]])) < 1e-6 * np.max(sp)


def test_kelvin_mode_history_records():
    g = Grid(16, 33, 16, Re=300.0, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.125)
    mode = KelvinMode(g, amp=0.1)
    c = Cross(g, explicit=1)
    c.set_state(*mode.state())
    c.etime = T_END
    c.run(100000)
    assert len(c.history) == c.nacum
    check_kelvin_history(mode, c.history, c.stats, c.spec)
    c.reset_statistics()
    assert c.nacum == 0 and not np.any(c.stats) and not np.any(c.spec)


def test_liftup_closed_form_solves_the_kelvin_equations():
    from oracle.rdt import liftup_velocity
    nu, t, ky, kz = 1.0 / 300.0, 0.9, 2 * math.pi, 4 * math.pi
    u0, v0 = 0.02 - 0.05j, 0.1 + 0.03j
    w0 = -ky * v0 / kz
    got = liftup_velocity(u0, v0, w0, ky, kz, nu, t)
    ref = _kelvin_ode(u0, v0, w0, (0.0, ky, kz), nu, t)
    assert np.linalg.norm(np.array(got) - ref) < 1e-11 * np.linalg.norm(ref)


def streamwise_run(CFL, explicit, amp=0.1, A=0.0, nfix=None, grid=(16, 33, 16), Re=300.0):
    """A run from the kx = 0 mode (and mean profiles); nfix > 0 fixes the
    step to T_END/nfix through the dump-time clipping (cross.f90:1686-1698)."""
    from oracle.rdt import StreamwiseMode
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=CFL)
    mode = StreamwiseMode(g, amp=amp, A=A)
    c = Cross(g, explicit=explicit)
    c.set_state(*mode.state())
    c.etime = T_END
    if nfix:
        c.dumptime = c.dtimag = T_END / nfix
    c.run(100000)
    assert c.time == pytest.approx(T_END, abs=1e-14)
    return mode.error(c.vor, c.phi, c.u00, c.w00, c.time), (c.vor, c.phi, c.u00, c.w00)


@pytest.mark.parametrize("explicit,order", [(1, 3), (0, 2)])
def test_streamwise_mode_converges(explicit, order):
    """the kx = 0 plane: KMM reconstruction with kx = 0, the Hermitian
    partner line, lift-up by the mean shear"""
    e1, _ = streamwise_run(0.25, explicit)
    e2, _ = streamwise_run(0.125, explicit)
    assert e1[0] / e2[0] > 0.75 * 2 ** order
    assert e2[0] < (1e-9 if explicit else 1e-7)
    assert e2[1] < (1e-9 if explicit else 1e-7)
    assert e2[2] == pytest.approx(e2[0], rel=1e-6)


@pytest.mark.parametrize("explicit,nfix,tol", [(1, None, 1e-8), (0, 50, 2e-8)])
def test_mean_profiles_diffuse(explicit, nfix, tol):
    """the 00 path (U, W with their RK3 accumulators and viscous terms):
    sinusoidal profiles decay as exp(-nu kappa^2 t).  A weak kx = 0 mode
    keeps the run above the laminar stop; its coupling to the profiles is
    below 3e-9 here.  A wrong 00-mode constant gives errors of order
    nu kappa^2 dt ~ 1e-3."""
    e, _ = streamwise_run(0.25, explicit, amp=1.5e-5, A=1e-3, nfix=nfix)
    assert e[3] < tol and e[4] < tol
