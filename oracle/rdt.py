"""Rapid-distortion (Kelvin mode) closed form of the shear-periodic box.

Test infrastructure: used by ``tests/`` only.

A single Fourier mode of a solenoidal velocity in the uniform shear U = S y
is an exact solution of the Navier-Stokes equations (u.grad u = 0, so u x
omega is a gradient and drops out of the curl).  Its wave vector tilts,
k(t) = (k0x, k0y - S k0x t, k0z), and its amplitude follows ``rdtfunc``
(operators.f90:1516-1562), restated in ``rdt_velocity``.  ``kelvin_state``
writes that mode as the time-advance state (omega_y, phi = lap v, U, W)
on the lines of a box, and ``kelvin_lines`` the state the closed form
predicts at time t, both on the lab-frame y grid y_j = j dy - Ly/2 that the
shear-periodic lines use (twist exp(-i kx xofft) with xofft = S Ly t).
``StreamwiseMode`` is the kx = 0 case (lift-up, ``liftup_velocity``), with
optional sinusoidal mean profiles U(y), W(y) that diffuse viscously.
"""
import math

import numpy as np


def rdt_velocity(u0, v0, w0, k0x, k0y, k0z, nu, t):
    """(u1, u2, u3) and the wave vector (kx, ky, kz) at time t for S = 1
    (operators.f90:1516-1562, including the viscous factor)."""
    s = 1.0
    kx, ky, kz = k0x, k0y - s * k0x * t, k0z
    div = k0x * u0 + k0y * v0 + k0z * w0
    if abs(div) > 1e-12:
        raise ValueError("initial mode is not solenoidal: k.u = %g" % abs(div))
    alp20 = k0x ** 2 + k0z ** 2
    alp0 = math.sqrt(alp20)
    kk0 = k0x ** 2 + k0y ** 2 + k0z ** 2
    kk = kx ** 2 + ky ** 2 + kz ** 2
    dpsi = math.atan(ky / alp0) - math.atan(k0y / alp0)
    u2 = kk0 / kk * v0
    u1 = (-kx * ky / kk / alp20 + kx * k0y / kk0 / alp20) * v0 \
        + kz * kz / kx / alp20 * dpsi / alp0 * v0
    u1 = u1 * kk0 + u0
    u3 = (-kz * ky / kk / alp20 + kz * k0y / kk0 / alp20) * v0 - kz * dpsi / alp20 / alp0 * v0
    u3 = u3 * kk0 + w0
    damp = math.exp(-nu * t * (kk0 - s * t * k0y * k0x + (s * t * k0x) ** 2 / 3.0))
    return (u1 * damp, u2 * damp, u3 * damp), (kx, ky, kz)


class KelvinMode:
    """One Kelvin mode at the (kx index i, kz index k) line of a grid, with
    n periods over Ly at t = 0 and velocity amplitude amp."""

    def __init__(self, g, i=1, k=1, n=1, amp=0.1):
        if i < 1:
            raise ValueError("the closed form needs kx != 0")
        self.g, self.i, self.k = g, i, k
        self.k0x, self.k0z = float(g.kx[i]), float(g.kz[k])
        self.k0y = 2.0 * math.pi * n / g.Ly
        self.v0 = amp * (1.0 + 0.3j)
        self.u0 = amp * (0.2 - 0.5j)
        self.w0 = -(self.k0x * self.u0 + self.k0y * self.v0) / self.k0z

    def lines(self, t):
        """(omega_y, phi) lines of the closed form at time t"""
        g = self.g
        (u1, u2, u3), (kx, ky, kz) = rdt_velocity(self.u0, self.v0, self.w0, self.k0x, self.k0y,
                                                  self.k0z, 1.0 / g.Re, t)
        e = np.exp(1j * ky * g.y)
        return (1j * kz * u1 - 1j * kx * u3) * e, -(kx * kx + ky * ky + kz * kz) * u2 * e

    def state(self):
        """(vor, phi, u00, w00) at t = 0 in the [y][kx][kz] layout"""
        g = self.g
        vor, phi = g.zeros(), g.zeros()
        vor[:, self.i, self.k], phi[:, self.i, self.k] = self.lines(0.0)
        return vor, phi, np.zeros(g.my), np.zeros(g.my)

    def error(self, vor, phi, t):
        """relative L2 errors of the mode's omega_y and phi lines at time t,
        and the norm of omega_y outside the mode relative to the mode"""
        ev, ep = self.lines(t)
        lv, lp = vor[:, self.i, self.k], phi[:, self.i, self.k]
        rv = np.linalg.norm(lv - ev) / np.linalg.norm(ev)
        rp = np.linalg.norm(lp - ep) / np.linalg.norm(ep)
        rest = math.sqrt(max(np.linalg.norm(vor) ** 2 - np.linalg.norm(lv) ** 2, 0.0))
        return rv, rp, rest / np.linalg.norm(ev)


def liftup_velocity(u0, v0, w0, ky, kz, nu, t, s=1.0):
    """A streamwise-independent mode (kx = 0) in the shear: its wave vector
    does not tilt and the pressure term 2 S kx v k/|k|^2 vanishes, so
    v, w decay viscously and u grows by lift-up, u = (u0 - S v0 t) e^(-nu|k|^2 t)."""
    if abs(ky * v0 + kz * w0) > 1e-12:
        raise ValueError("initial mode is not solenoidal")
    d = math.exp(-nu * (ky * ky + kz * kz) * t)
    return ((u0 - s * v0 * t) * d, v0 * d, w0 * d)


class StreamwiseMode:
    """The kx = 0 mode at kz index k (and its Hermitian partner at mz - k),
    n periods over Ly, velocity amplitude amp; optional mean profiles
    U = A sin(kappa y), W = A cos(kappa y + 0.4), kappa = 2 pi m / Ly, which
    decay as exp(-nu kappa^2 t)."""

    def __init__(self, g, k=1, n=1, amp=0.1, A=0.0, m=1):
        if not 1 <= k <= g.nz:
            raise ValueError("kz index must be in 1..nz")
        self.g, self.k = g, k
        self.kz, self.ky = float(g.kz[k]), 2.0 * math.pi * n / g.Ly
        self.v0 = amp * (1.0 + 0.3j)
        self.u0 = amp * (0.2 - 0.5j)
        self.w0 = -self.ky * self.v0 / self.kz
        self.A, self.kap = float(A), 2.0 * math.pi * m / g.Ly

    def lines(self, t):
        g = self.g
        u1, u2, _ = liftup_velocity(self.u0, self.v0, self.w0, self.ky, self.kz, 1.0 / g.Re, t, g.s)
        e = np.exp(1j * self.ky * g.y)
        return 1j * self.kz * u1 * e, -(self.ky ** 2 + self.kz ** 2) * u2 * e

    def means(self, t):
        g = self.g
        d = math.exp(-self.kap ** 2 * t / g.Re)
        return self.A * d * np.sin(self.kap * g.y), self.A * d * np.cos(self.kap * g.y + 0.4)

    def state(self):
        g = self.g
        vor, phi = g.zeros(), g.zeros()
        lv, lp = self.lines(0.0)
        vor[:, 0, self.k], phi[:, 0, self.k] = lv, lp
        vor[:, 0, g.mz - self.k], phi[:, 0, g.mz - self.k] = np.conj(lv), np.conj(lp)
        u00, w00 = self.means(0.0)
        return vor, phi, u00, w00

    def error(self, vor, phi, u00, w00, t):
        """relative L2 errors: omega_y and phi of the mode's line, of its
        Hermitian partner line, and, with profiles, of U and W"""
        g = self.g
        ev, ep = self.lines(t)
        rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
        out = [rel(vor[:, 0, self.k], ev), rel(phi[:, 0, self.k], ep),
               rel(vor[:, 0, g.mz - self.k], np.conj(ev))]
        if self.A:
            eu, ew = self.means(t)
            out += [rel(u00, eu), rel(w00, ew)]
        return out
