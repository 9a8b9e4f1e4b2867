"""Pins of the oracle's x-z plane algebra by mathematics, not by restating it.

* fftxz: textbook DFT sums with the 3/2-rule mode map (fou3D_thread.f:22-161)
  and the forward normalisation 1/(mgalx*mgalz);
* derived (KMM reconstruction, cross.f90:1221-1276): fields built from an
  analytic solenoidal velocity with exact y-derivatives must come back as
  that velocity and its curl;
* hvhg products (cross.f90:1453-1490): u . (u x (omega + 2 chi z)) = 0 at
  every point, so, u being inside the retained band, the spectral inner
  product of u with the dealiased H vanishes on every plane (Parseval);
* set_conj (operators.f90:1382-1429): the result transforms to a real field
  with no kx=0 imaginary part discarded.
"""
import math

import numpy as np
import pytest

from oracle import fftxz
from oracle.cross import Cross, set_conj
from oracle.grid import Grid


def _kzw(mz):
    nz = (mz - 1) // 2
    k = np.arange(mz)
    return np.where(k <= nz, k, k - mz)


@pytest.mark.parametrize("mgalx,mgalz", [(12, 12), (18, 24), (24, 18)])
def test_fftxz_textbook(mgalx, mgalz):
    kxn, mz = mgalx // 3, 2 * (mgalz // 3) - 1
    rng = np.random.default_rng(mgalx * 100 + mgalz)
    F = rng.normal(size=(kxn, mz)) + 1j * rng.normal(size=(kxn, mz))
    a = np.arange(mgalx)[:, None]
    b = np.arange(mgalz)[None, :]
    kzw = _kzw(mz)
    # backward: kx=0 column contributes its real part after the z sum
    ref = np.zeros((mgalx, mgalz))
    g0 = sum(F[0, k] * np.exp(2j * np.pi * kzw[k] * b / mgalz) for k in range(mz))
    ref += np.real(g0) * np.ones((mgalx, 1))
    for kx in range(1, kxn):
        for k in range(mz):
            ref += 2.0 * np.real(F[kx, k] * np.exp(2j * np.pi * (kx * a / mgalx + kzw[k] * b / mgalz)))
    got = fftxz.to_phys(F, mgalx, mgalz)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    # forward: normalised sum over the physical grid, retained modes only
    f = rng.normal(size=(mgalx, mgalz))
    fw = np.zeros((kxn, mz), dtype=np.complex128)
    for kx in range(kxn):
        for k in range(mz):
            fw[kx, k] = np.sum(f * np.exp(-2j * np.pi * (kx * a / mgalx + kzw[k] * b / mgalz)))
    fw /= mgalx * mgalz
    assert np.max(np.abs(fftxz.to_spec(f, kxn, mz) - fw)) < 1e-13


def test_set_conj_gives_exact_roundtrip():
    mgalx, mgalz = 24, 24
    kxn, mz = 8, 15
    nz = (mz - 1) // 2
    rng = np.random.default_rng(3)
    F = rng.normal(size=(1, kxn, mz)) + 1j * rng.normal(size=(1, kxn, mz))
    F[0, 0, 0] = F[0, 0, 0].real
    set_conj(F, nz)
    back = fftxz.to_spec(fftxz.to_phys(F, mgalx, mgalz), kxn, mz)
    assert np.max(np.abs(back - F)) < 1e-14


def test_kmm_reconstruction_is_velocity_and_curl():
    g = Grid(24, 33, 24, Re=100.0, alp=math.pi, gam=2 * math.pi, Ly=1.0)
    rng = np.random.default_rng(11)
    shape = (g.my, g.kxn, g.mz)
    y = g.y[:, None, None]
    kx = g.kx[None, :, None]
    kz = g.kz[None, None, :]
    k2 = kx ** 2 + kz ** 2
    u, v, w, duy, dv, dwy, d2v, d2uy, d2wy = (np.zeros(shape, complex) for _ in range(9))
    for n in (1, 2, -3):
        ky = 2 * np.pi * n / g.Ly
        e = np.exp(1j * ky * y)
        A = rng.normal(size=shape[1:]) + 1j * rng.normal(size=shape[1:])
        C = rng.normal(size=shape[1:]) + 1j * rng.normal(size=shape[1:])
        with np.errstate(divide="ignore", invalid="ignore"):
            vv = np.where(k2 > 0, A, 0.0)
            # i kx u + i ky v + i kz w = 0
            uu = np.where(k2 > 0, -kx * ky * vv / k2 + kz * C, 0.0)
            ww = np.where(k2 > 0, -kz * ky * vv / k2 - kx * C, 0.0)
        uu[0, 0, 0] = rng.normal()      # real mean profiles U(y), W(y)
        ww[0, 0, 0] = rng.normal()
        for f, d1, d2, amp in ((u, duy, d2uy, uu), (v, dv, d2v, vv), (w, dwy, d2wy, ww)):
            f += amp * e
            d1 += 1j * ky * amp * e
            d2 += -(ky ** 2) * amp * e
    for f in (u, w, duy, dwy):
        f[:, 0, 0] = f[:, 0, 0].real
    ix, iz = 1j * kx, 1j * kz
    ox = dwy - iz * v
    oy = iz * u - ix * w
    oz = ix * v - duy
    doy = iz * duy - ix * dwy                  # d(omega_y)/dy
    phi = d2v - k2 * v                         # lap v
    c = Cross(g)
    c.u00, c.w00 = u[:, 0, 0].real.copy(), w[:, 0, 0].real.copy()
    u1, u2, u3, o1, o2, o3 = c.derived(phi, oy, v, dv, doy, duy[:, 0, 0].real, dwy[:, 0, 0].real)
    scale = np.max(np.abs(u))
    for got, want in ((u1, u), (u2, v), (u3, w), (o2, oy)):
        assert np.max(np.abs(got - want)) < 1e-12 * scale
    oscale = np.max(np.abs(oz))
    assert np.max(np.abs(o1 - ox)) < 1e-12 * oscale
    assert np.max(np.abs(o3 - oz)) < 1e-12 * oscale


@pytest.mark.parametrize("chi", [0.0, 0.3])
def test_dealiased_lamb_vector_is_orthogonal_to_u(chi):
    from tests.boxes import box_ic
    g = Grid(24, 25, 24, Re=100.0, alp=math.pi, gam=2 * math.pi, Ly=1.0, chi=chi)
    vor, phi, u00, w00 = box_ic(24, 25, 24, uprim=0.5)
    c = Cross(g)
    c.set_state(vor, phi, u00, w00)
    # v, dv/dy, d(omega_y)/dy from the compact operators, as at step start
    shp = np.ones(g.kxn, complex)
    k2 = g.k2()
    c.hg = c._lines_helm(c.phi, k2, shp, shp)
    c.hg[:, 0, 0] = 0.0
    c.hv = c._lines_d(c.hg, shp, shp, 1)
    c.dvordy = c._lines_d(c.vor, shp, shp, 1)
    dudy = c.ops.deriv_real(c.u00, 1)[:, 0]
    dwdy = c.ops.deriv_real(c.w00, 1)[:, 0]
    u1, u2, u3, *_ = c.derived(set_conj(c.phi.copy(), g.nz), set_conj(c.vor.copy(), g.nz),
                               c.hg, c.hv, c.dvordy, dudy, dwdy)
    c.istep = 5
    c.hvhg(2, 0)
    h1, h2, h3 = c.last_H
    wgt = np.full(g.kxn, 2.0)
    wgt[0] = 1.0
    inner = sum(np.sum(wgt[None, :, None] * np.real(np.conj(a) * b), axis=(1, 2))
                for a, b in ((set_conj(u1, g.nz), h1), (set_conj(u2, g.nz), h2),
                             (set_conj(u3, g.nz), h3)))
    norm = sum(np.sum(wgt[None, :, None] * np.abs(a) * np.abs(b), axis=(1, 2))
               for a, b in ((u1, h1), (u2, h2), (u3, h3)))
    assert np.max(np.abs(inner) / norm) < 1e-13


def test_hist_second_moments_are_plane_means():
    """'f' statistics (cross.f90:1798-1860) by Parseval: the weighted spectral
    sums equal the x-z means of the physical products."""
    g = Grid(24, 9, 24, Re=100.0, alp=math.pi, gam=2 * math.pi, Ly=1.0)
    rng = np.random.default_rng(21)
    f = []
    for _ in range(6):
        a = rng.normal(size=(g.my, g.kxn, g.mz)) + 1j * rng.normal(size=(g.my, g.kxn, g.mz))
        a[:, 0, 0] = a[:, 0, 0].real
        f.append(set_conj(a, g.nz))
    u1, u2, u3, o1, o2, o3 = f
    c = Cross(g)
    c.Deltat = 0.5
    ener, ener2 = np.zeros(10), np.zeros(4)
    c._hist_f(u1, u2, u3, o1, o2, o3, u2, ener, ener2)
    P = [fftxz.to_phys(a, g.mgalx, g.mgalz) for a in f]
    mean = lambda a, b: np.mean(a * b, axis=(1, 2))
    pairs = {3: (0, 0), 4: (1, 1), 5: (2, 2), 6: (0, 1), 7: (0, 2), 8: (1, 2),
             12: (3, 3), 13: (4, 4), 14: (5, 5), 20: (3, 4)}
    for idx, (a, b) in pairs.items():
        want = mean(P[a], P[b])
        assert np.max(np.abs(c.stats[idx] - want)) < 1e-12 * np.max(np.abs(want)), idx
        assert np.max(np.abs(c.stats2[idx] - 0.5 * want)) < 1e-12 * np.max(np.abs(want)), idx
    for idx, a in zip((0, 1, 2, 9, 10, 11), range(6)):
        assert np.allclose(c.stats[idx], np.mean(P[a], axis=(1, 2)), rtol=0, atol=1e-13)
