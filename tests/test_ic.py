"""The seeded initial field (shear_nnf_amd/ic.py, the model of getini_tophat,
initcr.f90:954-1071) against what it is specified to be: a random solenoidal
velocity whose every Fourier mode carries the energy of the spectrum
E(k) ~ k^4 exp(-2 (k/k0)^2) spread over its shell (E / (4 pi k^2)), k in
units of the z fundamental, k0 = 4, normalised to u'_rms = uprim.

The check rebuilds the velocity from the returned (omega_y, phi, U, W) by
an independent route: an FFT in y (the field is a sum of y-Fourier modes
|n| <= ny), v = -phi / |k|^2, and u, w from omega_y = i kz u - i kx w and
continuity kx u + ky v + kz w = 0.
"""
import math

import numpy as np
import pytest

from shear_nnf_amd.ic import make_ic
from tests.boxes import ALP, GAM, LY


def _modes(mgalx, my, mgalz, seed, uprim):
    vor, phi, u00, w00 = make_ic(mgalx, my, mgalz, ALP, GAM, LY, seed=seed, uprim=uprim)
    kxn, mz = vor.shape[1], vor.shape[2]
    nz = (mz - 1) // 2
    ny = ((2 * (my // 3) - 1) - 1) // 2
    beta = 2 * math.pi / LY
    n = np.arange(-ny, ny + 1)
    # y_j = j dy - Ly/2: the FFT coefficient of mode n carries (-1)^n
    sgn = (-1.0) ** n
    fy = lambda f: np.fft.fft(f, axis=0)[n % my] / my * (sgn if f.ndim == 1 else sgn[:, None, None])
    omy, phs = fy(vor), fy(phi)
    # nothing outside |n| <= ny
    full = np.fft.fft(vor, axis=0) / my
    mask = np.ones(my, bool)
    mask[n % my] = False
    assert np.abs(full[mask]).max(initial=0.0) < 1e-14 * np.abs(full).max()
    kx = ALP * np.arange(kxn)
    kz = np.array([GAM * k if k <= nz else -GAM * (mz - k) for k in range(mz)])
    KY, KX, KZ = np.meshgrid(beta * n, kx, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    kh2 = KX ** 2 + KZ ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(k2 > 0, -phs / k2, 0.0)
        u = np.where(kh2 > 0, (-KY * KX * v - 1j * KZ * omy) / kh2, 0.0)
        w = np.where(kh2 > 0, (-KY * KZ * v + 1j * KX * omy) / kh2, 0.0)
    u[:, 0, 0] = fy(u00)
    w[:, 0, 0] = fy(w00)
    return n, KX, KY, KZ, u, v, w


@pytest.mark.parametrize("grid,uprim", [((32, 33, 32), 0.1), ((48, 49, 64), 0.25)])
def test_ic_spectrum_and_normalisation(grid, uprim):
    n, KX, KY, KZ, u, v, w = _modes(*grid, seed=20260001, uprim=uprim)
    e = np.abs(u) ** 2 + np.abs(v) ** 2 + np.abs(w) ** 2
    # u'_rms^2 = <u^2 + v^2 + w^2> / 3, kx > 0 modes counted twice (conjugates)
    wgt = np.where(KX > 0, 2.0, 1.0)
    assert np.sum(wgt * e) / 3.0 == pytest.approx(uprim ** 2, rel=1e-12)
    # every mode: energy E(k) / (4 pi k^2) with one constant
    kn = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) / GAM
    on = kn > 0
    shape = kn[on] ** 4 * np.exp(-2.0 * (kn[on] / 4.0) ** 2) / (4 * math.pi * kn[on] ** 2)
    c = np.sum(e[on] * shape) / np.sum(shape * shape)
    # the y-FFT rebuild leaves rounding of ~1e-16 of the largest mode on
    # every mode, so the far tail of the spectrum is checked in absolute terms
    assert np.all(np.abs(e[on] - c * shape) <= 1e-9 * e[on] + 1e-13 * e.max())
    assert e[~on].max() < 1e-30 * e.max()
    # the (kx, kz) = (0, 0) modes point along x or z only (continuity with
    # ky != 0 leaves v = 0 there)
    assert np.abs(v[:, 0, 0]).max() == 0.0


def test_ic_phases_depend_on_the_seed():
    a = make_ic(32, 33, 32, ALP, GAM, LY, seed=1)[0]
    b = make_ic(32, 33, 32, ALP, GAM, LY, seed=2)[0]
    c = make_ic(32, 33, 32, ALP, GAM, LY, seed=1)[0]
    assert np.array_equal(a, c)
    assert np.linalg.norm(a - b) > 0.5 * np.linalg.norm(a)
