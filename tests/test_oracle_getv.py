"""Pins the oracle's getv (operators.f90:735-793) and energy
(operators.f90:870-924) without the reference:

* getv: phi -> v -> phi is the identity on every line but (0,0) (lapcdy
  solves (R2 - k2 B2) v = B2 phi, visc3d applies B2^-1 (R2 - k2 B2));
  a single twisted Fourier mode of each line is an eigenvector of the
  twisted-cyclic operators, with eigenvalue R2(w)/B2(w) - k2 built from the
  stencils' symbols (a Fourier route independent of the banded solves);
  a smooth mode exp(i kappa y) is differentiated to the scheme's accuracy.
* energy: Parseval against the physical field on the padded x-z grid.
"""
import math

import numpy as np
import pytest

from oracle import fftxz
from oracle.cross import Cross, set_conj
from oracle.grid import Grid
from tests.boxes import ALP, GAM, LY, rel_l2


def _grid(mgalx=24, my=33, mgalz=24):
    return Grid(mgalx, my, mgalz, Re=800.0, alp=ALP, gam=GAM, Ly=LY, s=1.0)


def _field(g, rng):
    f = rng.normal(size=(g.my, g.kxn, g.mz)) + 1j * rng.normal(size=(g.my, g.kxn, g.mz))
    f[:, 0, 0] = f[:, 0, 0].real
    return set_conj(f, g.nz)


def _cross(g, xofft):
    c = Cross(g)
    c.xofft, c.xoffb = xofft, -xofft
    return c


def _sym(c, w):
    p = np.arange(-3, 4)
    return np.sum(c[:, None] * np.exp(1j * np.outer(p, w)), axis=0)


@pytest.mark.parametrize("xofft", [0.0, 0.37, 1.9])
def test_getv_round_trip(xofft):
    g = _grid()
    c = _cross(g, xofft)
    phi = _field(g, np.random.default_rng(11))
    v = c.getv(phi, 1)
    back = c.getv(v, -1)
    mask = np.ones((g.kxn, g.mz), bool)
    mask[0, 0] = False
    assert rel_l2(back[:, mask], phi[:, mask]) < 1e-12
    assert np.all(back[:, 0, 0] == 0) and np.all(v[:, 0, 0] == 0)


@pytest.mark.parametrize("xofft", [0.37, -2.3])
def test_getv_twisted_eigenmodes(xofft):
    """Line (i,k) set to exp(i w j), w = (arg shp + 2 pi m)/my with shp =
    exp(-i kx xofft) the top twist: getv(-1) multiplies it by
    R2(w)/B2(w) - k2 and getv(+1) by the reciprocal."""
    g = _grid()
    c = _cross(g, xofft)
    rng = np.random.default_rng(5)
    m = rng.integers(-g.my // 2, g.my // 2, size=(g.kxn, g.mz))
    th = np.angle(np.exp(-1j * g.kx * xofft))[:, None]
    w = (th + 2 * math.pi * m) / g.my                       # (kxn, mz)
    j = np.arange(g.my)[:, None, None]
    v = np.exp(1j * w[None] * j)
    k2 = g.alp2[:, None] + g.gam2[None, :]
    lam = (_sym(c.ops.rh2, w.ravel()) / _sym(c.ops.pr2, w.ravel())).reshape(w.shape) - k2
    phi = c.getv(v, -1)
    mask = np.ones((g.kxn, g.mz), bool)
    mask[0, 0] = False
    assert rel_l2(phi[:, mask], (lam[None] * v)[:, mask]) < 1e-12
    vv = c.getv(v, 1)
    assert rel_l2(vv[:, mask], (v / lam[None])[:, mask]) < 1e-12
    # the bottom twist in place of the top one would not reproduce the mode
    c.xofft, c.xoffb = -xofft, xofft
    assert rel_l2(c.getv(v, -1)[:, mask], (lam[None] * v)[:, mask]) > 1e-3


def test_getv_smooth_accuracy():
    """v = exp(i kappa y) with kappa = 2 pi / Ly (untwisted, xofft = 0):
    phi = -(kappa^2 + k2) v to the compact scheme's truncation error."""
    g = _grid(my=97)
    c = _cross(g, 0.0)
    kap = 2 * math.pi / g.Ly
    v = np.broadcast_to(np.exp(1j * kap * g.y)[:, None, None], (g.my, g.kxn, g.mz)).copy()
    k2 = g.alp2[:, None] + g.gam2[None, :]
    phi = c.getv(v, -1)
    ex = -(kap * kap + k2)[None] * v
    mask = np.ones((g.kxn, g.mz), bool)
    mask[0, 0] = False
    assert rel_l2(phi[:, mask], ex[:, mask]) < 1e-9


def test_energy_parseval():
    g = _grid(24, 17, 36)
    rng = np.random.default_rng(3)
    c = Cross(g)
    c.set_state(_field(g, rng), _field(g, rng), rng.normal(size=g.my), rng.normal(size=g.my))
    e = c.energy()
    for q, f in enumerate((c.vor, c.phi)):
        phys = fftxz.to_phys(f, g.mgalx, g.mgalz)
        assert abs(e[q] - np.sum(np.mean(phys ** 2, axis=(1, 2)))) < 1e-12 * e[q]
    assert e[2] == pytest.approx(np.dot(c.u00, c.u00), rel=1e-14)
    assert e[3] == pytest.approx(np.dot(c.w00, c.w00), rel=1e-14)
