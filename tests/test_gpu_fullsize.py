"""Checks at the measured size, 384x385x384 (Re_z = 8000, explicit), where the
numpy oracle is too slow to run a step: size-independent properties instead.

- Two independent implementations of the same operators must agree to
  rounding: the spectral y stage (twisted-Fourier multipliers, my = 385)
  against the banded twisted-cyclic L D L^H sweeps, and the wave x-z kernels
  (one-wave FFTs of 384 points) against the generic Stockham kernels.
  Bar: relative L2 <= 1e-12 per field after 2 RK3 steps.
- The energy the library computes on the device (snf_energy) equals the
  oracle's energy (operators.f90:870-924) of the downloaded state, to 1e-12.
- The mixed precision mode stays within 1e-5 of FP64 after 3 steps.
"""
import gc

import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

GRID = (384, 385, 384)
RE = 8000.0
PATH_TOL = 1e-12


@pytest.fixture(scope="module")
def ic():
    return box_ic(*GRID)


def _advance(ic, nstep, monkeypatch=None, env=None, **kw):
    import shear_nnf_amd as snf
    if env:
        for k, v in env.items():
            monkeypatch.setenv(k, v)
    box = snf.ShearBox(*GRID, RE, alp=ALP, gam=GAM, Ly=LY, explicit=1, **kw)
    box.set_state(*ic)
    st = box.cross(nstep)
    state = box.get_state()
    energy = box.energy()
    box.close()
    gc.collect()
    return st, state, energy


def test_fullsize_y_paths_agree(ic, monkeypatch):
    st_s, spec, _ = _advance(ic, 2, monkeypatch, {"SNF_Y_PATH": "spectral"})
    st_b, band, _ = _advance(ic, 2, monkeypatch, {"SNF_Y_PATH": "banded"})
    assert st_s.steps_done == st_b.steps_done == 2
    assert st_s.time == pytest.approx(st_b.time, rel=1e-14)
    for a, b in zip(spec, band):
        assert rel_l2(a, b) < PATH_TOL


def test_fullsize_xz_paths_agree(ic, monkeypatch):
    _, wave, _ = _advance(ic, 2, monkeypatch, {"SNF_XZ_PATH": "wave"})
    _, gen, _ = _advance(ic, 2, monkeypatch, {"SNF_XZ_PATH": "generic"})
    for a, b in zip(wave, gen):
        assert rel_l2(a, b) < PATH_TOL


def test_fullsize_device_energy(ic):
    from oracle.cross import Cross
    from oracle.grid import Grid
    _, (vor, phi, u00, w00), energy = _advance(ic, 1)
    c = Cross(Grid(*GRID, Re=RE, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5), explicit=1)
    c.set_state(vor, phi, u00, w00)
    ref = np.asarray(c.energy())
    assert np.all(ref > 0)
    np.testing.assert_allclose(np.asarray(energy), ref, rtol=1e-12)


def test_fullsize_mixed_close_to_fp64(ic):
    _, f64, _ = _advance(ic, 3, precision="fp64")
    _, fmx, _ = _advance(ic, 3, precision="mixed")
    for a, b in zip(fmx, f64):
        assert 0.0 < rel_l2(a, b) < 1e-5
