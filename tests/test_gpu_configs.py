"""GPU parity at the benchmark configurations of the path.

* 192x193x192, Re_z = 4000, explicit and implicit, 10 RK3 steps (the drop-in
  parity workload): against the committed oracle fixtures
  ``tests/golden/config2_192x193x192{,_implicit}.npz`` (made by
  ``tests/golden/make_config2.py``): the relative L2 error per field, estimated
  through 64 seeded complex Gaussian projections, <= 1e-12; U, W exactly as
  fields <= 1e-12; every UNER record to rtol 1e-11.  my = 193 is prime: the
  spectral y stage transforms its lines by Rader's algorithm.
* 32x193x32, explicit and implicit, 10 steps against the live oracle, on
  the Rader y stage and on the banded twisted-cyclic sweeps.
* 48x513x48, explicit and implicit, 5 steps against the live oracle, on
  the spectral y stage and on the banded sweeps (the 512^3 grid's line
  length).
* 48x49x48, Re_z = 1000, 100 steps against the live oracle, explicit and
  implicit.
* 512x513x512, Re_z = 8000 (too large for the oracle): the wave x-z kernels
  against the Stockham ones and the spectral y stage against the banded
  sweeps over 2 steps (<= 1e-12), the device energy
  against the oracle's energy of the downloaded state (<= 1e-12), and the
  mixed precision mode against FP64 over 3 steps (<= 1e-5).
* 48x49x48 Newton workload, mixed precision: 16 JVP directions of one batch,
  each bitwise the product of its own two shootings, and close to the FP64
  products (bound stated in the test).
"""
import gc
import math

import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12


def _box(grid, Re, **kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, **kw)


def _oracle(grid, Re, explicit=1, **kw):
    from oracle.cross import Cross
    from oracle.grid import Grid
    g = Grid(*grid, Re=Re, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5)
    return Cross(g, explicit=explicit, **kw)


@pytest.mark.parametrize("explicit", [1, 0])
def test_config2_192x193x192_10_steps(explicit):
    from tests.golden.make_config2 import GRID, NSTEP, RE, SEED, fixture_path, sketch
    ref = np.load(fixture_path(explicit))
    print(f"config 2: {GRID} Re_z={RE} explicit={explicit} {NSTEP} steps vs fixture", flush=True)
    box = _box(GRID, RE, explicit=explicit, nhist=1)
    box.set_state(*box_ic(*GRID, seed=SEED))
    st = box.cross(NSTEP)
    assert st.steps_done == NSTEP
    vor, phi, u00, w00 = box.get_state()
    for nm, f in (("vor", vor), ("phi", phi)):
        err = np.linalg.norm(sketch(f) - ref[nm + "_sketch"]) / np.linalg.norm(ref[nm + "_sketch"])
        print(f"  {nm}: sketched rel L2 {err:.3e}; norm {np.linalg.norm(f):.15e} vs "
              f"{float(ref[nm + '_norm']):.15e}")
        assert err < STEP_TOL
        assert abs(np.linalg.norm(f) - ref[nm + "_norm"]) <= 1e-13 * ref[nm + "_norm"]
    assert rel_l2(u00, ref["u00"]) < STEP_TOL and rel_l2(w00, ref["w00"]) < STEP_TOL
    assert st.time == pytest.approx(float(ref["time"]), rel=1e-13)
    assert st.xofft == pytest.approx(float(ref["xofft"]), rel=1e-13)
    np.testing.assert_allclose(np.array(box.history), ref["history"], rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("path", ["spectral", "banded"])
@pytest.mark.parametrize("explicit", [1, 0])
def test_y_my193_10_steps(explicit, path, monkeypatch):
    grid, Re = (32, 193, 32), 4000.0
    ic = box_ic(*grid)
    monkeypatch.setenv("SNF_Y_PATH", path)
    box = _box(grid, Re, explicit=explicit, nhist=1)
    box.set_state(*ic)
    st = box.cross(10)
    c = _oracle(grid, Re, explicit=explicit, nhist=1)
    c.set_state(*ic)
    c.run(10)
    assert st.steps_done == 10
    for a, b in zip(box.get_state(), (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("path", ["spectral", "banded"])
@pytest.mark.parametrize("explicit", [1, 0])
def test_y_my513_5_steps(explicit, path, monkeypatch):
    """my = 513 (the 512x513x512 grid's line length, 3^3 * 19) on a 48 x 48
    x-z plane, 5 steps against the oracle (about 17 s of oracle time per
    case): the spectral y stage (radix 9, 3 and a radix-19 pass spread over
    the lanes) and the banded twisted-cyclic sweeps."""
    grid, Re = (48, 513, 48), 8000.0
    ic = box_ic(*grid)
    monkeypatch.setenv("SNF_Y_PATH", path)
    box = _box(grid, Re, explicit=explicit, nhist=1)
    box.set_state(*ic)
    st = box.cross(5)
    c = _oracle(grid, Re, explicit=explicit, nhist=1)
    c.set_state(*ic)
    c.run(5)
    assert st.steps_done == 5
    for a, b in zip(box.get_state(), (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("explicit", [1, 0])
def test_config1_48x49x48_100_steps(explicit):
    grid, Re = (48, 49, 48), 1000.0
    ic = box_ic(*grid)
    box = _box(grid, Re, explicit=explicit, nhist=1)
    box.set_state(*ic)
    st = box.cross(100)
    c = _oracle(grid, Re, explicit=explicit, nhist=1)
    c.set_state(*ic)
    c.run(100)
    assert st.steps_done == 100
    errs = [rel_l2(a, b) for a, b in zip(box.get_state(), (c.vor, c.phi, c.u00, c.w00))]
    print(f"config 1, explicit={explicit}, 100 steps: rel L2 per field", errs)
    assert max(errs) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


# ---------------------------------------------------------------- 512^3
GRID512 = (512, 513, 512)


@pytest.fixture(scope="module")
def ic512():
    return box_ic(*GRID512)


def _advance512(ic, nstep, monkeypatch=None, env=None, **kw):
    if env:
        for k, v in env.items():
            monkeypatch.setenv(k, v)
    print(f"512^3: {nstep} steps {env or ''} {kw}", flush=True)
    box = _box(GRID512, 8000.0, explicit=1, **kw)
    box.set_state(*ic)
    st = box.cross(nstep)
    state = box.get_state()
    energy = box.energy()
    box.close()
    gc.collect()
    return st, state, energy


def test_512_xz_paths_agree(ic512, monkeypatch):
    st_w, wave, _ = _advance512(ic512, 2, monkeypatch, {"SNF_XZ_PATH": "wave"})
    st_g, gen, _ = _advance512(ic512, 2, monkeypatch, {"SNF_XZ_PATH": "generic"})
    assert st_w.steps_done == st_g.steps_done == 2
    for a, b in zip(wave, gen):
        assert rel_l2(a, b) < STEP_TOL


def test_512_y_paths_agree(ic512, monkeypatch):
    st_s, spec, _ = _advance512(ic512, 2, monkeypatch, {"SNF_Y_PATH": "spectral"})
    st_b, band, _ = _advance512(ic512, 2, monkeypatch, {"SNF_Y_PATH": "banded"})
    assert st_s.steps_done == st_b.steps_done == 2
    for a, b in zip(spec, band):
        assert rel_l2(a, b) < STEP_TOL


def test_512_device_energy(ic512):
    c = _oracle(GRID512, 8000.0)
    _, (vor, phi, u00, w00), energy = _advance512(ic512, 1)
    c.set_state(vor, phi, u00, w00)
    ref = np.asarray(c.energy())
    assert np.all(ref > 0)
    np.testing.assert_allclose(np.asarray(energy), ref, rtol=1e-12)


def test_512_mixed_close_to_fp64(ic512):
    _, f64, _ = _advance512(ic512, 3, precision="fp64")
    _, fmx, _ = _advance512(ic512, 3, precision="mixed")
    for a, b in zip(fmx, f64):
        assert 0.0 < rel_l2(a, b) < 1e-5


# ---------------------------------------------------------------- Newton
def test_newton_workload_mixed_jvp_16_directions():
    """48x49x48, Re_z = 1000: 16 directions in one batch on a mixed precision
    handle (the FP32 Jacobian-vector shootings).  With eps = 1e-7 |x| / |d|
    each product is bitwise that of its own two mixed shootings.  Accuracy:
    a forward difference of two evaluations with FP32 x-z stages carries
    their FP32 rounding divided by eps, so at eps = 1e-7 relative the mixed
    products are noise (measured 0.3 relative to FP64); at eps = 3e-4
    relative (about the square root of the FP32 unit roundoff, where the
    rounding and the curvature terms balance) they agree with the FP64
    products at eps = 1e-7 to the bound below."""
    grid, Re, T = (48, 49, 48), 1000.0, 0.01
    mx = _box(grid, Re, precision="mixed")
    f64 = _box(grid, Re)
    x0 = mx.pack_state(*box_ic(*grid))
    rng = np.random.default_rng(16)
    nd = 16
    dirs = rng.normal(size=(nd, x0.size))
    eps = 1e-7 * np.linalg.norm(x0) / np.linalg.norm(dirs, axis=1)
    y0, _ = mx.shoot(x0, T)
    out = mx.jvp_batch(x0, dirs, eps, T, y0=y0)
    for q in (0, 7, 15):
        yq, _ = mx.shoot(x0 + eps[q] * dirs[q], T)
        assert np.array_equal(out[q], (yq - y0) / eps[q])
    ref = f64.jvp_batch(x0, dirs[:4], eps[:4], T)
    big = mx.jvp_batch(x0, dirs[:4], 3e3 * eps[:4], T, y0=y0)
    ref_big = f64.jvp_batch(x0, dirs[:4], 3e3 * eps[:4], T)
    for q in range(4):
        e_noise = rel_l2(out[q], ref[q])
        e_mixed = rel_l2(big[q], ref[q])
        e_fd = rel_l2(ref_big[q], ref[q])
        e_same = rel_l2(big[q], ref_big[q])
        print(f"direction {q}: against FP64 at eps 1e-7: mixed eps 1e-7 {e_noise:.3e}, "
              f"mixed eps 3e-4 {e_mixed:.3e}, FP64 eps 3e-4 {e_fd:.3e}; mixed against FP64 "
              f"both at eps 3e-4 {e_same:.3e}")
        assert e_fd < 1e-2
        assert e_mixed < 2e-2
        assert e_same < 1e-3
