"""The shooting map (calc_dns_disk3d, gmres_shear.f90:2951-3120) and its
forward-difference Jacobian-vector products (matvec_dns_disk3d with
iopt_matvec=1, gmres_shear.f90:2666-2730) through the C ABI.

* Phi_T against the oracle's run of one cross call to t0 + T (<= 1e-11);
  a run long enough to span several advance chunks agrees with one
  uninterrupted advance call.
* JVP: a Kelvin mode is an exact solution and its amplitude enters
  linearly, so around the mode of amplitude A the derivative along the
  mode itself is the unit-amplitude closed form at T (operators.f90:1516-1562);
  the viscous step limit binds, so both shootings take the same steps.
* The batch equals separate shootings, direction by direction (bitwise).
"""
import math

import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu


def _box(grid, Re=1000.0, CFL=0.5, **kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, CFL=CFL, **kw)


def test_shoot_matches_oracle():
    from oracle.cross import Cross
    from oracle.grid import Grid
    grid, T = (32, 33, 32), 0.02
    box = _box(grid)
    ic = box_ic(*grid)
    y, st = box.shoot(box.pack_state(*ic), T, t0=0.3, xoffb0=-0.1, xofft0=0.1)
    assert st.time == pytest.approx(0.3 + T, abs=1e-14) and st.stop == 1
    c = Cross(Grid(*grid, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5))
    c.set_state(*ic, time=0.3, xoffb=-0.1, xofft=0.1)
    c.etime = 0.3 + T
    c.run(100000)
    for a, b in zip(box.unpack_state(y), (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < 1e-11
    assert st.xofft == pytest.approx(c.xofft, abs=1e-14)


def test_shoot_spans_chunks():
    """> 256 steps: the cross call runs in several advance chunks"""
    grid, T = (16, 33, 16), 2.0
    ic = box_ic(*grid)
    box = _box(grid, Re=300.0, CFL=0.125)
    y, st = box.shoot(box.pack_state(*ic), T)
    ref = _box(grid, Re=300.0, CFL=0.125, etime=T)
    ref.set_state(*ic)
    st2 = ref.cross(4000)
    assert st2.time == pytest.approx(T, abs=1e-14) and st2.steps_done > 256
    for a, b in zip(box.unpack_state(y), ref.get_state()):
        assert rel_l2(a, b) < 1e-12


def test_jvp_kelvin_mode():
    from oracle.grid import Grid
    from oracle.rdt import KelvinMode
    grid, Re, T = (16, 33, 16), 300.0, 0.5
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.125)
    box = _box(grid, Re=Re, CFL=0.125)
    base = KelvinMode(g, amp=0.1)
    unit = KelvinMode(g, amp=1.0)
    x0 = box.pack_state(*base.state())
    d = box.pack_state(*unit.state())
    jv = box.jvp_batch(x0, d, 1e-6, T)[0]
    vor, phi, u00, w00 = box.unpack_state(jv)
    e = unit.error(vor, phi, T)
    assert e[0] < 1e-6 and e[1] < 1e-6 and e[2] < 1e-5
    assert np.max(np.abs(u00)) < 1e-6 and np.max(np.abs(w00)) < 1e-6


def test_jvp_batch_equals_separate_shootings():
    grid, T = (32, 33, 32), 0.01
    box = _box(grid)
    x0 = box.pack_state(*box_ic(*grid))
    rng = np.random.default_rng(4)
    dirs = rng.normal(size=(2, x0.size)) * np.abs(x0).max()
    eps = np.array([1e-7, 3e-7])
    out = box.jvp_batch(x0, dirs, eps, T)
    y0, _ = box.shoot(x0, T)
    for q in range(2):
        yq, _ = box.shoot(x0 + eps[q] * dirs[q], T)
        assert np.array_equal(out[q], (yq - y0) / eps[q])
    # with y0 given, the same result
    out2 = box.jvp_batch(x0, dirs, eps, T, y0=y0)
    assert np.array_equal(out, out2)


def test_jvp_batch_groups_of_handles():
    """20 directions: the caller's handle and 15 copies shoot the first 16
    together, then 4 more; each product is bitwise that of its own shootings"""
    grid, T = (48, 49, 48), 0.005
    box = _box(grid)
    x0 = box.pack_state(*box_ic(*grid))
    rng = np.random.default_rng(9)
    nd = 20
    dirs = rng.normal(size=(nd, x0.size)) * np.abs(x0).max()
    eps = 1e-7 * (1.0 + np.arange(nd) / nd)
    y0, _ = box.shoot(x0, T)
    out = box.jvp_batch(x0, dirs, eps, T, y0=y0)
    for q in (0, 15, 16, 19):
        yq, _ = box.shoot(x0 + eps[q] * dirs[q], T)
        assert np.array_equal(out[q], (yq - y0) / eps[q])


@pytest.mark.parametrize("kw", [{"precision": "mixed"}, {"explicit": 0}])
def test_jvp_batch_mixed_and_implicit(kw):
    """the concurrent shootings of other handle kinds: bitwise those of
    separate shootings.  T spans several steps: a first step that already
    reaches t0 + T is followed by a zero-length one (see
    test_implicit_shoot_shorter_than_one_step)"""
    grid, T = (32, 33, 32), 0.05
    box = _box(grid, **kw)
    x0 = box.pack_state(*box_ic(*grid))
    dirs = np.random.default_rng(2).normal(size=(3, x0.size)) * np.abs(x0).max()
    eps = np.array([1e-6, 2e-6, 3e-6])
    y0, _ = box.shoot(x0, T)
    out = box.jvp_batch(x0, dirs, eps, T, y0=y0)
    for q in range(3):
        yq, _ = box.shoot(x0 + eps[q] * dirs[q], T)
        assert np.array_equal(out[q], (yq - y0) / eps[q])


def test_implicit_shoot_shorter_than_one_step():
    """The stop test is skipped before the second step (cross.f90:356,
    istep-1 /= 1), so when the first step already lands on t0 + T a second
    step of length zero runs; the implicit update divides by Deltat
    (cross.f90:849-865) and the state turns to NaN.  The library reproduces
    the reference here: the same entries are NaN, the rest agree."""
    from oracle.cross import Cross
    from oracle.grid import Grid
    grid, T = (32, 33, 32), 0.01
    ic = box_ic(*grid)
    box = _box(grid, explicit=0)
    y, st = box.shoot(box.pack_state(*ic), T)
    c = Cross(Grid(*grid, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5), explicit=0)
    c.set_state(*ic)
    c.etime = T
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        c.run(100000)
    for a, b in zip(box.unpack_state(y), (c.vor, c.phi, c.u00, c.w00)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        fin = ~np.isnan(b)
        np.testing.assert_allclose(a[fin], b[fin], rtol=0, atol=1e-12 * max(1.0, np.abs(b[fin]).max(initial=0.0)))
    assert np.isnan(c.vor).any()
    assert st.time == pytest.approx(T, abs=1e-14)
