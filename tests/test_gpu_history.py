"""Run-control corner cases of the C ABI.

* The UNER records (the .cfbin rows, cross.f90:2067-2119) of one advance
  call longer than the device ring's initial 4096 records all arrive, in
  order, the same as when they are fetched in small pieces.
* An operator evaluation (snf_op_hvhg) between a return at a dump and the
  call that continues the run leaves the run as it was: the continued run
  is bitwise the uninterrupted one.
"""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic

pytestmark = pytest.mark.gpu


def _box(grid=(16, 17, 16), Re=300.0, **kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, **kw)


def test_history_longer_than_the_ring():
    # a small CFL and little viscosity keep the box from laminarising (the
    # v' < 1e-5 stop) over the 5000 steps
    n = 5000
    ic = box_ic(16, 17, 16)
    one = _box(Re=1e4, CFL=0.01, nhist=1)
    one.set_state(*ic)
    st = one.cross(n)
    assert (st.steps_done, st.nhist_records, st.stop) == (n, n, 0), st.stop_reason
    h1 = np.array(one.history)
    assert h1.shape == (n, 24)
    assert np.all(np.diff(h1[:, 0]) > 0)
    pieces = _box(Re=1e4, CFL=0.01, nhist=1)
    pieces.set_state(*ic)
    for _ in range(5):
        pieces.cross(1000)
    # each call is its own cross call, which re-forms v, dv/dy and
    # d(omega_y)/dy from the state at its start: rounding-level differences
    h2 = np.array(pieces.history)
    assert h2.shape == h1.shape
    np.testing.assert_allclose(h1, h2, rtol=1e-9, atol=1e-15)


def test_hvhg_between_dump_and_continue():
    grid = (32, 33, 32)
    ic = box_ic(*grid)
    rng = np.random.default_rng(3)

    def run(probe):
        b = _box(grid, 1000.0, nhist=2, dumptime=0.004, dtimag=1e10)
        b.set_state(*ic)
        st = b.cross(12)
        assert st.stop == 2 and st.steps_done < 12
        if probe:
            shape = (b.my, b.kxn, b.mz)
            f = [rng.normal(size=shape) + 1j * rng.normal(size=shape) for _ in range(3)]
            b.hvhg(*f)
        st2 = b.cross(12 - st.steps_done)
        return st.steps_done + st2.steps_done, b.get_state(), np.array(b.history)

    n0, s0, h0 = run(False)
    n1, s1, h1 = run(True)
    assert n0 == n1 == 12
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)
    assert np.array_equal(h0, h1)
