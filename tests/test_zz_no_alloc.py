"""snf_energy_noalloc (partial sums in the library's static device scratch,
no allocation) against snf_energy through the C ABI: bitwise the same sums,
on two calls in a row, the second on a smaller box, and again after an
advance has changed the state."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic

pytestmark = pytest.mark.gpu


def _box(grid):
    import shear_nnf_amd as snf
    b = snf.ShearBox(*grid, 1000.0, alp=ALP, gam=GAM, Ly=LY)
    b.set_state(*box_ic(*grid))
    return b


def test_energy_noalloc_equals_energy():
    big, small = _box((192, 97, 128)), _box((32, 33, 32))
    for b in (big, small):          # two calls in a row, the second smaller
        e0 = b.energy_noalloc()
        print("noalloc", b.mgalx, b.my, b.mgalz, e0, flush=True)
        assert np.all(e0 > 0)
        assert np.array_equal(e0, b.energy())
    small.cross(3)
    big.cross(2)
    for b in (big, small):
        assert np.array_equal(b.energy_noalloc(), b.energy())
