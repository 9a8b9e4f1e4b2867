"""GPU parity of getv (operators.f90:735-793) and energy (operators.f90:870-924)
through the C ABI against the oracle.  Bars: relative L2 <= 1e-12 for getv
(a compact solve per line, cond up to ~1e4 at the finest k2), relative
1e-13 for the energy sums."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(mgalx, my, mgalz, **kw):
    import shear_nnf_amd as snf
    from oracle.grid import Grid
    from oracle.cross import Cross
    box = snf.ShearBox(mgalx, my, mgalz, 1000.0, alp=ALP, gam=GAM, Ly=LY, **kw)
    g = Grid(mgalx, my, mgalz, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5)
    return box, Cross(g)


def _field(c, rng):
    from oracle.cross import set_conj
    g = c.g
    f = rng.normal(size=(g.my, g.kxn, g.mz)) + 1j * rng.normal(size=(g.my, g.kxn, g.mz))
    f[:, 0, 0] = f[:, 0, 0].real
    return set_conj(f, g.nz)


@pytest.mark.parametrize("grid,xofft", [((48, 49, 48), 0.0), ((48, 49, 48), 0.61),
                                        ((96, 97, 64), -1.7), ((32, 193, 32), 2.9)])
def test_getv(grid, xofft):
    box, c = _pair(*grid)
    rng = np.random.default_rng(40 + grid[1])
    vor, phi = _field(c, rng), _field(c, rng)
    u00, w00 = rng.normal(size=c.g.my), rng.normal(size=c.g.my)
    box.set_state(vor, phi, u00, w00, xoffb=-xofft, xofft=xofft)
    c.set_state(vor, phi, u00, w00, xoffb=-xofft, xofft=xofft)
    v = box.getv(phi, 1)
    assert rel_l2(v, c.getv(phi, 1)) < TOL
    back = box.getv(v, -1)
    assert rel_l2(back, c.getv(v, -1)) < TOL
    mask = np.ones((c.g.kxn, c.g.mz), bool)
    mask[0, 0] = False
    assert rel_l2(back[:, mask], phi[:, mask]) < TOL
    assert np.all(v[:, 0, 0] == 0) and np.all(back[:, 0, 0] == 0)


@pytest.mark.parametrize("grid", [(48, 49, 48), (96, 97, 64)])
def test_energy(grid):
    box, c = _pair(*grid)
    vor, phi, u00, w00 = box_ic(*grid)
    box.set_state(vor, phi, u00, w00)
    c.set_state(vor, phi, u00, w00)
    got = box.energy()
    ref = c.energy()
    assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref))
    # after a few steps: the energy of the device state equals the oracle's
    # energy of the downloaded state
    box.cross(3)
    vor2, phi2, u2, w2 = box.get_state()
    c.set_state(vor2, phi2, u2, w2)
    ref2 = c.energy()
    got2 = box.energy()
    assert np.all(np.abs(got2 - ref2) <= 1e-13 * np.abs(ref2))
