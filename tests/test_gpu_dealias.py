"""GPU parity of the y-dealiasing variants 1 and 2 (iydealiasing, cross.f90:
459-475, dealiasing_y operators.f90:352-486) against the oracle: 3 RK3 steps
from a state with content at every y-mode (so the cuts remove something),
initial offsets beyond one period (so variant 2 wraps them), explicit and
implicit, odd and even my.  Bar: relative L2 <= 1e-12 per field, UNER
records (which carry the offsets) to 1e-11."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12


def _state(grid, seed=77):
    from oracle.cross import set_conj
    vor, phi, u00, w00 = box_ic(*grid)
    rng = np.random.default_rng(seed)

    def noise(f):
        n = rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape)
        n[:, 0, 0] = 0.0
        return set_conj(f + 0.05 * np.sqrt(np.mean(np.abs(f) ** 2)) * n, (f.shape[2] - 1) // 2)

    return (noise(vor), noise(phi), u00 + 0.01 * rng.normal(size=u00.shape),
            w00 + 0.01 * rng.normal(size=w00.shape))


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("grid,explicit", [((32, 33, 32), 1), ((48, 49, 48), 0), ((32, 40, 32), 1),
                                           ((24, 385, 24), 1), ((24, 513, 24), 0)])
def test_cross_ydealias(variant, grid, explicit):
    import shear_nnf_amd as snf
    from oracle.cross import Cross
    from oracle.grid import Grid
    vor, phi, u00, w00 = _state(grid)
    Lx = 2 * np.pi / ALP
    x0 = 2.6 * Lx + 0.3
    box = snf.ShearBox(*grid, 1000.0, alp=ALP, gam=GAM, Ly=LY, explicit=explicit, ydealias=variant)
    box.set_state(vor, phi, u00, w00, xoffb=-x0, xofft=x0)
    st = box.cross(3)
    got = box.get_state()
    g = Grid(*grid, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5)
    c = Cross(g, explicit=explicit, ydealias=variant)
    c.set_state(vor, phi, u00, w00, xoffb=-x0, xofft=x0)
    c.run(3)
    assert st.steps_done == 3
    for a, b in zip(got, (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < STEP_TOL
    assert st.xoffb == pytest.approx(c.xoffb, abs=1e-13)
    assert st.xofft == pytest.approx(c.xofft, abs=1e-13)
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)
    # the cut removed energy relative to the plain run of the same state
    if variant == 1:
        plain = Cross(g, explicit=explicit)
        plain.set_state(vor, phi, u00, w00, xoffb=-x0, xofft=x0)
        plain.run(3)
        assert np.linalg.norm(c.vor[:, 0, :]) < np.linalg.norm(plain.vor[:, 0, :])
