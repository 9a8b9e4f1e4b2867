"""Pins the oracle's y-dealiasing variants 1 and 2 (iydealiasing,
cross.f90:459-475, dealiasing_y operators.f90:352-486).

* The real-line cut of U, W (rfty, zeroing from real index 2 (my/3) on) and
  the complex cut of the kx = 0 lines (cfty, zeroing (ny, my - ny)) are two
  independent index computations of the reference that must keep the same
  modes, |m| <= ny: checked mode by mode on single-mode lines.
* Both variants are projections (idempotent for a fixed offset) that keep
  the kx = 0 plane Hermitian; variant 1 leaves kx != 0 lines untouched.
* Runs: a Kelvin mode (kx != 0) is untouched by variant 1 and converges to
  its closed form; the kx = 0 lift-up mode and U, W profiles below the cut
  pass both variants; variant 2 keeps the offsets wrapped to one period.
The values of the shifted windows djda/djdb of variant 2 for kx != 0
(operators.f90:433-460) are pinned mode by mode in tests/test_oracle_pins.py.
"""
import math

import numpy as np
import pytest

from oracle.cross import Cross, set_conj
from oracle.grid import Grid
from oracle.rdt import KelvinMode, StreamwiseMode
from tests.boxes import ALP, GAM, LY

T_END = 0.5


def _grid(my, mg=16, CFL=0.5):
    return Grid(mg, my, mg, Re=300.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=CFL)


@pytest.mark.parametrize("my", [33, 49, 97, 385])
def test_real_and_complex_cuts_keep_the_same_modes(my):
    g = _grid(my)
    c = Cross(g, ydealias=1)
    j = np.arange(my)
    kept_c, kept_r = [], []
    for m in range(my):
        c.vor = g.zeros()
        c.phi = g.zeros()
        c.vor[:, 0, 1] = np.exp(2j * math.pi * m * j / my)
        c.u00 = np.cos(2 * math.pi * m * j / my)
        c.w00 = np.zeros(my)
        c._dealias_lines(0.0)
        if np.linalg.norm(c.vor[:, 0, 1]) > 0.5 * math.sqrt(my):
            kept_c.append(m)
        else:
            assert np.linalg.norm(c.vor[:, 0, 1]) < 1e-13 * my
        if np.linalg.norm(c.u00) > 0.25 * math.sqrt(my):
            kept_r.append(m)
    assert kept_c == kept_r
    ny = ((2 * (my // 3) - 1) - 1) // 2
    assert len(kept_c) == 2 * ny + 1 <= 2 * my // 3 + 1
    assert all(min(m, my - m) <= ny for m in kept_c)


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("xwkb", [-0.3, -1.7, -3.9, 0.45])
def test_cut_is_a_hermitian_projection(variant, xwkb):
    g = Grid(32, 33, 32, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0)
    c = Cross(g, ydealias=variant)
    rng = np.random.default_rng(9)

    def field():
        f = rng.normal(size=(g.my, g.kxn, g.mz)) + 1j * rng.normal(size=(g.my, g.kxn, g.mz))
        f[:, 0, 0] = 0.0
        return set_conj(f, g.nz)

    vor, phi = field(), field()
    u00, w00 = rng.normal(size=g.my), rng.normal(size=g.my)
    c.set_state(vor, phi, u00, w00)
    c._dealias_y(xwkb)
    once = [a.copy() for a in (c.vor, c.phi, c.u00, c.w00)]
    c._dealias_y(xwkb)
    for a, b in zip((c.vor, c.phi, c.u00, c.w00), once):
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)
    for f in once[:2]:
        h = set_conj(f.copy(), g.nz)
        assert np.linalg.norm(h - f) <= 1e-13 * np.linalg.norm(f)
    # something was removed, and variant 1 removed it from kx = 0 only
    assert np.linalg.norm(once[0]) < 0.99 * np.linalg.norm(vor)
    assert np.linalg.norm(once[2]) < 0.99 * np.linalg.norm(u00)
    if variant == 1:
        assert np.linalg.norm(once[0][:, 1:, :] - vor[:, 1:, :]) <= 1e-14 * np.linalg.norm(vor)


def test_kelvin_mode_passes_variant_1():
    errs = []
    for CFL in (0.25, 0.125):
        g = _grid(33, CFL=CFL)
        mode = KelvinMode(g, amp=0.1)
        c = Cross(g, explicit=1, ydealias=1)
        c.set_state(*mode.state())
        c.etime = T_END
        c.run(100000)
        errs.append(mode.error(c.vor, c.phi, c.time))
    assert 6.0 < errs[0][0] / errs[1][0] < 10.5
    assert errs[1][0] < 1e-8 and errs[1][2] < 1e-7


def _run(g, mode, variant, explicit, x0, nfix=None):
    c = Cross(g, explicit=explicit, ydealias=variant)
    vor, phi, u00, w00 = mode.state()
    c.set_state(vor, phi, u00, w00, xoffb=-x0, xofft=x0)
    c.etime = T_END
    if nfix:
        c.dumptime = c.dtimag = T_END / nfix
    c.run(100000)
    return c, mode.error(c.vor, c.phi, c.u00, c.w00, c.time)


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("explicit", [1, 0])
def test_streamwise_mode_and_profiles_pass(variant, explicit):
    """a kx = 0 mode, and U, W below the cut, are untouched; with variant 2
    the offsets stay wrapped to one period of Lx (cross.f90:348-352, 733-736).
    Initial offsets of 3 Lx keep the twist 1, so the closed forms hold."""
    g = _grid(33, CFL=0.125)
    x0 = 3.0 * g.Lx
    c, e = _run(g, StreamwiseMode(g, amp=0.1), variant, explicit, x0)
    assert e[0] < (1e-9 if explicit else 1e-7)
    if variant == 2 and explicit:
        assert abs(c.xoffb) < g.Lx and c.xofft == -c.xoffb
        assert c.xofft == pytest.approx(g.s * g.Ly * T_END, abs=1e-12)
    elif variant == 2:
        # the implicit stages do not wrap their substep offsets (cross.f90:871-872)
        assert abs(c.xoffb) < g.Lx + g.s * g.Ly * T_END and c.xofft == pytest.approx(-c.xoffb)
    else:
        assert c.xofft == pytest.approx(x0 + g.s * g.Ly * T_END, abs=1e-12)
    _, e = _run(g, StreamwiseMode(g, amp=1.5e-5, A=1e-3), variant, explicit, x0,
                nfix=None if explicit else 50)
    assert e[3] < 2e-8 and e[4] < 2e-8
