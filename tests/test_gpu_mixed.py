"""The mixed precision mode (SNF_PREC_MIXED: FP32 x-z transforms, hybrid
planes and u x omega products; FP64 state, KMM reconstruction, y operators,
RK update and statistics sums) against the FP64 path and the oracle.

Bars: the FP32 nonlinear stage perturbs each substep's right-hand side at the
FP32 rounding level (~1e-7 relative), so fields after a few steps agree with
FP64 to <= 1e-5 relative L2 and the UNER statistics (energy, Reynolds stress,
dissipation) to <= 1e-5 relative, record by record.

SNF_Y32=1 (FP32 line transforms in the explicit y substeps as well,
test_mixed_fp32_y_transforms): the same field bar, energy and dissipation
record by record, and the Reynolds stress <uv>, which changes sign (a record
near zero has no relative accuracy), to <= 1e-5 of its largest magnitude over
the run."""
import math

import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-5


def _run(grid, precision, nstep, Re=1000.0, explicit=1):
    import shear_nnf_amd as snf
    box = snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, explicit=explicit, precision=precision)
    box.set_state(*box_ic(*grid))
    st = box.cross(nstep)
    out = box.get_state(), np.array(box.history), st
    box.close()
    return out


def _stats(hist):
    """energy 0.5 (u'^2 + v'^2 + w'^2), Reynolds stress <uv>, dissipation sum
    of each UNER record (time, Deltat, uner(10), ...)"""
    u = hist[:, 2:5]
    return {"energy": 0.5 * np.sum(u ** 2, axis=1), "reynolds_stress": hist[:, 5],
            "dissipation": hist[:, 11]}


def _check_stats(s64, smx, pointwise=False):
    for k in s64:
        d = np.abs(smx[k] - s64[k])
        if k == "reynolds_stress" and not pointwise:
            assert np.max(d) / np.max(np.abs(s64[k])) < MIX_TOL, k
        else:
            assert np.max(d / np.abs(s64[k])) < MIX_TOL, k


@pytest.mark.parametrize("grid,explicit", [((128, 49, 128), 1), ((192, 33, 128), 1),
                                           ((128, 49, 256), 0), ((512, 33, 512), 1),
                                           ((128, 513, 128), 1), ((128, 513, 128), 0)])
def test_mixed_agrees_with_fp64(grid, explicit):
    f64, h64, st64 = _run(grid, "fp64", 10, explicit=explicit)
    fmx, hmx, stmx = _run(grid, "mixed", 10, explicit=explicit)
    assert stmx.steps_done == st64.steps_done == 10
    for a, b in zip(fmx, f64):
        e = rel_l2(a, b)
        assert 0.0 < e < MIX_TOL        # FP32 rounding is there, and small
    _check_stats(_stats(h64), _stats(hmx), pointwise=True)
    assert abs(stmx.time - st64.time) < 1e-6 * st64.time


@pytest.mark.parametrize("grid", [(128, 49, 128), (192, 33, 128), (128, 513, 128),
                                  (64, 65, 64), (32, 33, 32)])
def test_mixed_fp32_y_transforms(grid, monkeypatch):
    """SNF_Y32=1: the explicit y substeps with FP32 line transforms (radix
    plans 7*7, 11*3, 9*3*19, 13*5; FP64 state, update and mean-flow lines)"""
    f64, h64, _ = _run(grid, "fp64", 10)
    monkeypatch.setenv("SNF_Y32", "1")
    fmx, hmx, st = _run(grid, "mixed", 10)
    assert st.steps_done == 10
    for a, b in zip(fmx, f64):
        assert 0.0 < rel_l2(a, b) < MIX_TOL
    _check_stats(_stats(h64), _stats(hmx))


def test_mixed_agrees_with_oracle():
    from oracle.cross import Cross
    from oracle.grid import Grid
    grid = (128, 33, 128)
    (vor, phi, u00, w00), _, _ = _run(grid, "mixed", 3)
    c = Cross(Grid(*grid, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5))
    c.set_state(*box_ic(*grid))
    c.run(3)
    for a, b in zip((vor, phi, u00, w00), (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < MIX_TOL


def test_mixed_kelvin_mode():
    """the closed-form Kelvin mode through the mixed path: the FP32 stage
    leaves the RK3 error of ~1e-8 at CFL 0.125 well under 1e-6"""
    import shear_nnf_amd as snf
    from oracle.grid import Grid
    from oracle.rdt import KelvinMode
    grid, Re = (128, 33, 128), 300.0
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.125)
    mode = KelvinMode(g, amp=0.1)
    box = snf.ShearBox(*grid, Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, CFL=0.125, etime=0.5,
                       precision="mixed")
    box.set_state(*mode.state())
    st = box.cross(2000)
    vor, phi, _, _ = box.get_state()
    box.close()
    assert st.time == pytest.approx(0.5, abs=1e-14)
    e = mode.error(vor, phi, st.time)
    assert e[0] < 1e-6 and e[1] < 1e-6 and e[2] < 1e-6


# sizes without a wave plan take the Stockham x-z kernels, FP32 there too:
# the ensemble sizes n^3, n x (n+1) x n (spectral and banded y), explicit and
# implicit, and a Newton-search box
@pytest.mark.parametrize("grid,explicit", [((48, 49, 48), 1), ((32, 33, 32), 1),
                                           ((64, 65, 64), 1), ((96, 96, 96), 0)])
def test_mixed_stockham_sizes_agree_with_fp64(grid, explicit):
    f64, h64, st64 = _run(grid, "fp64", 10, explicit=explicit)
    fmx, hmx, stmx = _run(grid, "mixed", 10, explicit=explicit)
    assert stmx.steps_done == st64.steps_done == 10
    for a, b in zip(fmx, f64):
        assert 0.0 < rel_l2(a, b) < MIX_TOL
    _check_stats(_stats(h64), _stats(hmx), pointwise=True)


def test_mixed_stockham_agrees_with_oracle():
    from oracle.cross import Cross
    from oracle.grid import Grid
    grid = (48, 49, 48)
    (vor, phi, u00, w00), _, _ = _run(grid, "mixed", 3)
    c = Cross(Grid(*grid, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5))
    c.set_state(*box_ic(*grid))
    c.run(3)
    for a, b in zip((vor, phi, u00, w00), (c.vor, c.phi, c.u00, c.w00)):
        assert rel_l2(a, b) < MIX_TOL
