"""The whole GPU time advance against the Kelvin-mode closed form (rdtfunc,
operators.f90:1516-1562): the same runs as tests/test_oracle_rdt.py, through
the C ABI, must converge to the closed form at the scheme's order and agree
with the oracle's run (relative L2 <= 1e-11 after ~90 steps)."""
import math

import numpy as np
import pytest

from tests.boxes import rel_l2
from tests.test_oracle_rdt import T_END, check_kelvin_history, streamwise_run

pytestmark = pytest.mark.gpu

STOP = 1


def gpu_kelvin_run(CFL, explicit, grid=(16, 33, 16), Re=300.0):
    import shear_nnf_amd as snf
    from oracle.grid import Grid
    from oracle.rdt import KelvinMode
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=CFL)
    mode = KelvinMode(g, amp=0.1)
    box = snf.ShearBox(*grid, Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, CFL=CFL,
                       explicit=explicit, etime=T_END)
    box.set_state(*mode.state())
    st = box.cross(400)     # one cross call runs to etime
    assert st.stop == STOP and st.time == pytest.approx(T_END, abs=1e-14)
    vor, phi, u00, w00 = box.get_state()
    stats = box.statistics()[0].T
    spec = box.spectra()[0]
    history = list(box.history)
    box.close()
    return mode.error(vor, phi, st.time), (vor, phi, u00, w00), (mode, history, stats, spec)


@pytest.mark.parametrize("explicit,order", [(1, 3), (0, 2)])
def test_gpu_kelvin_mode_converges(explicit, order):
    from oracle.cross import Cross
    from oracle.grid import Grid
    from oracle.rdt import KelvinMode
    e1, _, _ = gpu_kelvin_run(0.25, explicit)
    e2, fields, hist = gpu_kelvin_run(0.125, explicit)
    ratio = e1[0] / e2[0]
    assert 0.75 * 2 ** order < ratio < 1.3 * 2 ** order
    assert e2[0] < (1e-8 if explicit else 1e-6)
    assert e2[2] < 1e-7
    # the oracle's run of the same case
    g = Grid(16, 33, 16, Re=300.0, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.125)
    c = Cross(g, explicit=explicit)
    c.set_state(*KelvinMode(g, amp=0.1).state())
    c.etime = T_END
    c.run(100000)
    assert rel_l2(fields[0], c.vor) < 1e-11
    assert rel_l2(fields[1], c.phi) < 1e-11
    # U, W stay at rounding level (the mode's u x omega has no x-z mean)
    for a, b in zip(fields[2:], (c.u00, c.w00)):
        assert np.linalg.norm(a - b) < 1e-13
    if explicit:
        # the device UNER records and statistics against the closed form
        check_kelvin_history(*hist)


def gpu_streamwise_run(CFL, explicit, amp=0.1, A=0.0, nfix=None, grid=(16, 33, 16), Re=300.0):
    import shear_nnf_amd as snf
    from oracle.grid import Grid
    from oracle.rdt import StreamwiseMode
    g = Grid(*grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=CFL)
    mode = StreamwiseMode(g, amp=amp, A=A)
    kw = dict(dumptime=T_END / nfix, dtimag=T_END / nfix) if nfix else {}
    box = snf.ShearBox(*grid, Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, CFL=CFL,
                       explicit=explicit, etime=T_END, **kw)
    box.set_state(*mode.state())
    # every dump returns from the advance; the next call resumes the run
    for _ in range(400):
        st = box.cross(400)
        if st.stop == STOP:
            break
    assert st.stop == STOP and st.time == pytest.approx(T_END, abs=1e-14)
    vor, phi, u00, w00 = box.get_state()
    box.close()
    return mode.error(vor, phi, u00, w00, st.time), (vor, phi, u00, w00)


@pytest.mark.parametrize("explicit,order", [(1, 3), (0, 2)])
def test_gpu_streamwise_mode_converges(explicit, order):
    e1, _ = gpu_streamwise_run(0.25, explicit)
    e2, f2 = gpu_streamwise_run(0.125, explicit)
    assert e1[0] / e2[0] > 0.75 * 2 ** order
    assert e2[0] < (1e-9 if explicit else 1e-7)
    assert e2[1] < (1e-9 if explicit else 1e-7)
    assert e2[2] == pytest.approx(e2[0], rel=1e-6)
    _, ref = streamwise_run(0.125, explicit)
    assert rel_l2(f2[0], ref[0]) < 1e-11 and rel_l2(f2[1], ref[1]) < 1e-11


@pytest.mark.parametrize("explicit,nfix,tol", [(1, None, 1e-8), (0, 50, 2e-8), (0, 49, 2e-8)])
def test_gpu_mean_profiles_diffuse(explicit, nfix, tol):
    """with nfix, 2 steps per advance call between dumps, resumed each time;
    the run must equal the oracle's uninterrupted one"""
    e, f = gpu_streamwise_run(0.25, explicit, amp=1.5e-5, A=1e-3, nfix=nfix)
    assert e[3] < tol and e[4] < tol
    _, ref = streamwise_run(0.25, explicit, amp=1.5e-5, A=1e-3, nfix=nfix)
    for a, b in zip(f, ref):
        assert rel_l2(a, b) < 1e-11
