"""Why a run stopped (snf_run_status.stop_reason): the etime stop
(cross.f90:1699-1709) and the laminarisation stop (v'/Lz < 1e-5,
cross.f90:2106-2111), the first condition raised in the run."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic

pytestmark = pytest.mark.gpu

SNF_STOP, SNF_STOP_ETIME, SNF_STOP_LAMINAR = 1, 1, 4


def _box(**kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(48, 49, 48, 1000.0, alp=ALP, gam=GAM, Ly=LY, **kw)


def test_stop_reason_etime():
    box = _box(etime=0.02)
    box.set_state(*box_ic(48, 49, 48))
    st = box.cross(1000)
    box.close()
    assert st.stop == SNF_STOP and st.stop_reason == SNF_STOP_ETIME
    assert st.time == pytest.approx(0.02, abs=1e-14)


def test_stop_reason_laminar():
    vor, phi, u00, w00 = box_ic(48, 49, 48)
    box = _box(nhist=1)
    box.set_state(1e-9 * vor, 1e-9 * phi, 1e-9 * u00, 1e-9 * w00)
    st = box.cross(50)
    hist = np.array(box.history)
    box.close()
    assert st.stop == SNF_STOP and st.stop_reason == SNF_STOP_LAMINAR
    assert st.steps_done < 50
    assert hist[-1, 3] < 1e-5          # v' of the last record, Lz = 1


def test_no_stop_reason_when_running():
    box = _box()
    box.set_state(*box_ic(48, 49, 48))
    st = box.cross(3)
    box.close()
    assert st.stop == 0 and st.stop_reason == 0
