"""Ensembles: snf_advance_batch advances independent boxes of different
grids, Reynolds numbers and modes with all their work in flight at once;
every box must end exactly where its own snf_advance takes it (bitwise:
the same kernels on the same data, only the interleaving differs)."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic

pytestmark = pytest.mark.gpu

# the my = 385 box comes first: its y kernels need more LDS than the later,
# smaller boxes, and the per-kernel LDS limit must stay at the largest
CASES = [((48, 385, 48), 1500.0, {}), ((32, 33, 32), 500.0, {}), ((48, 49, 48), 3000.0, {"explicit": 0}),
         ((64, 65, 64), 1200.0, {}), ((96, 97, 96), 2000.0, {}),
         ((128, 49, 128), 1000.0, {"precision": "mixed"}), ((48, 49, 48), 800.0, {"ydealias": 3}),
         ((32, 150, 32), 700.0, {}), ((64, 64, 64), 900.0, {"explicit": 0, "precision": "mixed"}),
         ((48, 513, 48), 1000.0, {"precision": "mixed"})]


def _boxes():
    import shear_nnf_amd as snf
    out = []
    for q, (grid, Re, kw) in enumerate(CASES):
        b = snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, **kw)
        b.set_state(*box_ic(*grid, seed=20260001 + q))
        out.append(b)
    return out


def test_batch_equals_separate_advances():
    import shear_nnf_amd as snf
    batch, single = _boxes(), _boxes()
    st = snf.advance_batch(batch, 6)
    for b, s, r in zip(batch, single, st):
        r1 = s.cross(6)
        assert r.steps_done == r1.steps_done == 6 and r.time == r1.time
        for x, y in zip(b.get_state(), s.get_state()):
            assert np.array_equal(x, y)
        assert np.array_equal(np.array(b.history), np.array(s.history))


def test_batch_rejects_a_repeated_handle():
    import shear_nnf_amd as snf
    b = _boxes()[:1]
    with pytest.raises(snf.SnfError):
        snf.advance_batch([b[0], b[0]], 1)


# several boxes per size class: each class advances from box tables (one
# launch per kernel for all its boxes, pairs of steps from one graph)
TABLE_CASES = [((32, 33, 32), 500.0, {}), ((32, 33, 32), 2900.0, {}), ((32, 33, 32), 1700.0, {}),
               ((48, 49, 48), 900.0, {"explicit": 0}), ((48, 49, 48), 2500.0, {"explicit": 0}),
               ((64, 65, 64), 1200.0, {"precision": "mixed"}),
               ((64, 65, 64), 2200.0, {"precision": "mixed"}),
               ((96, 97, 96), 3000.0, {}), ((96, 97, 96), 600.0, {}), ((128, 49, 128), 800.0, {})]


def _table_boxes(nhist=1):
    import shear_nnf_amd as snf
    out = []
    for q, (grid, Re, kw) in enumerate(TABLE_CASES):
        b = snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, nhist=nhist, **kw)
        b.set_state(*box_ic(*grid, seed=20260100 + q))
        out.append(b)
    return out


@pytest.mark.parametrize("nsteps", [(5,), (2, 3, 1)])
def test_box_tables_equal_separate_advances(nsteps):
    """odd and even step counts, and calls that continue one another (the
    twist parity of each box carries over from call to call)"""
    import shear_nnf_amd as snf
    batch, single = _table_boxes(), _table_boxes()
    for k in nsteps:
        st = snf.advance_batch(batch, k)
        assert all(r.steps_done == k for r in st)
    for b, s in zip(batch, single):
        for k in nsteps:
            s.cross(k)
        for x, y in zip(b.get_state(), s.get_state()):
            assert np.array_equal(x, y)
        assert np.array_equal(np.array(b.history), np.array(s.history))


def test_batch_fp32_y_transforms(monkeypatch):
    """SNF_Y32=1: mixed boxes whose explicit y substeps transform in FP32, a
    size class of two (box tables) and a box of its own, bitwise against
    their own advances"""
    import shear_nnf_amd as snf
    monkeypatch.setenv("SNF_Y32", "1")
    cases = [((48, 385, 48), 1500.0), ((64, 65, 64), 1200.0), ((64, 65, 64), 2200.0),
             ((48, 513, 48), 1000.0)]

    def boxes():
        out = []
        for q, (grid, Re) in enumerate(cases):
            b = snf.ShearBox(*grid, Re, alp=ALP, gam=GAM, Ly=LY, nhist=1, precision="mixed")
            b.set_state(*box_ic(*grid, seed=20260200 + q))
            out.append(b)
        return out

    batch, single = boxes(), boxes()
    st = snf.advance_batch(batch, 5)
    for b, s, r in zip(batch, single, st):
        r1 = s.cross(5)
        assert r.steps_done == r1.steps_done == 5 and r.time == r1.time
        for x, y in zip(b.get_state(), s.get_state()):
            assert np.array_equal(x, y)
        assert np.array_equal(np.array(b.history), np.array(s.history))
