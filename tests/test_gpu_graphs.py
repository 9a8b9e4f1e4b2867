"""Steps replayed from captured HIP graphs against steps launched one by one
(SNF_GRAPHS=0): bitwise the same state and UNER records after every call of
a sequence of calls with host work and synchronisations in between, for the
spectral and banded y stages, the wave and Stockham x-z kernels, both
precisions, and the batched advance."""
import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic

pytestmark = pytest.mark.gpu


def _box(grid, **kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(*grid, 1000.0, alp=ALP, gam=GAM, Ly=LY, **kw)


def _calls(grid, monkeypatch, graphs, calls, **kw):
    monkeypatch.setenv("SNF_GRAPHS", "1" if graphs else "0")
    box = _box(grid, **kw)
    box.set_state(*box_ic(grid[0], grid[1], grid[2]))
    out = []
    for n in calls:
        st = box.cross(n)
        out.append((box.get_state(), np.array(box.history), st.steps_done, st.time))
    box.close()
    return out


@pytest.mark.parametrize("grid,kw", [((48, 49, 48), {}), ((32, 40, 32), {}),
                                     ((128, 33, 128), {}), ((48, 49, 48), {"explicit": 0}),
                                     ((128, 49, 128), {"precision": "mixed"}),
                                     ((48, 49, 48), {"nhist": 3})])
def test_graph_steps_equal_eager(grid, kw, monkeypatch):
    calls = (4, 1, 7, 2)       # pairs from graphs, odd remainders one by one
    eager = _calls(grid, monkeypatch, False, calls, **kw)
    graph = _calls(grid, monkeypatch, True, calls, **kw)
    for (se, he, ne, te), (sg, hg, ng, tg) in zip(eager, graph):
        assert ne == ng and te == tg
        for a, b in zip(se, sg):
            assert np.array_equal(a, b)
        assert np.array_equal(he, hg)


def test_graph_batch_equals_eager(monkeypatch):
    import shear_nnf_amd as snf
    grids = [(32, 33, 32), (48, 49, 48), (32, 32, 32)]
    res = {}
    for graphs in (False, True):
        monkeypatch.setenv("SNF_GRAPHS", "1" if graphs else "0")
        boxes = [_box(g) for g in grids]
        for b, g in zip(boxes, grids):
            b.set_state(*box_ic(*g))
        snf.advance_batch(boxes, 6)
        snf.advance_batch(boxes, 3)
        res[graphs] = [b.get_state() for b in boxes]
        for b in boxes:
            b.close()
    for se, sg in zip(res[False], res[True]):
        for a, b in zip(se, sg):
            assert np.array_equal(a, b)
