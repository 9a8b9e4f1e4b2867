"""The numpy restatement of the compact y-operators against the reference's
own derivyc / lapcdy (diffy_shear_hepta.f90:333-525) and compact filter
filterc (filter_shear_hepta.f90:128-165), compiled from its Fortran sources by
oracle/Makefile into oracle/_ref.

Twisted complex lines (m=2) and real periodic lines (m=1, the U(y), W(y)
means) at several my; relative L2 bar 1e-13 (both sides are FP64 LDL^H
solves of the same matrices, differing only in operation order).
"""
import math

import numpy as np
import pytest

from oracle import ref
from oracle.compact import YOps
from tests.boxes import rel_l2

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (needs the reference tree)")

TOL = 1e-13


@pytest.mark.parametrize("my", [33, 49, 97])
def test_twisted_lines(my):
    rng = np.random.default_rng(500 + my)
    nl = 12
    dy = 1.0 / my
    u = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    th = rng.uniform(-math.pi, math.pi, nl)
    th[0] = 0.0                     # untwisted line
    shp = np.exp(1j * th)
    shm = np.conj(shp)
    k2 = rng.uniform(5.0, 400.0, nl)
    d1, d2, lp = ref.run(my, dy, [("deriv", u, shp, shm, 1), ("deriv", u, shp, shm, 2),
                                  ("lap", u, k2, shp, shm)])
    ops = YOps(my, dy)
    assert rel_l2(ops.deriv(u, shp, shm, 1), d1) < TOL
    assert rel_l2(ops.deriv(u, shp, shm, 2), d2) < TOL
    assert rel_l2(ops.helmholtz(u, k2, shp, shm), lp) < TOL


@pytest.mark.parametrize("my", [33, 65])
def test_real_lines(my):
    rng = np.random.default_rng(900 + my)
    nl = 4
    dy = 1.0 / my
    u = rng.normal(size=(my, nl))
    k2 = np.full(nl, 37.5)
    d1, d2, lp = ref.run(my, dy, [("deriv_real", u, 1), ("deriv_real", u, 2), ("lap_real", u, k2)])
    ops = YOps(my, dy)
    assert rel_l2(ops.deriv_real(u, 1), d1) < TOL
    assert rel_l2(ops.deriv_real(u, 2), d2) < TOL
    assert rel_l2(ops.helmholtz_real(u, 37.5), lp) < TOL


@pytest.mark.parametrize("my", [33, 49, 97])
def test_compact_filter(my):
    rng = np.random.default_rng(700 + my)
    nl = 8
    dy = 1.0 / my
    u = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    th = rng.uniform(-math.pi, math.pi, nl)
    shp = np.exp(1j * th)
    ur = rng.normal(size=(my, 3))
    f, fr = ref.run(my, dy, [("filter", u, shp, np.conj(shp)), ("filter_real", ur)])
    ops = YOps(my, dy)
    assert rel_l2(ops.filter(u, shp, np.conj(shp)), f) < TOL
    assert rel_l2(ops.filter_real(ur), fr) < TOL
