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


@pytest.mark.parametrize("order", [1, 2])
def test_factor_reuse_ias0(order):
    """ias=0 (diffy_shear_hepta.f90:389-406, 431-445) reuses the factor of the
    previous ias=1 call; the reference pairs the calls with the same twist
    (cross.f90:440-442, 659-660), where the oracle's one solve per twist gives
    the same numbers.  Reusing a factor of another twist does not: the
    reference really skips the refactor."""
    my = 49
    rng = np.random.default_rng(40 + order)
    nl = 6
    dy = 1.0 / my
    u = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    ua = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    shp = np.exp(1j * rng.uniform(-math.pi, math.pi, nl))
    other = np.exp(1j * rng.uniform(-math.pi, math.pi, nl))
    same, mixed = ref.run(my, dy, [("deriv_reuse", u, ua, shp, shp, np.conj(shp), order),
                                   ("deriv_reuse", u, ua, other, shp, np.conj(shp), order)])
    ops = YOps(my, dy)
    want = ops.deriv(u, shp, np.conj(shp), order)
    assert rel_l2(same, want) < TOL
    assert rel_l2(mixed, want) > 1e-6
