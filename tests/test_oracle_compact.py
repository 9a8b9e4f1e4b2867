"""Pins the oracle's compact y-operators (diffy_shear_hepta.f90) without the
reference: design conditions evaluated independently, the twisted-Fourier
eigenvalue identity, and dense brute-force solves of the assembled matrices."""
import math

import numpy as np
import pytest

from oracle.compact import (YOps, dense_matrix, design_first, design_second, modchol,
                            solve, twist_base)


def _symbol(c, kap):
    m = np.arange(-3, 4)
    return np.sum(c * np.exp(1j * kap * m))


def test_first_derivative_design():
    pr, rh = design_first()
    assert pr[3] == 1.0 and rh[3] == 0.0
    assert np.allclose(pr, pr[::-1]) and np.allclose(rh, -rh[::-1])
    # modified wavenumber exact at the three design points (diffy_shear_hepta.f90:81-88)
    for kap in np.array([0.5, 0.9, 0.7]) * math.pi:
        mw = (_symbol(rh, kap) / _symbol(pr, kap)).imag
        assert abs(mw - kap) < 1e-13
    # Taylor: exact derivative of x, x^3, x^5 at 0
    x = np.arange(-3, 4, dtype=float)
    for order in (1, 3, 5):
        lhs = np.sum(pr * order * x ** (order - 1)) if order > 1 else np.sum(pr)
        assert abs(lhs - np.sum(rh * x ** order)) < 1e-12
    # a wrong design point (kappa = 0.8 pi) must not be exact
    kap = 0.8 * math.pi
    assert abs((_symbol(rh, kap) / _symbol(pr, kap)).imag - kap) > 1e-6


def test_second_derivative_design():
    pr, rh = design_second()
    assert pr[3] == 1.0
    assert np.allclose(pr, pr[::-1]) and np.allclose(rh, rh[::-1])
    assert abs(np.sum(rh)) < 1e-13                       # d2/dy2 of a constant
    for kap in np.array([0.9, 0.5, 0.7]) * math.pi:
        mw = -(_symbol(rh, kap) / _symbol(pr, kap)).real
        assert abs(mw - kap * kap) < 1e-12
    x = np.arange(-3, 4, dtype=float)
    for order in (2, 4, 6):
        lhs = np.sum(pr * order * (order - 1) * x ** (order - 2))
        assert abs(lhs - np.sum(rh * x ** order)) < 1e-10


@pytest.mark.parametrize("N", [33, 49, 64])
def test_factorisation_reconstructs_matrix(N):
    rng = np.random.default_rng(N)
    pr, _ = design_first()
    t = np.exp(1j * rng.uniform(-3, 3, 3))
    f = modchol(pr, N, t)
    for l in range(3):
        L = np.eye(N, dtype=complex)
        for i in range(1, N - 3):
            L[i, i - 1] = f.La[i, l]
        for i in range(2, N - 3):
            L[i, i - 2] = f.Lb[i, l]
        for i in range(3, N - 3):
            L[i, i - 3] = f.Lc[i, l]
        L[N - 3, :N - 3] = f.Le[:N - 3, l]
        L[N - 2, :N - 2] = f.Lf[:N - 2, l]
        L[N - 1, :N - 1] = f.Lg[:N - 1, l]
        A = L @ np.diag(f.D[:, l]) @ L.conj().T
        assert np.max(np.abs(A - dense_matrix(pr, N, t[l]))) < 1e-13


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("N", [33, 49, 97])
def test_deriv_matches_dense_solve(order, N):
    rng = np.random.default_rng(10 * N + order)
    h = 1.0 / N
    ops = YOps(N, h)
    nb = 6
    t = np.exp(1j * rng.uniform(-math.pi, math.pi, nb))
    u = rng.normal(size=(N, nb)) + 1j * rng.normal(size=(N, nb))
    d = ops.deriv(u, t, np.conj(t), order)
    pr = ops.pr1 if order == 1 else ops.pr2
    rh = ops.rh1 if order == 1 else ops.rh2
    for l in range(nb):
        ref = np.linalg.solve(dense_matrix(pr, N, t[l]), dense_matrix(rh, N, t[l]) @ u[:, l])
        assert np.linalg.norm(d[:, l] - ref) <= 1e-12 * np.linalg.norm(ref)


def test_twisted_base_equals_full_factorisation():
    N = 49
    rng = np.random.default_rng(5)
    pr, _ = design_second()
    t = np.exp(1j * rng.uniform(-3, 3, 4))
    base = modchol(pr, N, np.ones(1, dtype=complex))
    tw = twist_base(base, pr, t)
    full = modchol(pr, N, t)
    for nm in ("D", "Le", "Lf", "Lg"):
        assert np.max(np.abs(getattr(tw, nm) - getattr(full, nm))) < 1e-14


@pytest.mark.parametrize("order", [1, 2])
def test_twisted_fourier_identity(order):
    """u_j = exp(i kappa j), exp(i kappa N) = shp  ->  D u = lambda(kappa) u"""
    N = 49
    h = 1.0 / N
    ops = YOps(N, h)
    pr = ops.pr1 if order == 1 else ops.pr2
    rh = ops.rh1 if order == 1 else ops.rh2
    theta = 0.731
    shp = np.array([np.exp(1j * theta)])
    for m in (0, 3, 11, -7):
        kap = (theta + 2 * math.pi * m) / N
        u = np.exp(1j * kap * np.arange(N))[:, None]
        lam = _symbol(rh, kap) / _symbol(pr, kap)
        d = ops.deriv(u, shp, np.conj(shp), order)
        scale = max(abs(lam), np.sum(np.abs(rh)))   # rounding scales with the stencil norm
        assert np.max(np.abs(d[:, 0] - lam * u[:, 0])) <= 1e-13 * scale  # LHS cond ~ 400


def test_helmholtz_identity_and_dense():
    N = 49
    h = 1.0 / N
    ops = YOps(N, h)
    theta = -1.2
    shp = np.array([np.exp(1j * theta)])
    k2 = 37.0
    for m in (0, 2, -5):
        kap = (theta + 2 * math.pi * m) / N
        u = np.exp(1j * kap * np.arange(N))[:, None]
        lam = _symbol(ops.pr2, kap) / (_symbol(ops.rh2, kap) - k2 * _symbol(ops.pr2, kap))
        p = ops.helmholtz(u, np.array([k2]), shp, np.conj(shp))
        assert np.max(np.abs(p[:, 0] - lam * u[:, 0])) < 1e-13
    rng = np.random.default_rng(1)
    u = rng.normal(size=(N, 1)) + 1j * rng.normal(size=(N, 1))
    p = ops.helmholtz(u, np.array([k2]), shp, np.conj(shp))
    A = dense_matrix(ops.rh2 - k2 * ops.pr2, N, shp[0])
    ref = np.linalg.solve(A, dense_matrix(ops.pr2, N, shp[0]) @ u[:, 0])
    assert np.linalg.norm(p[:, 0] - ref) < 1e-13 * np.linalg.norm(ref)


def test_derivative_converges_on_smooth_function():
    """high-order accuracy (error ratio > 40 per halving of h) on a smooth periodic line"""
    errs = []
    for N in (65, 129):
        h = 1.0 / N
        ops = YOps(N, h)
        y = np.arange(N) * h
        f = np.exp(np.sin(2 * math.pi * y))[:, None].astype(complex)
        df = (2 * math.pi * np.cos(2 * math.pi * y) * np.exp(np.sin(2 * math.pi * y)))
        one = np.ones(1, dtype=complex)
        d = ops.deriv(f, one, one, 1)[:, 0].real
        errs.append(np.max(np.abs(d - df)))
    assert errs[1] < errs[0] / 40.0


def test_real_line_equals_untwisted_complex_line():
    N = 49
    ops = YOps(N, 1.0 / N)
    rng = np.random.default_rng(3)
    u = rng.normal(size=N)
    one = np.ones(1, dtype=complex)
    for order in (1, 2):
        a = ops.deriv_real(u, order)[:, 0]
        b = ops.deriv(u[:, None].astype(complex), one, one, order)[:, 0]
        assert np.max(np.abs(a - b.real)) < 1e-12 and np.max(np.abs(b.imag)) < 1e-12
