"""Compact hepta-diagonal finite differences in y with shear-periodic wrap.

Restates diffy_shear_hepta.f90 (reference):
  * coefficient design (prederiv1 68-185, prederiv2 187-296, update_dy 535-573)
  * cyclic Hermitian LDL^H factorisation (modChol7cyclic_cmplx 578-755)
  * twisted re-factorisation of the border rows (modChol_addshear 760-870,
    with the border pre-multiplication done in derivyc 389-406 / 431-445)
  * forward/backward substitution (solve7cyclic 900-954)
  * derivyc 333-450, lapcdy 466-525, visc3d 298-318.

Lines are batched: every routine takes arrays of shape (N, nb) -- row j of
all nb lines is ``u[j]`` -- and per-line twist factors ``shp`` of shape (nb,).
Indices are 0-based: interior rows 0..N-4, border rows N-3, N-2, N-1.

The twisted-cyclic operator with symmetric band a(0..3) and twist t maps
row r to sum_m a(|m|) u(r+m) with u(r+N) = t u(r) and u(r-N) = conj(t) u(r)
(the RHS wrap in derivyc uses ``shp`` above and ``shm`` below).
"""
import math
import numpy as np


# ----------------------------------------------------------------------------
# coefficient design (diffy_shear_hepta.f90:68-296)
# ----------------------------------------------------------------------------
def design_first():
    """(pr1, rh1h) over offsets -3..3 (index m+3).

    Scheme  sum_j b_j f'_{i+j} = sum_j a_j f_{i+j}, b_0 = 1, b even, a odd.
    Conditions: Taylor orders 1, 3, 5 and exact modified wavenumber at
    kappa = (0.5, 0.9, 0.7)*pi  (diffy_shear_hepta.f90:81-143).
    """
    n = 3
    kap = np.array([0.5, 0.9, 0.7]) * math.pi
    jj = np.arange(1, n + 1, dtype=np.float64)
    mat = np.zeros((2 * n, 2 * n))
    rhs = np.zeros(2 * n)
    for row, order in enumerate((1, 3, 5)):
        mat[row, :n] = order * jj ** (order - 1)
        mat[row, n:] = -(jj ** order)
    rhs[0] = -0.5
    for q in range(3):
        mat[3 + q, :n] = kap[q] * np.cos(kap[q] * jj)
        mat[3 + q, n:] = -np.sin(kap[q] * jj)
        rhs[3 + q] = -0.5 * kap[q]
    sol = np.linalg.solve(mat, rhs)
    pr = np.zeros(7)
    rh = np.zeros(7)
    pr[3] = 1.0
    for j in range(1, 4):
        pr[3 + j] = pr[3 - j] = sol[j - 1]
        rh[3 + j] = sol[n + j - 1]
        rh[3 - j] = -sol[n + j - 1]
    return pr, rh


def design_second():
    """(pr2, rh2h) over offsets -3..3.

    Conditions: consistency (a_0 + 2 sum a_j = 0), Taylor orders 2, 4, 6 and
    exact modified wavenumber at kappa = (0.9, 0.5, 0.7)*pi
    (diffy_shear_hepta.f90:199-279).
    """
    n = 3
    kap = np.array([0.9, 0.5, 0.7]) * math.pi
    jj = np.arange(1, n + 1, dtype=np.float64)
    mat = np.zeros((2 * n + 1, 2 * n + 1))
    rhs = np.zeros(2 * n + 1)
    # unknowns: b_1..b_3, a_0, a_1..a_3
    mat[0, n + 1:] = 1.0
    mat[0, n] = 0.5
    for row, order in zip((1, 2, 3), (2, 4, 6)):
        mat[row, :n] = order * (order - 1) * jj ** (order - 2)
        mat[row, n + 1:] = -(jj ** order)
    rhs[1] = -1.0
    for q in range(3):
        mat[4 + q, :n] = kap[q] ** 2 * np.cos(kap[q] * jj)
        mat[4 + q, n + 1:] = np.cos(kap[q] * jj)
        mat[4 + q, n] = 0.5
        rhs[4 + q] = -0.5 * kap[q] ** 2
    sol = np.linalg.solve(mat, rhs)
    pr = np.zeros(7)
    rh = np.zeros(7)
    pr[3] = 1.0
    rh[3] = sol[n]
    for j in range(1, 4):
        pr[3 + j] = pr[3 - j] = sol[j - 1]
        rh[3 + j] = rh[3 - j] = sol[n + j]
    return pr, rh


# ----------------------------------------------------------------------------
# LDL^H of the twisted-cyclic Hermitian hepta matrix
# ----------------------------------------------------------------------------
class Factor:
    """L D L^H factors. Interior (real): D, Dm1, La, Lb, Lc of shape (N, nb)
    (or broadcastable); border rows (complex): Le (row N-3), Lf (row N-2),
    Lg (row N-1), each of shape (N, nb) with unused tails zero."""

    def __init__(self, N, nb):
        self.N = N
        z = lambda: np.zeros((N, nb))
        self.D, self.Dm1, self.La, self.Lb, self.Lc = z(), z(), z(), z(), z()
        zc = lambda: np.zeros((N, nb), dtype=np.complex128)
        self.Le, self.Lf, self.Lg = zc(), zc(), zc()

    def copy(self):
        f = Factor.__new__(Factor)
        f.N = self.N
        for nm in ("D", "Dm1", "La", "Lb", "Lc", "Le", "Lf", "Lg"):
            setattr(f, nm, getattr(self, nm).copy())
        return f


def _band(a):
    """a: (7,) stencil over -3..3 or (7, nb) -> four (nb,) / scalar rows a0..a3"""
    a = np.asarray(a)
    return a[3], a[4], a[5], a[6]


def modchol(a, N, t, nb=None):
    """Full factorisation (modChol7cyclic_cmplx, diffy_shear_hepta.f90:578-755).

    a : symmetric stencil, shape (7,) or (7, nb); t : twist, shape (nb,).
    """
    t = np.asarray(t, dtype=np.complex128)
    if nb is None:
        nb = t.shape[0]
    a0, a1, a2, a3 = _band(a)
    f = Factor(N, nb)
    D, Dm1, La, Lb, Lc = f.D, f.Dm1, f.La, f.Lb, f.Lc
    zero = np.zeros(nb)

    def g(arr, i):
        return arr[i] if i >= 0 else zero

    # interior rows 0..N-4 (1-based j = 1..N-3)
    last = N - 4
    for j in range(0, last + 1):
        D[j] = a0 - (g(Lc, j) * g(D, j - 3) * g(Lc, j)
                     + g(Lb, j) * g(D, j - 2) * g(Lb, j)
                     + g(La, j) * g(D, j - 1) * g(La, j))
        Dm1[j] = 1.0 / D[j]
        if j + 1 <= last:
            La[j + 1] = Dm1[j] * (a1 - (g(Lc, j + 1) * g(D, j - 2) * g(Lb, j)
                                        + g(Lb, j + 1) * g(D, j - 1) * g(La, j)))
        if j + 2 <= last:
            Lb[j + 2] = Dm1[j] * (a2 - g(Lc, j + 2) * g(D, j - 1) * g(La, j))
        if j + 3 <= last:
            Lc[j + 3] = Dm1[j] * a3

    def border(Lx, src):
        # Lx(i) = Dm1(i) * (A(row, i) - Lx(i-3) D(i-3) Lc(i) - Lx(i-2) D(i-2) Lb(i)
        #                   - Lx(i-1) D(i-1) La(i)),  i = 0..N-4
        for i in range(0, last + 1):
            acc = src(i) - g(Lx, i - 3) * g(D, i - 3) * g(Lc, i) \
                - g(Lx, i - 2) * g(D, i - 2) * g(Lb, i) \
                - g(Lx, i - 1) * g(D, i - 1) * g(La, i)
            Lx[i] = Dm1[i] * acc

    zc = np.zeros(nb, dtype=np.complex128)

    # row N-3 couples to columns 0 (a3*t), N-6 (a3), N-5 (a2), N-4 (a1)
    def src_e(i):
        v = zc.copy()
        if i == 0:
            v = v + a3 * t
        if i == N - 6:
            v = v + a3
        if i == N - 5:
            v = v + a2
        if i == N - 4:
            v = v + a1
        return v

    def src_f(i):
        v = zc.copy()
        if i == 0:
            v = v + a2 * t
        if i == 1:
            v = v + a3 * t
        if i == N - 5:
            v = v + a3
        if i == N - 4:
            v = v + a2
        return v

    def src_g(i):
        v = zc.copy()
        if i == 0:
            v = v + a1 * t
        if i == 1:
            v = v + a2 * t
        if i == 2:
            v = v + a3 * t
        if i == N - 4:
            v = v + a3
        return v

    border(f.Le, src_e)
    _tail(f, a, last)
    return _tail_fg(f, a, last, src_f, src_g, border)


def _tail(f, a, last):
    a0, a1, a2, a3 = _band(a)
    N = f.N
    # D(N-3) = a0 - sum_{k<=N-4} |Le|^2 D
    s = np.zeros_like(f.D[0])
    for k in range(0, last + 1):
        s = s + (f.Le[k] * np.conj(f.Le[k])).real * f.D[k]
    f.D[N - 3] = a0 - s
    f.Dm1[N - 3] = 1.0 / f.D[N - 3]


def _tail_fg(f, a, last, src_f, src_g, border):
    a0, a1, a2, a3 = _band(a)
    N = f.N
    border(f.Lf, src_f)
    acc = np.zeros_like(f.Le[0])
    for k in range(0, last + 1):
        acc = acc + f.Lf[k] * np.conj(f.Le[k]) * f.D[k]
    f.Lf[N - 3] = f.Dm1[N - 3] * (a1 - acc)
    s = np.zeros_like(f.D[0])
    for k in range(0, N - 2):
        s = s + (f.Lf[k] * np.conj(f.Lf[k])).real * f.D[k]
    f.D[N - 2] = a0 - s
    f.Dm1[N - 2] = 1.0 / f.D[N - 2]
    border(f.Lg, src_g)
    acc = np.zeros_like(f.Le[0])
    for k in range(0, N - 3):
        acc = acc + f.Lg[k] * f.D[k] * np.conj(f.Le[k])
    f.Lg[N - 3] = f.Dm1[N - 3] * (a2 - acc)
    acc = np.zeros_like(f.Le[0])
    for k in range(0, N - 2):
        acc = acc + f.Lg[k] * f.D[k] * np.conj(f.Lf[k])
    f.Lg[N - 2] = f.Dm1[N - 2] * (a1 - acc)
    s = np.zeros_like(f.D[0])
    for k in range(0, N - 1):
        s = s + (f.Lg[k] * np.conj(f.Lg[k])).real * f.D[k]
    f.D[N - 1] = a0 - s
    f.Dm1[N - 1] = 1.0 / f.D[N - 1]
    return f


def twist_base(base, a, t):
    """Re-factor the border rows of an untwisted base factor for twist t.

    The first border entries are the base (real) entries times t
    (derivyc, diffy_shear_hepta.f90:398-405); the last three entries of each
    border row and D(N-3..N-1) are recomputed (modChol_addshear 760-870).
    """
    N = base.N
    t = np.asarray(t, dtype=np.complex128)
    nb = t.shape[0]
    a0, a1, a2, a3 = _band(a)
    f = Factor(N, nb)
    for nm in ("D", "Dm1", "La", "Lb", "Lc"):
        getattr(f, nm)[...] = getattr(base, nm)
    f.Le[...] = base.Le
    f.Lf[...] = base.Lf
    f.Lg[...] = base.Lg
    f.Le[:N - 6] = base.Le[:N - 6].real * t
    f.Lf[:N - 5] = base.Lf[:N - 5].real * t
    f.Lg[:N - 4] = base.Lg[:N - 4].real * t
    D, La, Lb, Lc, Dm1 = f.D, f.La, f.Lb, f.Lc, f.Dm1

    def rec(Lx, i, aval):
        return Dm1[i] * (aval - Lx[i - 3] * D[i - 3] * Lc[i]
                         - Lx[i - 2] * D[i - 2] * Lb[i] - Lx[i - 1] * D[i - 1] * La[i])

    last = N - 4
    f.Le[N - 6] = rec(f.Le, N - 6, a3)
    f.Le[N - 5] = rec(f.Le, N - 5, a2)
    f.Le[N - 4] = rec(f.Le, N - 4, a1)
    _tail(f, a, last)
    f.Lf[N - 5] = rec(f.Lf, N - 5, a3)
    f.Lf[N - 4] = rec(f.Lf, N - 4, a2)
    acc = np.zeros(nb, dtype=np.complex128)
    for k in range(0, last + 1):
        acc = acc + f.Lf[k] * np.conj(f.Le[k]) * D[k]
    f.Lf[N - 3] = Dm1[N - 3] * (a1 - acc)
    s = np.zeros(nb)
    for k in range(0, N - 2):
        s = s + (f.Lf[k] * np.conj(f.Lf[k])).real * D[k]
    D[N - 2] = a0 - s
    Dm1[N - 2] = 1.0 / D[N - 2]
    f.Lg[N - 4] = rec(f.Lg, N - 4, a3)
    acc = np.zeros(nb, dtype=np.complex128)
    for k in range(0, N - 3):
        acc = acc + f.Lg[k] * np.conj(f.Le[k]) * D[k]
    f.Lg[N - 3] = Dm1[N - 3] * (a2 - acc)
    acc = np.zeros(nb, dtype=np.complex128)
    for k in range(0, N - 2):
        acc = acc + f.Lg[k] * np.conj(f.Lf[k]) * D[k]
    f.Lg[N - 2] = Dm1[N - 2] * (a1 - acc)
    s = np.zeros(nb)
    for k in range(0, N - 1):
        s = s + (f.Lg[k] * np.conj(f.Lg[k])).real * D[k]
    D[N - 1] = a0 - s
    Dm1[N - 1] = 1.0 / D[N - 1]
    return f


def solve(f, b):
    """Forward/backward substitution (solve7cyclic, diffy_shear_hepta.f90:900-954)."""
    N = f.N
    b = np.asarray(b, dtype=np.complex128)
    y = np.zeros_like(b)
    y[0] = b[0]
    y[1] = b[1] - f.La[1] * y[0]
    y[2] = b[2] - (f.Lb[2] * y[0] + f.La[2] * y[1])
    for i in range(3, N - 3):
        y[i] = b[i] - (f.Lc[i] * y[i - 3] + f.Lb[i] * y[i - 2] + f.La[i] * y[i - 1])
    for row, Lx in ((N - 3, f.Le), (N - 2, f.Lf), (N - 1, f.Lg)):
        acc = np.zeros_like(b[0])
        for k in range(0, row):
            acc = acc + Lx[k] * y[k]
        y[row] = b[row] - acc
    cLe, cLf, cLg = np.conj(f.Le), np.conj(f.Lf), np.conj(f.Lg)
    x = np.zeros_like(b)
    x[N - 1] = y[N - 1] * f.Dm1[N - 1]
    x[N - 2] = y[N - 2] * f.Dm1[N - 2] - cLg[N - 2] * x[N - 1]
    x[N - 3] = y[N - 3] * f.Dm1[N - 3] - cLf[N - 3] * x[N - 2] - cLg[N - 3] * x[N - 1]
    x[N - 4] = (y[N - 4] * f.Dm1[N - 4] - cLe[N - 4] * x[N - 3] - cLf[N - 4] * x[N - 2]
                - cLg[N - 4] * x[N - 1])
    x[N - 5] = (y[N - 5] * f.Dm1[N - 5] - f.La[N - 4] * x[N - 4] - cLe[N - 5] * x[N - 3]
                - cLf[N - 5] * x[N - 2] - cLg[N - 5] * x[N - 1])
    x[N - 6] = (y[N - 6] * f.Dm1[N - 6] - f.La[N - 5] * x[N - 5] - f.Lb[N - 4] * x[N - 4]
                - cLe[N - 6] * x[N - 3] - cLf[N - 6] * x[N - 2] - cLg[N - 6] * x[N - 1])
    for i in range(N - 7, -1, -1):
        x[i] = (y[i] * f.Dm1[i] - f.La[i + 1] * x[i + 1] - f.Lb[i + 2] * x[i + 2]
                - f.Lc[i + 3] * x[i + 3] - cLe[i] * x[N - 3] - cLf[i] * x[N - 2]
                - cLg[i] * x[N - 1])
    return x


def stencil(c, u, shp, shm, skip_center=False):
    """Twisted-cyclic (2 na + 1)-point product, na = (len(c) - 1) / 2:
    out(j) = sum_m c(m) u(j+m), with u(j+N) = shp u(j), u(j-N) = shm u(j)
    (derivyc 358-383, lapcdy 487-514, matvecNcyclic 960-996)."""
    N = u.shape[0]
    na = (len(c) - 1) // 2
    ext = np.concatenate([u[N - na:] * shm, u, u[:na] * shp], axis=0)
    out = np.zeros_like(u, dtype=np.complex128)
    for m in range(-na, na + 1):
        if m == 0 and skip_center:
            continue
        out = out + c[m + na] * ext[na + m:na + m + N]
    return out


# Compact filter of filter_shear_hepta.f90:53-69 (prefilter): symmetric
# hepta-diagonal LHS, symmetric 9-point RHS, not scaled by the grid step.
FILTER_PR = np.array([0.053003478448814, 0.315003368623046, 0.738981897704866, 1.0,
                      0.738981897704866, 0.315003368623046, 0.053003478448814])
FILTER_RH = np.array([-0.000179828066167, 0.054442102978149, 0.309968182770373,
                      0.749052269410218, 0.987412035368317, 0.749052269410218,
                      0.309968182770373, 0.054442102978149, -0.000179828066167])


class YOps:
    """Per-grid compact operators (derivadas, diffy_shear_hepta.f90:535-573)."""

    def __init__(self, my, dy):
        self.N = my
        self.h = dy
        self.pr1, rh1h = design_first()
        self.pr2, rh2h = design_second()
        self.rh1 = rh1h / dy
        self.rh2 = rh2h / (dy * dy)
        one = np.ones(1, dtype=np.complex128)
        self.base1 = modchol(self.pr1, my, one)
        self.base2 = modchol(self.pr2, my, one)

    def twisted(self, order, shp):
        base = self.base1 if order == 1 else self.base2
        pr = self.pr1 if order == 1 else self.pr2
        return twist_base(base, pr, shp)

    def deriv(self, u, shp, shm, order):
        """derivyc with m=2 (complex lines, twisted)."""
        f = self.twisted(order, shp)
        if order == 1:
            r = stencil(self.rh1, u, shp, shm, skip_center=True)
        else:
            r = stencil(self.rh2, u, shp, shm)
        return solve(f, r)

    def deriv_real(self, u, order):
        """derivyc with m=1: real periodic line (the 00 modes)."""
        u = np.asarray(u, dtype=np.float64).reshape(self.N, -1)
        one = np.ones(u.shape[1], dtype=np.complex128)
        if order == 1:
            r = stencil(self.rh1, u.astype(np.complex128), one, one, skip_center=True)
            f = _bcast(self.base1, u.shape[1])
        else:
            r = stencil(self.rh2, u.astype(np.complex128), one, one)
            f = _bcast(self.base2, u.shape[1])
        return solve(f, r).real

    def filter(self, u, shp, shm):
        """filterc m=2 (filter_shear_hepta.f90:128-165): B_f x = R_f u on
        twisted lines, factor re-used from the untwisted one as in derivyc."""
        if not hasattr(self, "basef"):
            self.basef = modchol(FILTER_PR, self.N, np.ones(1, dtype=np.complex128))
        f = twist_base(self.basef, FILTER_PR, shp)
        return solve(f, stencil(FILTER_RH, u, shp, shm))

    def filter_real(self, u):
        """filterc m=1: real periodic line (the 00 modes)."""
        if not hasattr(self, "basef"):
            self.basef = modchol(FILTER_PR, self.N, np.ones(1, dtype=np.complex128))
        u = np.asarray(u, dtype=np.float64).reshape(self.N, -1)
        one = np.ones(u.shape[1], dtype=np.complex128)
        r = stencil(FILTER_RH, u.astype(np.complex128), one, one)
        return solve(_bcast(self.basef, u.shape[1]), r).real

    def lap_matrix(self, k2):
        k2 = np.atleast_1d(np.asarray(k2, dtype=np.float64))
        return self.rh2[:, None] - k2[None, :] * self.pr2[:, None]

    def helmholtz(self, u, k2, shp, shm):
        """lapcdy m=2: solve (d_yy - k2) p = u on twisted lines."""
        f = modchol(self.lap_matrix(k2), self.N, shp)
        r = stencil(self.pr2, u, shp, shm)
        return solve(f, r)

    def helmholtz_real(self, u, k2):
        """lapcdy m=1: real periodic line."""
        u = np.asarray(u, dtype=np.float64).reshape(self.N, -1)
        one = np.ones(u.shape[1], dtype=np.complex128)
        f = modchol(self.lap_matrix(np.full(u.shape[1], k2)), self.N, one)
        r = stencil(self.pr2, u.astype(np.complex128), one, one)
        return solve(f, r).real


def _bcast(f, nb):
    g = Factor(f.N, nb)
    for nm in ("D", "Dm1", "La", "Lb", "Lc", "Le", "Lf", "Lg"):
        getattr(g, nm)[...] = getattr(f, nm)
    return g


def dense_matrix(c, N, t):
    """Dense twisted-cyclic matrix of a 7-point stencil (test helper)."""
    A = np.zeros((N, N), dtype=np.complex128)
    for r in range(N):
        for m in range(-3, 4):
            col = r + m
            fac = 1.0
            if col >= N:
                col -= N
                fac = t
            elif col < 0:
                col += N
                fac = np.conj(t)
            A[r, col] += c[m + 3] * fac
    return A
