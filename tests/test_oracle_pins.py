"""Pins of three oracle functions that the closed-form runs leave open.

* The 00-mode forcing rf0u, rf0w of ``Cross.hvhg`` (cross.f90:1453-1490:
  rf0u(j) = H1(0,0), rf0w(j) = H3(0,0) with H1 = v (omega_3 + 2 chi) -
  w omega_2, H3 = u omega_2 - v omega_1): on a random dealiased state of a
  tiny grid, against plane means of the products formed in physical space
  by direct double DFT sums (no FFT).  A swapped pair of components or a
  flipped sign fails by orders of magnitude.
* The adaptive time step ``Cross._new_dt`` (cross.f90:1598-1712,
  initcr.f90:474-495): states whose physical maxima are known exactly (a
  pure mean profile U(y); single Fourier modes that peak on grid points),
  so each branch is checked against the number it must give: umax(0), not
  umax(1), in the x term (iadd_mode = 5, cross.f90:1645-1648); the y and z
  terms; the explicit viscous cap (cross.f90:1665-1673), absent in implicit
  runs; the dt < 1e-10 stop; the dump and etime clips (cross.f90:1693-1712);
  and the umax >= 1 first step of a generated field (cross.f90:1211-1219).
* The shifted windows of y-dealiasing variant 2 for kx != 0
  (operators.f90:433-460): the set of y-Fourier modes each line keeps,
  mode by mode, against sets worked out by hand from the window bounds
  (commented per case) at my = 33, mx1 = 9, Lx/Ly = 2.
"""
import math

import numpy as np
import pytest

from oracle.cross import Cross, set_conj
from oracle.grid import Grid
from tests.boxes import ALP, GAM, LY


# ------------------------------------------------------------------ rf0u, rf0w
def _phys(g, F):
    """real field on the mgalx x mgalz grid of a half spectrum F[kx][kz]
    (kx >= 0, Hermitian kx = 0 line) by direct double sums:
    f = sum_kz F[0,kz] e^{i kz z} + 2 Re sum_{kx>0} sum_kz F[kx,kz] e^{i (kx x + kz z)}"""
    x = np.arange(g.mgalx) * (2 * math.pi / g.alp / g.mgalx)
    z = np.arange(g.mgalz) * (2 * math.pi / g.gam / g.mgalz)
    ex = np.exp(1j * np.outer(g.kx, x))          # [kx][x]
    ez = np.exp(1j * np.outer(g.kz, z))          # [kz][z]
    kx0 = np.real(np.einsum("k,kz->z", F[0], ez))
    half = np.einsum("ik,ix,kz->xz", F[1:], ex[1:], ez)
    return kx0[None, :] + 2.0 * np.real(half)


def _forcing_case(chi, vbulk, seed):
    g = Grid(12, 17, 12, Re=500.0, alp=ALP, gam=GAM, Ly=LY, s=1.0, chi=chi)
    rng = np.random.default_rng(seed)

    def field():
        f = rng.normal(size=(g.my, g.kxn, g.mz)) + 1j * rng.normal(size=(g.my, g.kxn, g.mz))
        f[:, 0, 0] = f[:, 0, 0].real
        return set_conj(f, g.nz)

    c = Cross(g, explicit=1, nhist=10 ** 9, vbulk=vbulk)
    c.set_state(field(), field(), rng.normal(size=g.my), rng.normal(size=g.my))
    c.hg, c.hv, c.dvordy = field(), field(), field()
    # the fields hvhg reconstructs u, v, w and omega from
    dudy = c.ops.deriv_real(c.u00, 1)[:, 0]
    dwdy = c.ops.deriv_real(c.w00, 1)[:, 0]
    u, v, w, o1, o2, o3 = c.derived(set_conj(c.phi.copy(), g.nz), set_conj(c.vor.copy(), g.nz),
                                    c.hg, c.hv, c.dvordy, dudy, dwdy)
    c.istep = 3
    c.hvhg(2, 0)
    return g, c, [np.array([_phys(g, f[j]) for j in range(g.my)]) for f in (u, v, w, o1, o2, o3)]


@pytest.mark.parametrize("chi,vbulk,seed", [(0.0, 0.0, 1), (0.37, 0.21, 2)])
def test_forcing_is_the_plane_mean_of_u_cross_omega(chi, vbulk, seed):
    g, c, (U, V, W, O1, O2, O3) = _forcing_case(chi, vbulk, seed)
    H1 = V * (O3 + 2 * chi) - W * O2
    H3 = U * O2 - V * O1
    ref_u = H1.mean(axis=(1, 2))
    ref_w = H3.mean(axis=(1, 2))
    scale = np.linalg.norm(ref_u) + np.linalg.norm(ref_w)
    assert np.linalg.norm(c.rf0u - ref_u) < 1e-13 * scale
    assert np.linalg.norm(c.rf0w - ref_w) < 1e-13 * scale
    # planted mistakes fail: the components swapped, a sign flipped, the
    # rotation term dropped (when chi != 0)
    assert np.linalg.norm(c.rf0u - ref_w) > 1e-3 * scale
    assert np.linalg.norm(c.rf0u - (-V * (O3 + 2 * chi) - W * O2).mean(axis=(1, 2))) > 1e-3 * scale
    assert np.linalg.norm(c.rf0w - (U * O2 + V * O1).mean(axis=(1, 2))) > 1e-3 * scale
    if chi:
        assert np.linalg.norm(c.rf0u - (V * O3 - W * O2).mean(axis=(1, 2))) > 1e-3 * scale


# ------------------------------------------------------------------ _new_dt
def _dt_case(g, c, u00=None, w00=None, lines=()):
    """run the rkstep-1 part of hvhg on a state given by its mean profiles and
    single (field, kx index, kz index, value) entries of v, dv/dy, omega_y"""
    c.vor, c.phi = g.zeros(), g.zeros()
    c.hg, c.hv, c.dvordy = g.zeros(), g.zeros(), g.zeros()
    c.u00 = np.zeros(g.my) if u00 is None else np.asarray(u00, float)
    c.w00 = np.zeros(g.my) if w00 is None else np.asarray(w00, float)
    for name, i, k, val in lines:
        getattr(c, name)[:, i, k] = val
    c.hvhg(1, 0)
    return c


def _grid(Re=4000.0, CFL=0.5):
    return Grid(24, 25, 24, Re=Re, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=CFL)


def _consts(g):
    """cflx, cfly, cflz, cflvv as initcr.f90:483-497 defines them"""
    mx, mz = 2 * (g.mgalx // 3), 2 * (g.mgalz // 3) - 1
    dx, dz, dy = 2 * math.pi / g.alp / mx, 2 * math.pi / g.gam / mz, g.Ly / g.my
    cflvv = 1.0 / g.Re / g.CFL * max(math.pi ** 2 / dx ** 2, 6.25 / dy ** 2, math.pi ** 2 / dz ** 2)
    return math.pi / g.CFL / dx, 2.5 / g.CFL / dy, math.pi / g.CFL / dz, cflvv


def test_dt_uses_the_fluctuation_maximum_in_x():
    g = _grid()
    cflx, cfly, cflz, cflvv = _consts(g)
    y = g.y
    u00 = 0.3 * np.sin(2 * math.pi * y / g.Ly)          # max over the samples, exactly
    c = _dt_case(g, Cross(g, explicit=1), u00=u00)
    umax0 = np.max(np.abs(u00))
    umax1 = np.max(np.abs(u00 + g.s * y))
    assert c.umax[0] == umax0 and c.umax[1] == umax1 and c.umax[2] == 0.0 and c.umax[3] == 0.0
    assert c.Deltat == pytest.approx(1.0 / (umax0 * cflx), rel=1e-15)
    # umax(1) (with the shear) would give another step
    assert abs(1.0 / (umax1 * cflx) - c.Deltat) > 0.1 * c.Deltat
    assert c.CFLv == pytest.approx(g.CFL * cflvv * c.Deltat, rel=1e-15)
    assert c.dtr == pytest.approx(g.Re / c.Deltat, rel=1e-15) and not c.istop


@pytest.mark.parametrize("term", ["x", "y", "z"])
def test_dt_each_direction(term):
    """u = a cos(alpha x) (from dv/dy on kx = 1, kz = 0), v = 2 b cos(gamma z)
    (kx = 0, kz = +-1), w = 2 c cos(alpha x) (from omega_y on kx = 1, kz = 0):
    each peaks on the grid point x = 0 or z = 0, so umax is exact"""
    g = _grid()
    cflx, cfly, cflz, _ = _consts(g)
    amp = {"x": (0.4, 0.01, 0.01), "y": (0.01, 0.5, 0.01), "z": (0.01, 0.01, 0.6)}[term]
    a, b, cc = amp
    # u1 = i alpha dv / alpha^2 = a/2  ->  dv = -i alpha a / 2
    # u3 = i alpha ome2 / alpha^2 = cc  ->  ome2 = -i alpha cc
    lines = [("hv", 1, 0, -1j * g.alp * a / 2), ("hg", 0, 1, b), ("hg", 0, g.mz - 1, b),
             ("vor", 1, 0, -1j * g.alp * cc)]
    c = _dt_case(g, Cross(g, explicit=1), lines=lines)
    assert c.umax[0] == pytest.approx(a, rel=1e-14)
    assert c.umax[2] == pytest.approx(2 * b, rel=1e-14)
    assert c.umax[3] == pytest.approx(2 * cc, rel=1e-14)
    want = 1.0 / max(a * cflx, 2 * b * cfly, 2 * cc * cflz)
    assert c.Deltat == pytest.approx(want, rel=1e-13)


def test_dt_viscous_cap_explicit_only():
    g = _grid(Re=20.0)
    cflx, _, _, cflvv = _consts(g)
    u00 = 0.01 * np.cos(2 * math.pi * g.y / g.Ly)
    ce = _dt_case(g, Cross(g, explicit=1), u00=u00)
    ci = _dt_case(g, Cross(g, explicit=0), u00=u00)
    dt_adv = 1.0 / (np.max(np.abs(u00)) * cflx)
    assert g.CFL * cflvv * dt_adv > g.CFL          # the cap binds
    assert ce.Deltat == pytest.approx(1.0 / cflvv, rel=1e-15)
    assert ci.Deltat == pytest.approx(dt_adv, rel=1e-15)
    assert ce.CFLv == pytest.approx(g.CFL * cflvv * dt_adv, rel=1e-15)


def test_dt_divergence_stop():
    g = _grid()
    c = _dt_case(g, Cross(g, explicit=0), u00=np.full(g.my, 1e9))
    assert c.Deltat < 1e-10 and c.istop == 1


def test_dt_dump_and_etime_clips():
    g = _grid()
    cflx = _consts(g)[0]
    u00 = 0.2 * np.ones(g.my)
    dt = 1.0 / (0.2 * cflx)
    # dump: time + dt passes dumptime -> dt = dumptime - time, next dump
    c = Cross(g, explicit=1)
    c.time, c.dumptime, c.dtimag, c.etime = 1.0, 1.0 + 0.25 * dt, 0.5, 10.0
    _dt_case(g, c, u00=u00)
    assert c.Deltat == pytest.approx(0.25 * dt, rel=1e-12) and c.idump == 1 and c.istop == 0
    assert c.dumptime == pytest.approx(1.5 + 0.25 * dt, rel=1e-15)
    assert c.dtr == pytest.approx(g.Re / c.Deltat, rel=1e-15)
    # a dump at etime also stops
    c = Cross(g, explicit=1)
    c.time, c.dumptime, c.dtimag, c.etime = 1.0, 1.0 + 0.25 * dt, 0.5, 1.0 + 0.25 * dt
    _dt_case(g, c, u00=u00)
    assert c.idump == 1 and c.istop == 1
    # etime without a dump
    c = Cross(g, explicit=1)
    c.time, c.dumptime, c.etime = 1.0, 100.0, 1.0 + 0.5 * dt
    _dt_case(g, c, u00=u00)
    assert c.Deltat == pytest.approx(0.5 * dt, rel=1e-12) and c.istop == 1 and c.idump == 0
    # neither
    c = Cross(g, explicit=1)
    c.time, c.dumptime, c.etime = 1.0, 100.0, 100.0
    _dt_case(g, c, u00=u00)
    assert c.Deltat == pytest.approx(dt, rel=1e-15) and c.istop == 0 and c.idump == 0


@pytest.mark.parametrize("s,want", [(1.0, 1.0), (0.0, 0.1)])
def test_dt_generated_field_first_step(s, want):
    """readflag = 0, istep = 1: umax starts at 1 (sheared) or uprim"""
    g = Grid(24, 25, 24, Re=4000.0, alp=ALP, gam=GAM, Ly=LY, s=s)
    cflx, cfly, cflz, _ = _consts(g)
    c = Cross(g, explicit=1, readflag=0, uprim=0.1)
    c.vor, c.phi = g.zeros(), g.zeros()
    c.hg, c.hv, c.dvordy = g.zeros(), g.zeros(), g.zeros()
    c.u00 = 0.01 * np.ones(g.my)
    c.w00 = np.zeros(g.my)
    c.istep = 1
    c.hvhg(1, 0)
    assert c.umax[0] == want and c.umax[2] == want and c.umax[3] == want
    assert c.Deltat == pytest.approx(1.0 / (want * max(cflx, cfly, cflz)), rel=1e-15)


# ------------------------------------------------------------------ variant 2
def _kept_modes(fac, i, my=33):
    g = Grid(32, my, 32, Re=1000.0, alp=ALP, gam=GAM, Ly=LY, s=1.0)
    c = Cross(g, ydealias=2)
    j = np.arange(my)
    kept = []
    for m in range(my):
        c.vor, c.phi = g.zeros(), g.zeros()
        c.u00, c.w00 = np.zeros(my), np.zeros(my)
        c.vor[:, i, 3] = np.exp(2j * math.pi * m * j / my)
        c._dealias_lines(fac)
        n = np.linalg.norm(c.vor[:, i, 3])
        if n > 0.5 * math.sqrt(my):
            kept.append(m)
        else:
            assert n < 1e-12 * math.sqrt(my)
    return kept


# my = 33: my23 = 22, ny1 = 10, ny2 = 23; mx = 20, mx1 = 9; Lx = 2, Ly = 1.
# djdb = ny1/mx1 (-fac) Ly/Lx i; djda = ny1/mx1 (-fac -+ Lx/Ly) Ly/Lx i
# (- for fac <= 0, + for fac > 0); cut: ny1 + djdb <= j <= ny2 + djdb, or
# ny1 + djda < j < ny2 + djda.
CASES = [
    # fac = -0.5, i = 3: djdb = 10/9 * 0.75 = 0.833, djda = 10/9 * (-2.25) = -2.5
    #   cut 11..23 and 8..20 -> kept 0..7, 24..32
    (-0.5, 3, list(range(0, 8)) + list(range(24, 33))),
    # fac = 0.3, i = 9: djdb = 10/9 * (-1.35) = -1.5, djda = 10/9 * 7.65 = 8.5
    #   cut 9..21 and 19..31 -> kept 0..8, 32
    (0.3, 9, list(range(0, 9)) + [32]),
    # fac = -1.9, i = 5: djdb = 10/9 * 4.75 = 5.278, djda = 10/9 * (-0.25) = -0.278
    #   cut 16..28 and 10..22 -> kept 0..9, 29..32
    (-1.9, 5, list(range(0, 10)) + list(range(29, 33))),
    # fac = 0.3, i = 0: both windows unshifted; the inclusive one cuts 10..23
    #   -> kept 0..9, 24..32 (variant 1 keeps 10 and 23 as well)
    (0.3, 0, list(range(0, 10)) + list(range(24, 33))),
]


@pytest.mark.parametrize("fac,i,kept", CASES)
def test_variant2_windows(fac, i, kept):
    assert _kept_modes(fac, i) == kept
