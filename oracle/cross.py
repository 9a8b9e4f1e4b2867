"""RK3 time advance of the shear-periodic box (restates cross.f90).

``Cross`` holds the state (vor = omega_y, phi = lap v, u00 = U(y), w00 = W(y)),
the work fields and the scalars of modules ``running``/``bcs`` and runs
``cross`` (cross.f90:3-1101): low-storage SMR RK3 with iadd_mode=5 phase
maps, explicit (default) or implicit viscous terms, ``hvhg`` nonlinear
terms (cross.f90:1127-1765) and ``hist`` statistics (cross.f90:1773-2121).

Spectral layout: ``[y][kx][kz]`` complex; lines are ``f[:, i, k]``.
"""
import math
import numpy as np

from .grid import Grid
from .compact import YOps, modchol, stencil, solve
from . import fftxz

# cross.f90:18-24
GAMA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
ALPHA = (29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0)
BETA = (37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0)
IBETA = (160.0 / 37.0, 24.0 / 5.0, 6.0)
ZETA = (-17.0 / 60.0, -5.0 / 12.0, 0.0)
CA = (8.0 / 15.0, 2.0 / 3.0, 1.0)
TINY = 1e-13
NSTAT = 23
NSPEC = 8


def _div(a, b):
    """a / b with IEEE semantics (b = 0 gives inf, as the compiled reference
    does when a step is clipped to zero length at the etime stop)"""
    return a / b if b != 0.0 else math.copysign(math.inf, a)


def set_conj(f, nz):
    """kx=0 Hermitian symmetry f(0,k) = conj f(0,mz-k), k > nz
    (operators.f90:1382-1429); in place on a [y][kx][kz] array."""
    mz = f.shape[2]
    for k in range(nz + 1, mz):
        f[:, 0, k] = np.conj(f[:, 0, mz - k])
    return f


def moving_frame(g, f, m):
    """f(y,kx,kz) *= exp(i kx y m)  (operators.f90:86-145), in place."""
    th = (g.kx[None, :] * g.y[:, None]) * m
    f *= np.exp(1j * th)[:, :, None]
    return f


class Cross:
    def __init__(self, grid: Grid, explicit=1, nhist=1, nimag=10 ** 9,
                 readflag=1, uprim=0.1, vbulk=0.0, ydealias=0):
        self.g = grid
        self.ops = YOps(grid.my, grid.dy)
        self.explicit = int(explicit)
        self.nhist = int(nhist)
        self.nimag = int(nimag)
        self.readflag = int(readflag)
        self.uprim = float(uprim)
        self.vbulk = float(vbulk)
        if ydealias not in (0, 1, 2, 3):
            raise ValueError("iydealiasing must be 0, 1, 2 or 3")
        self.ydealias = int(ydealias)
        g = grid
        self.vor = g.zeros()
        self.phi = g.zeros()
        self.u00 = np.zeros(g.my)
        self.w00 = np.zeros(g.my)
        self.hv, self.hg = g.zeros(), g.zeros()
        self.vorwk, self.phiwk, self.dvordy = g.zeros(), g.zeros(), g.zeros()
        self.u00wk = np.zeros(g.my)
        self.w00wk = np.zeros(g.my)
        self.rf0u = np.zeros(g.my)
        self.rf0w = np.zeros(g.my)
        self.time = 0.0
        self.Deltat = 0.0
        self.dtr = 0.0
        self.xoffb = 0.0
        self.xofft = 0.0
        self.etime = 1e10
        self.dumptime = 1e10
        self.dtimag = 1e10
        self.istop = 0
        self.idump = 0
        self.umax = np.zeros(4)
        self.CFLv = 0.0
        self.history = []       # 24-double UNER records (the .cfbin rows)
        self.stats = np.zeros((NSTAT, g.my))
        self.stats2 = np.zeros((NSTAT, g.my))
        self.spec = np.zeros((NSPEC, g.kxn, g.mz))     # sp(nspec, 0:mx1, 0:mz1)
        self.spec2 = np.zeros((NSPEC, g.kxn, g.mz))
        self.nacum = 0
        self.total_dt = 0.0
        self.istep = 0
        self.last_H = None       # (H1, H2, H3) spectral planes of the last hvhg

    # ---------------------------------------------------------------- helpers
    def set_state(self, vor, phi, u00, w00, time=0.0, xoffb=0.0, xofft=0.0):
        self.vor = np.array(vor, dtype=np.complex128)
        self.phi = np.array(phi, dtype=np.complex128)
        self.u00 = np.array(u00, dtype=np.float64)
        self.w00 = np.array(w00, dtype=np.float64)
        self.time, self.xoffb, self.xofft = float(time), float(xoffb), float(xofft)

    def _shift(self, xoff):
        return np.exp(-self.g.xalp * xoff)

    def _lines_d(self, f, shp, shm, order):
        """derivyc (m=2) on every line of a [y][kx][kz] field; twist per kx"""
        fac = self.ops.twisted(order, shp)
        for nm in ("D", "Dm1", "La", "Lb", "Lc", "Le", "Lf", "Lg"):
            setattr(fac, nm, getattr(fac, nm)[:, :, None])
        if order == 1:
            r = stencil(self.ops.rh1, f, shp[None, :, None], shm[None, :, None],
                        skip_center=True)
        else:
            r = stencil(self.ops.rh2, f, shp[None, :, None], shm[None, :, None])
        return solve(fac, r)

    def _lines_helm(self, f, k2, shp, shm):
        """lapcdy (m=2) on every line except (0,0); k2 is (kxn, mz)"""
        g = self.g
        nl = g.kxn * g.mz
        t = np.repeat(shp, g.mz)
        k2l = np.array(k2, dtype=np.float64).reshape(-1)
        if k2l[0] == 0.0:
            k2l[0] = 1.0   # the (0,0) line is never solved; keep it regular
        fac = modchol(self.ops.lap_matrix(k2l), g.my, t)
        fl = f.reshape(g.my, nl)
        r = stencil(self.ops.pr2, fl, t[None, :], np.repeat(shm, g.mz)[None, :])
        out = solve(fac, r).reshape(g.my, g.kxn, g.mz)
        return out

    def reset_statistics(self):
        """readwrite.f90:1322-1334: zero the statistics and spectra after a dump"""
        self.stats[:] = 0.0
        self.stats2[:] = 0.0
        self.spec[:] = 0.0
        self.spec2[:] = 0.0
        self.nacum = 0
        self.total_dt = 0.0

    def getv(self, f, iopt):
        """getv (operators.f90:735-793) with the twist of the current offsets:
        iopt=+1 returns v = lapcdy(phi=f); iopt=-1 returns phi = d2v/dy2 -
        k2 v (visc3d with scal=1, diffy_shear_hepta.f90:298-318) for v=f.
        The (0,0) line of the result is zero (the reference leaves it
        untouched)."""
        g = self.g
        sht, shb = self._shift(self.xofft), self._shift(self.xoffb)
        k2 = g.alp2[:, None] + g.gam2[None, :]
        if iopt == 1:
            out = self._lines_helm(f, k2, sht, shb)
        elif iopt == -1:
            out = self._lines_d(f, sht, shb, 2) - k2[None, :, :] * f
        else:
            raise ValueError("iopt must be +1 or -1")
        out[:, 0, 0] = 0.0
        return out

    def energy(self):
        """energy (operators.f90:870-924): sums of |vor|^2 and |phi|^2 with
        kx > 0 counted twice, and of U^2, W^2."""
        w = np.ones(self.g.kxn)
        w[1:] = 2.0
        e1 = float(np.sum(w[None, :, None] * np.abs(self.vor) ** 2))
        e2 = float(np.sum(w[None, :, None] * np.abs(self.phi) ** 2))
        return np.array([e1, e2, float(np.sum(self.u00 ** 2)), float(np.sum(self.w00 ** 2))])

    # ------------------------------------------------------------------- hvhg
    def derived(self, phic, ome2c, rhgc, rhvc, ome1c, dudy, dwdy):
        """KMM reconstruction at every plane (cross.f90:1221-1276).
        Returns u1..u3 (u, v, w) and o1..o3 (omega_x, omega_y, omega_z)."""
        g = self.g
        xg = g.xgam[None, None, :]
        xa = g.xalp[None, :, None]
        kkw = (g.alp2[:, None] + g.gam2[None, :])[None, :, :]
        u1, u3, o1, o3 = (np.zeros_like(phic) for _ in range(4))
        with np.errstate(divide="ignore", invalid="ignore"):
            o3[:, 0, 1:] = -ome1c[:, 0, 1:] / xg[:, 0, 1:]
            o1[:, 0, 1:] = -phic[:, 0, 1:] / xg[:, 0, 1:]
            u3[:, 0, 1:] = -rhvc[:, 0, 1:] / xg[:, 0, 1:]
            u1[:, 0, 1:] = ome2c[:, 0, 1:] / xg[:, 0, 1:]
            o3[:, 1:, :] = (ome1c[:, 1:, :] * xg - phic[:, 1:, :] * xa[:, 1:, :]) / kkw[:, 1:, :]
            o1[:, 1:, :] = (ome1c[:, 1:, :] * xa[:, 1:, :] + phic[:, 1:, :] * xg) / kkw[:, 1:, :]
            u3[:, 1:, :] = (rhvc[:, 1:, :] * xg + ome2c[:, 1:, :] * xa[:, 1:, :]) / kkw[:, 1:, :]
            u1[:, 1:, :] = (rhvc[:, 1:, :] * xa[:, 1:, :] - ome2c[:, 1:, :] * xg) / kkw[:, 1:, :]
        u1[:, 0, 0] = self.u00
        u3[:, 0, 0] = self.w00
        o3[:, 0, 0] = -dudy
        o1[:, 0, 0] = dwdy
        u2 = rhgc.copy()
        u2[:, 0, 0] = self.vbulk
        o2 = ome2c.copy()
        return u1, u2, u3, o1, o2, o3

    def hvhg(self, rkstep, ihist):
        g = self.g
        one = np.ones(1)
        dudy = self.ops.deriv_real(self.u00, 1)[:, 0]
        dwdy = self.ops.deriv_real(self.w00, 1)[:, 0]
        phic = set_conj(self.phi.copy(), g.nz)
        ome2c = set_conj(self.vor.copy(), g.nz)
        u1, u2, u3, o1, o2, o3 = self.derived(phic, ome2c, self.hg, self.hv,
                                              self.dvordy, dudy, dwdy)
        ener = np.zeros(10)
        ener2 = np.zeros(4)
        if ihist:
            self._hist_f(u1, u2, u3, o1, o2, o3, self.hv, ener, ener2)
        if self.istep == 1 and self.readflag == 0:
            umaxl = np.full(4, 1.0 if abs(g.s) > 0.5 else self.uprim)
        else:
            umaxl = np.zeros(4)
        mgx, mgz = g.mgalx, g.mgalz
        U = fftxz.to_phys(u1, mgx, mgz)
        V = fftxz.to_phys(u2, mgx, mgz)
        W = fftxz.to_phys(u3, mgx, mgz)
        O1 = fftxz.to_phys(o1, mgx, mgz)
        O2 = fftxz.to_phys(o2, mgx, mgz)
        O3 = fftxz.to_phys(o3, mgx, mgz)
        if ihist:
            # cross.f90:2030-2045 ('p')
            self.stats[16] += np.sum(V * U ** 2, axis=(1, 2))
            self.stats[17] += np.sum(V * W ** 2, axis=(1, 2))
            self.stats[18] += np.sum(V ** 3, axis=(1, 2))
            self.stats2[16] += np.sum(V * U ** 2, axis=(1, 2)) * self.Deltat
            self.stats2[17] += np.sum(V * W ** 2, axis=(1, 2)) * self.Deltat
            self.stats2[18] += np.sum(V ** 3, axis=(1, 2)) * self.Deltat
        if rkstep == 1:
            umaxl[0] = max(umaxl[0], np.max(np.abs(U)))
            umaxl[1] = max(umaxl[1], np.max(np.abs(U + g.s * g.y[:, None, None])))
            umaxl[2] = max(umaxl[2], np.max(np.abs(V)))
            umaxl[3] = max(umaxl[3], np.max(np.abs(W)))
        # u x omega (cross.f90:1453-1462)
        c2 = 2.0 * g.chi
        H1 = V * (O3 + c2) - W * O2
        H2 = W * O1 - U * (O3 + c2)
        H3 = U * O2 - V * O1
        h1 = fftxz.to_spec(H1, g.kxn, g.mz)
        h2 = fftxz.to_spec(H2, g.kxn, g.mz)
        h3 = fftxz.to_spec(H3, g.kxn, g.mz)
        self.last_H = (h1, h2, h3)
        self.rf0u = h1[:, 0, 0].real.copy()
        self.rf0w = h3[:, 0, 0].real.copy()
        xa = g.xalp[None, :, None]
        xg = g.xgam[None, None, :]
        v = self.hg
        self.dvordy = xa * h1 + xg * h3
        self.hg = -xa * h3 + xg * (h1 - g.s * v)
        self.hv = -h2 * (g.alp2[:, None] + g.gam2[None, :])[None]
        if rkstep == 1:
            self._new_dt(umaxl)
        if ihist:
            self._hist_w(ener, ener2)

    def _hist_f(self, u1, u2, u3, o1, o2, o3, dv, ener, ener2):
        g = self.g
        w = np.full(g.kxn, 2.0)
        w[0] = 1.0
        wt = w[None, :, None]

        def psum(a, b):
            return np.sum((a * np.conj(b)).real * wt, axis=(1, 2))

        up, vp, wp = psum(u1, u1), psum(u2, u2), psum(u3, u3)
        uvr, uwr, vwr = psum(u1, u2), psum(u1, u3), psum(u2, u3)
        o1p, o2p, o3p = psum(o1, o1), psum(o2, o2), psum(o3, o3)
        o12p = psum(o1, o2)
        st, st2, dt = self.stats, self.stats2, self.Deltat
        for idx, val in zip((3, 4, 5, 6, 7, 8, 12, 13, 14),
                            (up, vp, wp, uvr, uwr, vwr, o1p, o2p, o3p)):
            st[idx] += val
            st2[idx] += val * dt
        for idx, f in zip((0, 1, 2, 9, 10, 11), (u1, u2, u3, o1, o2, o3)):
            st[idx] += f[:, 0, 0].real
            st2[idx] += f[:, 0, 0].real * dt
        ener[0:9] += [up.sum(), vp.sum(), wp.sum(), uvr.sum(), uwr.sum(), vwr.sum(),
                      o1p.sum(), o2p.sum(), o3p.sum()]
        # spectra of the same products, summed over the planes
        # (cross.f90:1970-2016; o3 without the shear correction)
        c2 = lambda a, b: (a * np.conj(b)).real
        spec = np.stack([c2(u1, u1), c2(u2, u2), c2(u3, u3), c2(u1, u2),
                         c2(o1, o1), c2(o2, o2), c2(o3, o3), c2(o1, o2)])
        spec = np.sum(spec * wt[None], axis=1)
        self.spec += spec
        self.spec2 += spec * dt
        o3s = o3.copy()
        o3s[:, 0, 0] -= g.s
        o300 = o3s[:, 0, 0].real
        pk = -o300 * (-uvr) + o300 * o300 * (1.0 / g.Re)
        st[19] += pk
        st[20] += o12p
        st2[19] += pk * dt
        st2[20] += o12p * dt
        ener2[0] += pk.sum()
        ener2[1] += o12p.sum()
        # dissipation (cross.f90:1944-1966)
        k2 = (g.alp2[:, None] + g.gam2[None, :])[None]
        a2 = lambda a: (a * np.conj(a)).real
        uu = a2(u1) + a2(u2) + a2(u3)
        eps = np.sum(g.gam2[None, :] * uu[:, 0, :], axis=1) \
            + np.sum(a2(dv[:, 0, :]) + a2(o3s[:, 0, :]), axis=1)
        eps = eps + 2.0 * np.sum(k2[:, 1:, :] * uu[:, 1:, :] + a2(dv[:, 1:, :]), axis=(1, 2))
        xg = g.xgam[None, None, :]
        xa = g.xalp[None, :, None]
        eps = eps + np.sum(a2(o1[:, 0, :] + xg[:, 0, :] * u2[:, 0, :]), axis=1)
        eps = eps + 2.0 * np.sum(a2(o1[:, 1:, :] + xg * u2[:, 1:, :])
                                 + a2(o3s[:, 1:, :] - xa[:, 1:, :] * u2[:, 1:, :]), axis=(1, 2))
        st[15] += eps
        st2[15] += eps * dt
        ener[9] += eps.sum()

    def _hist_w(self, ener, ener2):
        g = self.g
        self.nacum += 1
        self.total_dt += self.Deltat
        my = float(g.my)
        uner = np.empty(10)
        uner[0:3] = np.sqrt(np.abs(ener[0:3] / my))
        uner[3:6] = ener[3:6] / my
        uner[6:9] = np.sqrt(np.abs(ener[6:9] / my))
        uner[9] = ener[9] / my
        uner2 = ener2 / my
        rec = np.concatenate([[self.time, self.Deltat], uner, [g.CFL, self.CFLv],
                              self.umax[1:4], [self.xoffb, self.xofft, self.umax[0]], uner2])
        self.history.append(rec)
        if np.isnan(uner[0]):
            self.istop = 1
        if uner[1] / g.Lz < 1e-5:
            self.istop = 1

    def _new_dt(self, umaxl):
        """cross.f90:1598-1712 (adaptive dt, iadd_mode=5, no LES)"""
        g = self.g
        self.umax = umaxl.copy()
        dt = 1.0 / max(self.umax[0] * g.cflx, self.umax[2] * g.cfly, self.umax[3] * g.cflz)
        self.CFLv = g.CFL * g.cflvv * dt
        if self.CFLv > g.CFL and self.explicit == 1:
            dt = 1.0 / g.cflvv
        self.Deltat = dt
        self.dtr = _div(g.Re, self.Deltat)
        if self.Deltat < 1e-10:
            self.istop = 1
        if self.time + self.Deltat > self.dumptime:
            self.Deltat = abs(self.dumptime - self.time)
            self.dtr = _div(g.Re, self.Deltat)
            if self.dumptime > self.etime - TINY:
                self.istop = 1
            self.dumptime += self.dtimag
            self.idump = 1
        elif self.time + self.Deltat > self.etime:
            self.Deltat = abs(self.etime - self.time)
            self.dtr = _div(g.Re, self.Deltat)
            self.istop = 1

    def _wrap(self, xoff):
        """(xoff/Lx - int(xoff/Lx)) Lx, Fortran int() truncating toward zero
        (cross.f90:349, 734)"""
        Lx = self.g.Lx
        return (xoff / Lx - math.trunc(xoff / Lx)) * Lx

    def _dealias_y(self, xwkb):
        """iydealiasing=1, 2 at the top of a substep (cross.f90:459-475):
        the state is taken to the moving frame exp(i kx y fac) with fac from
        the bottom offset, cut in y-Fourier space by dealiasing_y
        (operators.f90:352-486), and taken back."""
        g = self.g
        Lx, Ly = g.Lx, g.Ly
        fac = self._wrap(xwkb) / Ly
        if fac < -Lx / Ly * 0.5:
            fac = Lx / Ly + fac
        moving_frame(g, self.vor, fac)
        moving_frame(g, self.phi, fac)
        self._dealias_lines(fac)
        moving_frame(g, self.vor, -fac)
        moving_frame(g, self.phi, -fac)

    def _dealias_lines(self, fac):
        """dealiasing_y (operators.f90:352-486), variants 1 and 2.  rfty and
        cfty are FFTW transforms scaled by 1/n forward (rftdoublew.F:133-158,
        cftdoublew.F:151-201): numpy's forward/inverse pair has the same
        product.  U, W: the real-to-complex coefficients from real index
        my23 = 2 (my/3) on are zeroed, i.e. modes k >= my/3.  vor, phi:
        coefficient j of a line is zeroed inside the window of its kx index."""
        g = self.g
        my, Lx, Ly = g.my, g.Lx, g.Ly
        kcut = (2 * (my // 3)) // 2
        for name in ("u00", "w00"):
            c = np.fft.rfft(getattr(self, name))
            c[kcut:] = 0.0
            setattr(self, name, np.fft.irfft(c, n=my))
        my23 = 2 * (my // 3)
        ny1 = ((my23 - 1) - 1) // 2
        ny2 = my - ny1
        mx1 = g.kxn - 1
        j = np.arange(my)
        rows = (0,) if self.ydealias == 1 else range(g.kxn)
        for i in rows:
            if self.ydealias == 1 or fac <= 0.0:
                djda = ny1 / mx1 * (-fac - Lx / Ly) * Ly / Lx * i
            else:
                djda = ny1 / mx1 * (-fac + Lx / Ly) * Ly / Lx * i
            djdb = ny1 / mx1 * (-fac) * Ly / Lx * i
            if self.ydealias == 1:
                cut = (j > ny1 + djdb) & (j < ny2 + djdb)
            else:
                cut = ((j >= ny1 + djdb) & (j <= ny2 + djdb)) | \
                      ((j > ny1 + djda) & (j < ny2 + djda))
            for f in (self.vor, self.phi):
                c = np.fft.fft(f[:, i, :], axis=0)
                c[cut, :] = 0.0
                f[:, i, :] = np.fft.ifft(c, axis=0)

    def _filter_state(self, sht, shb):
        """iydealiasing=3: compact filter of vor, phi (twisted) and U, W
        (cross.f90:214-225, 476-488; filterc filter_shear_hepta.f90:128-165)"""
        g = self.g
        ops = self.ops
        self.u00 = ops.filter_real(self.u00)[:, 0]
        self.w00 = ops.filter_real(self.w00)[:, 0]
        t = np.repeat(sht, g.mz)
        tm = np.repeat(shb, g.mz)
        for name in ("vor", "phi"):
            f = getattr(self, name).reshape(g.my, -1)
            setattr(self, name, ops.filter(f, t, tm).reshape(g.my, g.kxn, g.mz))

    # ------------------------------------------------------------------ cross
    def run(self, nstep, on_dump=None):
        """cross.f90:3-1101. Returns the number of completed steps."""
        g = self.g
        ops = self.ops
        s = g.s
        k2 = g.k2()
        nz00 = np.ones((g.kxn, g.mz), dtype=bool)
        nz00[0, 0] = False
        m00 = nz00[None]
        ihist = 1
        irun = 0
        self.istop = 0
        self.idump = 0
        set_conj(self.vor, g.nz)
        set_conj(self.phi, g.nz)
        if self.ydealias == 3:
            # before the step loop, with the twist of the current offsets
            self._filter_state(self._shift(self.xofft), self._shift(self.xoffb))
        done = 0
        for istep in range(1, nstep + 1):
            self.istep = istep
            if (istep - 1) % self.nhist == 0:
                ihist = 1
            if self.ydealias == 2:
                self.xoffb = self._wrap(self.xoffb)
                self.xofft = -self.xoffb
            xwkt, xwkb = self.xofft, self.xoffb
            if istep - 1 != 1:
                if (((istep - 1) % self.nimag == 0 and istep != 1)
                        or self.istop == 1 or self.idump == 1):
                    if on_dump is not None:
                        on_dump(self)
                    if self.istop == 1:
                        break
                    self.idump = 0
            u00b, u00t = s * (-g.Ly), s * g.Ly
            if abs(s) > 0.001:
                self.shb = self._shift(xwkb)
                self.sht = self._shift(xwkt)
            else:
                self.shb = getattr(self, "shb", np.ones(g.kxn, dtype=np.complex128))
                self.sht = getattr(self, "sht", np.ones(g.kxn, dtype=np.complex128))
            shwkt, shwkb = self.sht.copy(), self.shb.copy()
            if irun == 0:
                irun = 1
                v = self._lines_helm(self.phi, k2, self.sht, self.shb)
                self.hg = np.where(m00, v, self.hg)
            for rk in (1, 2, 3):
                ir = rk - 1
                self.dvordy = self._lines_d(self.vor, shwkt, shwkb, 1)
                self.hv = self._lines_d(self.hg, shwkt, shwkb, 1)
                if self.ydealias == 3:
                    self._filter_state(shwkt, shwkb)
                elif self.ydealias in (1, 2):
                    self._dealias_y(xwkb)
                self.hvhg(rk, ihist)
                if ihist and rk == 1:
                    ihist = 0
                self.dvordy = self._lines_d(self.dvordy, shwkt, shwkb, 1)
                self.hv = self.hv - self.dvordy
                r1 = CA[ir] * self.Deltat
                dtr1 = self.dtr * IBETA[ir]
                if self.explicit == 1:
                    fac = 1.0 / g.Re
                    d2v = self._lines_d(self.vor, shwkt, shwkb, 2)
                    d2p = self._lines_d(self.phi, shwkt, shwkb, 2)
                    self.hg = np.where(m00, self.hg + fac * (-k2[None] * self.vor + d2v), self.hg)
                    self.hv = np.where(m00, self.hv + fac * (-k2[None] * self.phi + d2p), self.hv)
                    self.rf0u = self.rf0u + fac * (-0.0 * self.u00 + ops.deriv_real(self.u00, 2)[:, 0])
                    self.rf0w = self.rf0w + fac * (-0.0 * self.w00 + ops.deriv_real(self.w00, 2)[:, 0])
                    if rk == 1:
                        self.vorwk = GAMA[ir] * self.hg
                        self.phiwk = GAMA[ir] * self.hv
                        self.u00wk = GAMA[ir] * self.rf0u
                        self.w00wk = GAMA[ir] * self.rf0w
                    else:
                        self.vorwk = self.vorwk + GAMA[ir] * self.hg
                        self.phiwk = self.phiwk + GAMA[ir] * self.hv
                        self.u00wk = self.u00wk + GAMA[ir] * self.rf0u
                        self.w00wk = self.w00wk + GAMA[ir] * self.rf0w
                    self.vor = self.vor + self.Deltat * self.vorwk
                    self.phi = self.phi + self.Deltat * self.phiwk
                    self.u00 = self.u00 + self.Deltat * self.u00wk
                    self.w00 = self.w00 + self.Deltat * self.w00wk
                    if abs(s) > 0.01:
                        xwkb = self.xoffb + r1 * u00b
                        xwkt = self.xofft + r1 * u00t
                        if self.ydealias == 2:
                            xwkb = self._wrap(xwkb)
                            xwkt = -xwkb
                        shwkb = self._shift(xwkb)
                        shwkt = self._shift(xwkt)
                    if rk != 3:
                        self.phiwk = self.hv * ZETA[ir]
                        self.vorwk = self.hg * ZETA[ir]
                        if rk == 1:
                            f1 = -GAMA[0] * self.Deltat * s
                        else:
                            f1 = -(GAMA[1] + ZETA[0]) * self.Deltat * s
                        moving_frame(g, self.vorwk, f1)
                        moving_frame(g, self.phiwk, f1)
                        self.u00wk = self.rf0u * ZETA[ir]
                        self.w00wk = self.rf0w * ZETA[ir]
                    f2 = -s * r1 if rk == 1 else -s * (CA[ir] - CA[ir - 1]) * self.Deltat
                    moving_frame(g, self.vor, f2)
                    moving_frame(g, self.phi, f2)
                    v = self._lines_helm(self.phi, k2, shwkt, shwkb)
                    self.hg = np.where(m00, v, self.hg)
                else:
                    if rk == 1:
                        self.vorwk = GAMA[ir] * self.hg
                        self.phiwk = GAMA[ir] * self.hv
                    else:
                        self.vorwk = self.vorwk + GAMA[ir] * self.hg
                        self.phiwk = self.phiwk + GAMA[ir] * self.hv
                    fac = ALPHA[ir] * (1.0 / g.Re)
                    d2v = self._lines_d(self.vor, shwkt, shwkb, 2)
                    d2p = self._lines_d(self.phi, shwkt, shwkb, 2)
                    self.vorwk = np.where(m00, self.vorwk + fac * (-k2[None] * self.vor + d2v), self.vorwk)
                    self.phiwk = np.where(m00, self.phiwk + fac * (-k2[None] * self.phi + d2p), self.phiwk)
                    if rk == 1:
                        self.u00wk = GAMA[ir] * self.rf0u
                        self.w00wk = GAMA[ir] * self.rf0w
                    else:
                        self.u00wk = self.u00wk + GAMA[ir] * self.rf0u
                        self.w00wk = self.w00wk + GAMA[ir] * self.rf0w
                    self.u00wk = self.u00wk + fac * (-0.0 * self.u00 + ops.deriv_real(self.u00, 2)[:, 0])
                    self.w00wk = self.w00wk + fac * (-0.0 * self.w00 + ops.deriv_real(self.w00, 2)[:, 0])
                    self.vorwk = self.vorwk + _div(1.0, self.Deltat) * self.vor
                    self.phiwk = self.phiwk + _div(1.0, self.Deltat) * self.phi
                    fac = -g.Re * IBETA[ir]
                    self.vorwk = fac * self.vorwk
                    self.phiwk = fac * self.phiwk
                    self.u00wk = fac * (self.u00wk + _div(1.0, self.Deltat) * self.u00)
                    self.w00wk = fac * (self.w00wk + _div(1.0, self.Deltat) * self.w00)
                    if abs(s) > 0.5:
                        xwkb = self.xoffb + r1 * u00b
                        xwkt = self.xofft + r1 * u00t
                        shwkb = self._shift(xwkb)
                        shwkt = self._shift(xwkt)
                    f2 = -s * r1 if rk == 1 else -s * (CA[ir] - CA[ir - 1]) * self.Deltat
                    moving_frame(g, self.vorwk, f2)
                    moving_frame(g, self.phiwk, f2)
                    k2i = dtr1 + k2
                    nv = self._lines_helm(self.vorwk, k2i, shwkt, shwkb)
                    npi = self._lines_helm(self.phiwk, k2i, shwkt, shwkb)
                    self.vor = np.where(m00, nv, self.vor)
                    self.phi = np.where(m00, npi, self.phi)
                    if rk != 3:
                        self.vorwk = ZETA[ir] * self.hg
                        self.phiwk = ZETA[ir] * self.hv
                        if rk == 1:
                            f1 = -GAMA[0] * self.Deltat * s
                        else:
                            f1 = -(GAMA[1] + ZETA[0]) * self.Deltat * s
                        moving_frame(g, self.vorwk, f1)
                        moving_frame(g, self.phiwk, f1)
                    v = self._lines_helm(self.phi, k2, shwkt, shwkb)
                    self.hg = np.where(m00, v, self.hg)
                    self.u00 = ops.helmholtz_real(self.u00wk, dtr1)[:, 0]
                    self.w00 = ops.helmholtz_real(self.w00wk, dtr1)[:, 0]
                    if rk != 3:
                        self.u00wk = ZETA[ir] * self.rf0u
                        self.w00wk = ZETA[ir] * self.rf0w
            self.xoffb, self.xofft = xwkb, xwkt
            self.sht, self.shb = shwkt, shwkb
            self.time = self.time + self.Deltat
            done += 1
        return done
