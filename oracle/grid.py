"""Grid sizes, wavenumbers and CFL constants of the shear-periodic box.

Follows initcr.f90:186-216 (sizes) and initcr.f90:443-519 (set_xyz_dy).
Spectral arrays in the oracle use the layout ``[y][kx][kz]`` (kz fastest),
the same as the device layout, so a y-line is ``f[:, i, k]``.
"""
import math
import numpy as np


class Grid:
    def __init__(self, mgalx, my, mgalz, Re, alp, gam, Ly, s=1.0, chi=0.0,
                 CFL=0.5):
        self.mgalx, self.my, self.mgalz = int(mgalx), int(my), int(mgalz)
        self.Re, self.alp, self.gam, self.Ly = float(Re), float(alp), float(gam), float(Ly)
        self.s, self.chi, self.CFL = float(s), float(chi), float(CFL)
        # initcr.f90:190-203
        self.mx = 2 * (self.mgalx // 3)
        self.mz = 2 * (self.mgalz // 3) - 1
        self.kxn = self.mx // 2            # complex kx modes 0..mx1
        self.nz = (self.mz - 1) // 2       # kz = 0..nz positive, nz+1..mz-1 negative
        mz = self.mz
        # initcr.f90:455-472: xalp = i*alp*i, xgam = i*gam*k (negative half)
        kxv = self.alp * np.arange(self.kxn, dtype=np.float64)
        kzv = np.empty(mz)
        for k in range(mz):
            kzv[k] = self.gam * k if k <= self.nz else -self.gam * (mz - k)
        self.kx = kxv
        self.kz = kzv
        self.xalp = 1j * kxv
        self.xgam = 1j * kzv
        self.alp2 = kxv * kxv
        self.gam2 = kzv * kzv
        # initcr.f90:474-495
        pi = math.pi
        self.dx = 2 * pi / self.alp / self.mx
        self.dz = 2 * pi / self.gam / self.mz
        self.dy = self.Ly / self.my
        self.y = np.arange(self.my) * self.dy - 0.5 * self.Ly
        self.cflx = pi / self.CFL / self.dx
        self.cflz = pi / self.CFL / self.dz
        self.cfly = 2.5 / self.CFL / self.dy
        self.cflvv = 1.0 / self.Re / self.CFL * max(pi * pi / self.dx / self.dx,
                                                    6.25 / self.dy / self.dy,
                                                    pi * pi / self.dz / self.dz)
        self.Lx = 2 * pi / self.alp
        self.Lz = 2 * pi / self.gam

    @property
    def nc(self):
        """complex coefficients per spectral field"""
        return self.kxn * self.mz * self.my

    def k2(self):
        """kx^2 + kz^2 per (i, k), shape (kxn, mz)"""
        return self.alp2[:, None] + self.gam2[None, :]

    def zeros(self):
        return np.zeros((self.my, self.kxn, self.mz), dtype=np.complex128)
