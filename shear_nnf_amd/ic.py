"""Seeded synthetic initial fields (omega_y, phi, U, W) for the shear box.

A random solenoidal velocity field with spectrum E(k) ~ k^4 exp(-2 (k/k0)^2)
(k in units of the z fundamental 2 pi / Lz, k0 = 4), uniform random phases
from a counter-based SplitMix64 generator, Rogallo/Craya construction
(the same construction as getini_tophat, initcr.f90:954-1071), Hermitian
symmetry on the kx = 0 plane, normalised to u'_rms = uprim
(u'_rms^2 = <u^2 + v^2 + w^2>/3 including the mean profiles).

Returned arrays use the device layout [y][kx][kz]: vor, phi complex
(my, kxn, mz); u00, w00 real (my,).  At t = 0 the shear offset is zero,
so the field is periodic in y and is built from y-Fourier modes
|n| <= ny (initcr.f90:205-210 y-dealiasing bound).
"""
import math
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed, idx):
    """uniform doubles in [0,1) for counter values idx (uint64 array)"""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * _G + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def grid_sizes(mgalx, my, mgalz):
    mx = 2 * (mgalx // 3)
    mz = 2 * (mgalz // 3) - 1
    return mx // 2, mz, (mz - 1) // 2


def make_ic(mgalx, my, mgalz, alp, gam, Ly, seed, uprim=0.1, k0=4.0):
    kxn, mz, nz = grid_sizes(mgalx, my, mgalz)
    my23 = 2 * (my // 3)
    ny = ((my23 - 1) - 1) // 2
    bet = 2.0 * math.pi / Ly
    Lz = 2.0 * math.pi / gam
    nvals = np.arange(-ny, ny + 1)
    nn = nvals.size
    kx = alp * np.arange(kxn)
    kz = np.array([gam * k if k <= nz else -gam * (mz - k) for k in range(mz)])
    KX = kx[:, None, None]
    KY = (bet * nvals)[None, :, None]
    KZ = kz[None, None, :]
    KX, KY, KZ = np.broadcast_arrays(KX, KY, KZ)
    kmag = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    kn = kmag * Lz / (2.0 * math.pi)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = kn ** 4 * np.exp(-2.0 * (kn / k0) ** 2)
        A = np.where(kn > 0, np.sqrt(E / (4.0 * math.pi * kn * kn)), 0.0)
    idx = (np.arange(kxn)[:, None, None] * nn + np.arange(nn)[None, :, None]) * mz \
        + np.arange(mz)[None, None, :]
    idx = idx.astype(np.uint64) * np.uint64(3)
    th1 = 2.0 * math.pi * splitmix64_uniform(seed, idx)
    th2 = 2.0 * math.pi * splitmix64_uniform(seed, idx + np.uint64(1))
    ph = 2.0 * math.pi * splitmix64_uniform(seed, idx + np.uint64(2))
    a = A * np.cos(ph) * np.exp(1j * th1)
    b = A * np.sin(ph) * np.exp(1j * th2)
    kxy = np.sqrt(KX * KX + KY * KY)
    with np.errstate(divide="ignore", invalid="ignore"):
        big = kxy > 1e-12
        e1 = [np.where(big, KY / kxy, 1.0), np.where(big, -KX / kxy, 0.0), np.zeros_like(KX)]
        e2 = [np.where(big, KX * KZ / (kxy * kmag), 0.0),
              np.where(big, KY * KZ / (kxy * kmag), 1.0),
              np.where(big, -kxy / kmag, 0.0)]
    u = a * e1[0] + b * e2[0]
    v = a * e1[1] + b * e2[1]
    w = a * e1[2] + b * e2[2]
    # kx = 0 plane: keep the half (kz > 0) or (kz = 0, n > 0), mirror the rest
    for f in (u, v, w):
        f[0, :, 0][nvals <= 0] = 0.0
        f[0, :, nz + 1:] = 0.0
        for k in range(1, nz + 1):
            f[0, ::-1, mz - k] = np.conj(f[0, :, k])
        half = f[0, :, 0].copy()
        f[0, :, 0] = half + np.conj(half[::-1])
    # energy, Parseval: kx > 0 modes count twice
    wgt = np.full(kxn, 2.0)
    wgt[0] = 1.0
    etot = np.sum(wgt[:, None, None] * (np.abs(u) ** 2 + np.abs(v) ** 2 + np.abs(w) ** 2))
    scale = uprim / math.sqrt(etot / 3.0) if etot > 0 else 0.0
    u, v, w = u * scale, v * scale, w * scale
    omy = 1j * KZ * u - 1j * KX * w
    phs = -(kmag * kmag) * v
    y = np.arange(my) * (Ly / my) - 0.5 * Ly
    Ey = np.exp(1j * bet * np.outer(y, nvals))         # (my, nn)
    vor = (Ey @ omy.transpose(1, 0, 2).reshape(nn, kxn * mz)).reshape(my, kxn, mz)
    phi = (Ey @ phs.transpose(1, 0, 2).reshape(nn, kxn * mz)).reshape(my, kxn, mz)
    vor[:, 0, 0] = 0.0
    phi[:, 0, 0] = 0.0
    u00 = (Ey @ u[0, :, 0]).real
    w00 = (Ey @ w[0, :, 0]).real
    return vor, phi, u00, w00
