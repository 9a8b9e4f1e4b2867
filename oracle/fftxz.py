"""x-z plane transforms with the 3/2 dealiasing map (fou3D_thread.f:1-161).

Spectral plane: complex ``(kxn, mz)`` indexed [kx][kz].
Physical plane: real ``(mgalx, mgalz)`` indexed [x][z].

Backward (spectral -> physical, fourxz iopt=+1): fill00 places kz = 0..nz at
rows 0..nz and the negative kz (indices nz+1..mz-1) at rows mgalz-mz+k;
everything else is zero; then an unnormalised complex inverse DFT along z and
a complex-to-real inverse DFT along x (the imaginary part of the kx=0 element
of every row is ignored, as a c2r transform does).

Forward (physical -> spectral, fourxz iopt=-1): real-to-complex DFT along x,
complex DFT along z, scaled by 1/(mgalx*mgalz), then pack00 keeps the same
modes.
"""
import numpy as np


def kz_rows(mz, mgalz):
    """row (0..mgalz-1) of the padded z-transform for each kz index (fill00)"""
    nz = (mz - 1) // 2
    k = np.arange(mz)
    return np.where(k <= nz, k, k - mz + mgalz)


def to_phys(F, mgalx, mgalz):
    """F: (..., kxn, mz) complex -> (..., mgalx, mgalz) real"""
    kxn, mz = F.shape[-2:]
    pad = np.zeros(F.shape[:-2] + (mgalx // 2 + 1, mgalz), dtype=np.complex128)
    pad[..., :kxn, kz_rows(mz, mgalz)] = F
    G = np.fft.ifft(pad, axis=-1, norm="forward")
    return np.fft.irfft(G, n=mgalx, axis=-2, norm="forward")


def to_spec(f, kxn, mz):
    """f: (..., mgalx, mgalz) real -> (..., kxn, mz) complex"""
    mgalx, mgalz = f.shape[-2:]
    H = np.fft.rfft(f, axis=-2)
    H = np.fft.fft(H, axis=-1)
    H = H * (1.0 / (mgalx * mgalz))
    return H[..., :kxn, :][..., kz_rows(mz, mgalz)]
