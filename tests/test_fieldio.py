"""Restart files in the reference's format (escru readwrite.f90:1336-1393,
getfil readwrite.f90:72-447): the record layout written by escru, a write ->
read round trip, and getfil's resolution change (kx and both kz halves
copied up to the smaller size, the rest zero)."""
import math
import struct

import numpy as np
import pytest

from shear_nnf_amd import fieldio
from tests.boxes import ALP, GAM, LY, box_ic


def _write(tmp_path, grid=(48, 49, 48), **kw):
    vor, phi, u00, w00 = box_ic(*grid)
    p = tmp_path / "box.0001.bin"
    fieldio.write_restart(p, vor, phi, u00, w00, mgalx=grid[0], mgalz=grid[2], time=1.25,
                          Re=1000.0, alp=ALP, gam=GAM, Ly=LY, xoffb=-0.7, xofft=0.7,
                          Deltat=3e-3, **kw)
    return p, (vor, phi, u00, w00)


def test_record_layout(tmp_path):
    p, (vor, _, _, _) = _write(tmp_path)
    my, kxn, mz = vor.shape
    mx = 2 * kxn
    data = p.read_bytes()
    lengths, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        assert struct.unpack_from("<i", data, pos + 4 + n)[0] == n
        lengths.append(n)
        pos += n + 8
    assert lengths == [16, 7 * 8 + 3 * 4, 8 * my, 32, 16 * my] + [16 * mx * mz] * my + \
        [32, 12, 4 * fieldio.NOPT, 8 * fieldio.NPARAMS]
    # the first plane record interleaves vor and phi element by element
    off = 4 * 5 + sum(lengths[:5]) + 4 * 5 + 4
    first = np.frombuffer(data[off:off + 32], "<f8")
    assert first[0] == vor[0, 0, 0].real and first[2] == vor[0, 0, 0].imag
    hdr = struct.unpack_from("<4i", data, 4)
    assert hdr == (48, my, 48, 1)


def test_round_trip(tmp_path):
    p, (vor, phi, u00, w00) = _write(tmp_path)
    v2, p2, u2, w2, h = fieldio.read_restart(p)
    assert np.array_equal(v2, vor) and np.array_equal(p2, phi)
    assert np.array_equal(u2, u00) and np.array_equal(w2, w00)
    assert (h["mgalx"], h["my"], h["mgalz"]) == (48, 49, 48)
    assert h["time"] == 1.25 and h["xoffb"] == -0.7 and h["xofft"] == 0.7
    assert h["Ly"] == pytest.approx(LY, rel=1e-15) and h["alp"] == ALP and h["gam"] == GAM
    assert h["Deltat"] == 3e-3 and h["explicit"] == 1
    assert h["y"][0] == -0.5 * LY


@pytest.mark.parametrize("target", [(32, 32), (96, 72)])
def test_resolution_change(tmp_path, target):
    p, (vor, phi, _, _) = _write(tmp_path)
    v2, p2, _, _, _ = fieldio.read_restart(p, mgalx=target[0], mgalz=target[1])
    _, kxn, mz = fieldio.grid_sizes(*target)
    kxe, mze = vor.shape[1], vor.shape[2]
    nz, nze = (mz - 1) // 2, (mze - 1) // 2
    for kx in range(kxn):
        for k in range(mz):
            kz = k if k <= nz else k - mz          # signed z wavenumber index
            ke = kz if kz >= 0 else kz + mze
            inside = kx < kxe and -nze <= kz <= nze
            want = vor[:, kx, ke] if inside else 0.0
            assert np.array_equal(v2[:, kx, k], np.broadcast_to(want, (vor.shape[0],)))
    assert np.count_nonzero(v2) == np.count_nonzero(p2)
