"""Restart files of the shear-periodic box in the reference's own format.

``write_restart`` writes what ``escru`` writes (readwrite.f90:1336-1393) and
``read_restart`` reads it as ``getfil`` does (readwrite.f90:72-447), so that
fields pass between the Fortran driver and this package unchanged.  The
file is Fortran unformatted sequential (a 4-byte length before and after
every record, little endian):

  1  mgalx, my, mgalz, numerop                                  int32 x 4
  2  time, Re, alp, Ly/pi, gam, s, chi, mx, my, mz              f64 x 7, int32 x 3
  3  y(0:my-1)                                                  f64 x my
  4  xoffb, xofft, zoffb, zofft                                 f64 x 4
  5  u00(0:my-1), w00(0:my-1)                                   f64 x 2 my
  6  one record per plane j: (vor(i), phi(i), i = 1..mx*mz)     f64 x 2 mx mz
     with vor, phi of the plane as real(mx*mz): re/im fastest, then kx, kz
  7  CFL, uprim, vbulk, xforce                                  f64 x 4
  8  irev, nopt, nparams                                        int32 x 3
  9  explicit, ifix_dt, iadd_force, iadd_mode, iadd_sym,
     iadd_damping, iuse_LES, iadd_visc, idynamic                int32 x nopt (9)
 10  Deltat, force_roll, xforce, zforce, damp_aa, damp_up,
     Cles, Deltag, cutoffx, cutoffz, LESflt(1:3)                f64 x nparams (13)

Fields here are in the package's [y][kx][kz] complex layout.  A file of
another x-z resolution is read the way getfil reads it: kx modes and the
positive and negative halves of kz are copied up to the smaller size and
the rest is zero (readwrite.f90:254-270); another my is not interpolated.
"""
import math
import struct

import numpy as np

NOPT, NPARAMS = 9, 13


def _record(f, payload):
    n = len(payload)
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _records(path):
    with open(path, "rb") as f:
        data = f.read()
    out, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        payload = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n or len(payload) != n:
            raise ValueError(f"{path}: broken record at byte {pos}")
        out.append(payload)
        pos += 8 + n
    return out


def grid_sizes(mgalx, mgalz):
    """mx (reals per kx row), kxn, mz of the 3/2-rule grid (initcr.f90:186-203)"""
    mx = 2 * (mgalx // 3)
    mz = 2 * (mgalz // 3) - 1
    return mx, mx // 2, mz


def write_restart(path, vor, phi, u00, w00, *, mgalx, mgalz, time, Re, alp, gam, Ly=1.0,
                  s=1.0, chi=0.0, xoffb=0.0, xofft=0.0, zoffb=0.0, zofft=0.0, CFL=0.5,
                  uprim=0.1, vbulk=0.0, Deltat=0.0, explicit=1, irev=0):
    vor = np.asarray(vor, dtype=np.complex128)
    phi = np.asarray(phi, dtype=np.complex128)
    my = vor.shape[0]
    mx, kxn, mz = grid_sizes(mgalx, mgalz)
    if vor.shape != (my, kxn, mz) or phi.shape != vor.shape:
        raise ValueError(f"fields must be ({my}, {kxn}, {mz}) complex")
    y = np.arange(my) * (Ly / my) - 0.5 * Ly
    with open(path, "wb") as f:
        _record(f, struct.pack("<4i", mgalx, my, mgalz, 1))
        _record(f, struct.pack("<7d3i", time, Re, alp, Ly / math.pi, gam, s, chi, mx, my, mz))
        _record(f, y.astype("<f8").tobytes())
        _record(f, struct.pack("<4d", xoffb, xofft, zoffb, zofft))
        _record(f, np.concatenate([np.asarray(u00, "<f8"), np.asarray(w00, "<f8")]).tobytes())
        for j in range(my):
            # plane as real(2, kxn, mz) with kx before kz: Fortran order = C order of [kz][kx][re/im]
            a = np.empty((mz, kxn, 2, 2))
            a[:, :, 0, 0] = vor[j].T.real
            a[:, :, 0, 1] = phi[j].T.real
            a[:, :, 1, 0] = vor[j].T.imag
            a[:, :, 1, 1] = phi[j].T.imag
            _record(f, a.astype("<f8").tobytes())
        _record(f, struct.pack("<4d", CFL, uprim, vbulk, 0.0))
        _record(f, struct.pack("<3i", irev, NOPT, NPARAMS))
        _record(f, struct.pack("<9i", explicit, 0, 0, 5, 0, 0, 0, 0, 0))
        _record(f, struct.pack("<13d", Deltat, *([0.0] * 12)))


def read_restart(path, mgalx=None, mgalz=None):
    """Returns (vor, phi, u00, w00, header).  With mgalx, mgalz given, the
    fields are mapped to that x-z resolution as getfil does."""
    rec = _records(path)
    mgalxe, mye, mgalze, _ = struct.unpack("<4i", rec[0])
    t, Re, alp, Lypi, gam, s, chi, mxe, my, mze = struct.unpack("<7d3i", rec[1])
    if my != mye:
        raise ValueError(f"{path}: header my {mye} != {my}")
    y = np.frombuffer(rec[2], "<f8")
    xoffb, xofft, zoffb, zofft = struct.unpack("<4d", rec[3])
    u0w0 = np.frombuffer(rec[4], "<f8")
    u00, w00 = u0w0[:my].copy(), u0w0[my:].copy()
    kxe, mze_ = mxe // 2, mze
    planes = rec[5:5 + my]
    if len(planes) != my:
        raise ValueError(f"{path}: {len(planes)} planes for my = {my}")
    tail = rec[5 + my:]
    header = {"mgalx": mgalxe, "my": my, "mgalz": mgalze, "time": t, "Re": Re, "alp": alp,
              "Ly": Lypi * math.pi, "gam": gam, "s": s, "chi": chi, "y": y,
              "xoffb": xoffb, "xofft": xofft, "zoffb": zoffb, "zofft": zofft}
    if len(tail) >= 4:
        header["CFL"], header["uprim"], header["vbulk"], _ = struct.unpack("<4d", tail[0])
        header["irev"], nopt, nparams = struct.unpack("<3i", tail[1])
        header["ioptions"] = struct.unpack(f"<{nopt}i", tail[2])
        header["params"] = struct.unpack(f"<{nparams}d", tail[3])
        header["Deltat"] = header["params"][0]
        header["explicit"] = header["ioptions"][0]
    if abs(header["Ly"] + 2.0 * y[0]) > 1e-13:
        raise ValueError(f"{path}: y(0) does not match Ly")
    mgalx = mgalxe if mgalx is None else mgalx
    mgalz = mgalze if mgalz is None else mgalz
    _, kxn, mz = grid_sizes(mgalx, mgalz)
    # readwrite.f90:254-270: kx up to the smaller size; kz in two halves
    mxm = min(kxn, kxe)
    mzm = min(mz, mze_)
    klen = (mzm + 1) // 2
    vor = np.zeros((my, kxn, mz), dtype=np.complex128)
    phi = np.zeros_like(vor)
    for j, p in enumerate(planes):
        a = np.frombuffer(p, "<f8").reshape(mze_, kxe, 2, 2)   # [kz][kx][re/im][field]
        fv = (a[:, :, 0, 0] + 1j * a[:, :, 1, 0]).T                # (kxe, mze)
        fp = (a[:, :, 0, 1] + 1j * a[:, :, 1, 1]).T
        for dst, src in ((vor, fv), (phi, fp)):
            dst[j, :mxm, :klen] = src[:mxm, :klen]
            if klen > 1:
                dst[j, :mxm, mz - (klen - 1):] = src[:mxm, mze_ - (klen - 1):]
    return vor, phi, u00, w00, header
