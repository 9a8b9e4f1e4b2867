"""MI355X (gfx950) RK3 time advance for the shear-periodic box.

Python host side of ``libshear_rk3.so`` (C ABI in ``include/shear_rk3.h``).
It mirrors the reference's Fortran interface for the hot path
(hiiragiyua/shear_nnf):

* ``ShearBox``            -- module state (ctes/running/bcs, mod.f90) + the
                              device handle; ``cross(nstep)`` is the drop-in
                              for ``cross`` (cross.f90:3-1101)
* ``ShearBox.derivyc``     -- diffy_shear_hepta.f90:333-450
* ``ShearBox.lapcdy``      -- diffy_shear_hepta.f90:466-525
* ``ShearBox.fourxz``      -- fou3D_thread.f:22-73
* ``ShearBox.hvhg``        -- cross.f90:1127-1765 (one nonlinear evaluation)

Everything runs in the HIP library; when it is missing the calls raise.
"""
import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshear_rk3.so")

SNF_OK, SNF_BAD_PARAM, SNF_DEVICE_ERROR, SNF_NO_DEVICE, SNF_NOT_READY = 0, 1, 2, 3, 4
SNF_RUNNING, SNF_STOP, SNF_DUMP = 0, 1, 2
SNF_PREC_FP64, SNF_PREC_MIXED = 0, 1
UNER_LEN = 24
EXPORTS = (
    "snf_create", "snf_destroy", "snf_last_error", "snf_sizes",
    "snf_state_upload", "snf_state_download", "snf_state_upload_ykxkz",
    "snf_state_download_ykxkz", "snf_advance", "snf_history", "snf_statistics",
    "snf_op_derivy", "snf_op_helmholtz", "snf_op_fourxz", "snf_op_hvhg",
    "snf_set_timing", "snf_get_timing", "snf_copy_probe", "snf_op_getv", "snf_energy",
    "snf_spectra", "snf_reset_statistics", "snf_state_len", "snf_shoot", "snf_jvp_batch",
    "snf_advance_batch", "snf_energy_noalloc",
)
KERNEL_CLASSES = ("xz_A", "xz_B", "xz_C", "y_lines", "small")


class SnfParams(ctypes.Structure):
    _fields_ = [
        ("mgalx", ctypes.c_int32), ("my", ctypes.c_int32), ("mgalz", ctypes.c_int32),
        ("Re", ctypes.c_double), ("alp", ctypes.c_double), ("gam", ctypes.c_double),
        ("Ly", ctypes.c_double), ("s", ctypes.c_double), ("chi", ctypes.c_double),
        ("CFL", ctypes.c_double),
        ("explicit_visc", ctypes.c_int32), ("nhist", ctypes.c_int32),
        ("readflag", ctypes.c_int32),
        ("uprim", ctypes.c_double), ("vbulk", ctypes.c_double),
        ("etime", ctypes.c_double), ("dumptime", ctypes.c_double), ("dtimag", ctypes.c_double),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
        ("ydealias", ctypes.c_int32),
    ]


class SnfRunStatus(ctypes.Structure):
    _fields_ = [
        ("steps_done", ctypes.c_int32), ("stop", ctypes.c_int32),
        ("time", ctypes.c_double), ("Deltat", ctypes.c_double),
        ("xoffb", ctypes.c_double), ("xofft", ctypes.c_double),
        ("umax", ctypes.c_double * 4), ("CFLv", ctypes.c_double),
        ("nhist_records", ctypes.c_int32), ("stop_reason", ctypes.c_int32),
    ]


_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libshear_rk3.so not built ({path}); run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    H = ctypes.c_void_p
    lib.snf_create.argtypes = [P(SnfParams), P(H)]
    lib.snf_destroy.argtypes = [H]
    lib.snf_last_error.restype = ctypes.c_char_p
    lib.snf_sizes.argtypes = [H, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int64)]
    for nm in ("snf_state_upload", "snf_state_upload_ykxkz"):
        getattr(lib, nm).argtypes = [H, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double]
    lib.snf_state_download.argtypes = [H, dp, dp, dp, dp, dp, dp, dp]
    lib.snf_state_download_ykxkz.argtypes = [H, dp, dp, dp, dp]
    lib.snf_advance.argtypes = [H, ctypes.c_int32, P(SnfRunStatus)]
    lib.snf_history.argtypes = [H, dp, ctypes.c_int32, P(ctypes.c_int32)]
    lib.snf_statistics.argtypes = [H, dp, dp, P(ctypes.c_int32), dp]
    lib.snf_op_derivy.argtypes = [H, dp, dp, ctypes.c_int32, dp, dp, ctypes.c_int32]
    lib.snf_op_helmholtz.argtypes = [H, dp, dp, ctypes.c_int32, dp, dp, dp]
    lib.snf_op_fourxz.argtypes = [H, dp, dp, ctypes.c_int32, ctypes.c_int32]
    lib.snf_op_hvhg.argtypes = [H, dp, dp, dp, dp, dp, dp, dp, dp]
    lib.snf_set_timing.argtypes = [H, ctypes.c_int32]
    lib.snf_get_timing.argtypes = [H, dp, P(ctypes.c_int32), ctypes.c_int32]
    lib.snf_copy_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, dp]
    lib.snf_op_getv.argtypes = [H, dp, dp, ctypes.c_int32]
    lib.snf_energy.argtypes = [H, dp]
    lib.snf_energy_noalloc.argtypes = [H, dp]
    lib.snf_spectra.argtypes = [H, dp, dp]
    lib.snf_reset_statistics.argtypes = [H]
    lib.snf_state_len.argtypes = [H, P(ctypes.c_int64)]
    lib.snf_advance_batch.argtypes = [P(H), ctypes.c_int32, ctypes.c_int32, P(SnfRunStatus)]
    d = ctypes.c_double
    lib.snf_shoot.argtypes = [H, dp, d, d, d, d, dp, P(SnfRunStatus)]
    lib.snf_jvp_batch.argtypes = [H, dp, dp, dp, ctypes.c_int32, dp, d, d, d, d, dp]
    _lib = lib
    return lib


def copy_probe(nbytes, iters=20, device=0):
    """mean ms of a plain device copy of nbytes (read nbytes + write nbytes)"""
    lib = load_library()
    ms = ctypes.c_double()
    rc = lib.snf_copy_probe(device, int(nbytes), int(iters), ctypes.byref(ms))
    if rc != SNF_OK:
        raise SnfError(lib.snf_last_error().decode())
    return ms.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _need(a, shape, name):
    """the C side reads exactly this many elements: refuse anything else"""
    if tuple(np.shape(a)) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(np.shape(a))}, expected {tuple(shape)}")


def advance_batch(boxes, nstep):
    """Advances every ShearBox of `boxes` by up to nstep steps with all their
    work in flight at once (one stream each); returns the statuses and
    appends the UNER records to each box's history."""
    lib = load_library()
    n = len(boxes)
    hs = (ctypes.c_void_p * n)(*[b.h.value for b in boxes])
    st = (SnfRunStatus * n)()
    rc = lib.snf_advance_batch(hs, n, int(nstep), st)
    if rc != SNF_OK:
        raise SnfError(f"snf error {rc}: {lib.snf_last_error().decode()}")
    for b in boxes:
        b.fetch_history()
    return list(st)


class SnfError(RuntimeError):
    pass


class ShearBox:
    """One shear-periodic box on one GPU (the state of cross.f90 + modules)."""

    def __init__(self, mgalx, my, mgalz, Re, alp=np.pi, gam=2 * np.pi, Ly=1.0, s=1.0,
                 chi=0.0, CFL=0.5, explicit=1, nhist=1, readflag=1, uprim=0.1, vbulk=0.0,
                 etime=1e10, dumptime=1e10, dtimag=1e10, device=0, ydealias=0,
                 precision="fp64"):
        """precision: "fp64", or "mixed" (FP32 x-z transforms and nonlinear
        products, FP64 state, y operators and statistics sums)."""
        self.lib = load_library()
        prec = {"fp64": SNF_PREC_FP64, "mixed": SNF_PREC_MIXED}[precision]
        p = SnfParams(mgalx, my, mgalz, Re, alp, gam, Ly, s, chi, CFL, explicit, nhist,
                      readflag, uprim, vbulk, etime, dumptime, dtimag, prec, device, ydealias)
        h = ctypes.c_void_p()
        self._check(self.lib.snf_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        kxn, mz, nc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        self.lib.snf_sizes(h, ctypes.byref(kxn), ctypes.byref(mz), ctypes.byref(nc))
        self.kxn, self.mz, self.nc = kxn.value, mz.value, nc.value
        self.my, self.mgalx, self.mgalz = my, mgalx, mgalz
        self.params = p
        self.history = []

    def _check(self, rc):
        if rc != SNF_OK:
            raise SnfError(f"snf error {rc}: {self.lib.snf_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.snf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def set_state(self, vor, phi, u00, w00, time=0.0, xoffb=0.0, xofft=0.0):
        """vor, phi in the device layout [y][kx][kz]"""
        vor, phi = _c128(vor), _c128(phi)
        u00 = np.ascontiguousarray(u00, dtype=np.float64)
        w00 = np.ascontiguousarray(w00, dtype=np.float64)
        for a, nm in ((vor, "vor"), (phi, "phi")):
            _need(a, (self.my, self.kxn, self.mz), nm)
        for a, nm in ((u00, "u00"), (w00, "w00")):
            _need(a, (self.my,), nm)
        self._check(self.lib.snf_state_upload_ykxkz(self.h, _ptr(vor.view(np.float64)),
                                                    _ptr(phi.view(np.float64)), _ptr(u00),
                                                    _ptr(w00), time, xoffb, xofft))

    def set_state_jik(self, vor, phi, u00, w00, time=0.0, xoffb=0.0, xofft=0.0):
        """vor, phi in the reference layout complex(0:my-1, 0:mx1, 0:mz-1)
        (Fortran order, i.e. numpy shape (mz, kxn, my) C-contiguous)"""
        vor, phi = _c128(vor), _c128(phi)
        u00 = np.ascontiguousarray(u00, dtype=np.float64)
        w00 = np.ascontiguousarray(w00, dtype=np.float64)
        for a, nm in ((vor, "vor"), (phi, "phi")):
            _need(a, (self.mz, self.kxn, self.my), nm)
        for a, nm in ((u00, "u00"), (w00, "w00")):
            _need(a, (self.my,), nm)
        self._check(self.lib.snf_state_upload(self.h, _ptr(vor.view(np.float64)),
                                              _ptr(phi.view(np.float64)), _ptr(u00), _ptr(w00),
                                              time, xoffb, xofft))

    def get_state(self):
        vor = np.empty((self.my, self.kxn, self.mz), dtype=np.complex128)
        phi = np.empty_like(vor)
        u00 = np.empty(self.my)
        w00 = np.empty(self.my)
        self._check(self.lib.snf_state_download_ykxkz(self.h, _ptr(vor.view(np.float64)),
                                                      _ptr(phi.view(np.float64)), _ptr(u00),
                                                      _ptr(w00)))
        return vor, phi, u00, w00

    def get_state_jik(self):
        vor = np.empty((self.mz, self.kxn, self.my), dtype=np.complex128)
        phi = np.empty_like(vor)
        u00 = np.empty(self.my)
        w00 = np.empty(self.my)
        t, xb, xt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.snf_state_download(self.h, _ptr(vor.view(np.float64)),
                                                _ptr(phi.view(np.float64)), _ptr(u00),
                                                _ptr(w00), ctypes.byref(t), ctypes.byref(xb),
                                                ctypes.byref(xt)))
        return vor, phi, u00, w00, t.value, xb.value, xt.value

    # ---------------------------------------------------------------- cross
    def cross(self, nstep):
        """Advance up to nstep RK3 steps (cross.f90:3-1101); returns the
        run status and appends the UNER records to ``self.history``."""
        st = SnfRunStatus()
        self._check(self.lib.snf_advance(self.h, int(nstep), ctypes.byref(st)))
        self.fetch_history()
        return st

    def advance_async_status(self, nstep):
        st = SnfRunStatus()
        self._check(self.lib.snf_advance(self.h, int(nstep), ctypes.byref(st)))
        return st

    def fetch_history(self):
        buf = np.empty((4096, UNER_LEN))
        n = ctypes.c_int32()
        while True:
            self._check(self.lib.snf_history(self.h, _ptr(buf), buf.shape[0], ctypes.byref(n)))
            self.history.extend(buf[:n.value].copy())
            if n.value < buf.shape[0]:
                break
        return self.history

    def statistics(self):
        st = np.empty((self.my, 23))
        st2 = np.empty((self.my, 23))
        nacum = ctypes.c_int32()
        tdt = ctypes.c_double()
        self._check(self.lib.snf_statistics(self.h, _ptr(st), _ptr(st2), ctypes.byref(nacum),
                                            ctypes.byref(tdt)))
        return st, st2, nacum.value, tdt.value

    def spectra(self):
        """(sp, sp2), each (8, kxn, mz): the accumulated spectra of the
        statistics (cross.f90:1970-2016)"""
        sp = np.empty((self.mz, self.kxn, 8))
        sp2 = np.empty((self.mz, self.kxn, 8))
        self._check(self.lib.snf_spectra(self.h, _ptr(sp), _ptr(sp2)))
        return sp.transpose(2, 1, 0).copy(), sp2.transpose(2, 1, 0).copy()

    def reset_statistics(self):
        """zero the statistics, spectra and their counters (readwrite.f90:1322-1334)"""
        self._check(self.lib.snf_reset_statistics(self.h))

    # ---- shooting: state vectors [vor | phi | u00 | w00] ([y][kx][kz] complex)
    def pack_state(self, vor, phi, u00, w00):
        return np.concatenate([_c128(vor).view(np.float64).ravel(),
                               _c128(phi).view(np.float64).ravel(),
                               np.asarray(u00, np.float64), np.asarray(w00, np.float64)])

    def unpack_state(self, x):
        nc2 = 2 * self.my * self.kxn * self.mz
        shp = (self.my, self.kxn, self.mz)
        vor = x[:nc2].copy().view(np.complex128).reshape(shp)
        phi = x[nc2:2 * nc2].copy().view(np.complex128).reshape(shp)
        return vor, phi, x[2 * nc2:2 * nc2 + self.my].copy(), x[2 * nc2 + self.my:].copy()

    def shoot(self, x, T, t0=0.0, xoffb0=0.0, xofft0=0.0):
        """Phi_T(x): the shooting map of calc_dns_disk3d (gmres_shear.f90:2951-3120)"""
        x = np.ascontiguousarray(x, np.float64)
        _need(x, (self.state_len(),), "x")
        y = np.empty_like(x)
        st = SnfRunStatus()
        self._check(self.lib.snf_shoot(self.h, _ptr(x), t0, xoffb0, xofft0, T, _ptr(y),
                                       ctypes.byref(st)))
        return y, st

    def jvp_batch(self, x0, dirs, eps, T, y0=None, t0=0.0, xoffb0=0.0, xofft0=0.0):
        """(Phi_T(x0 + eps_d d) - Phi_T(x0)) / eps_d for the rows d of dirs
        (matvec_dns_disk3d, gmres_shear.f90:2666-2730)"""
        n = self.state_len()
        x0 = np.ascontiguousarray(x0, np.float64)
        dirs = np.ascontiguousarray(np.atleast_2d(dirs), np.float64)
        _need(x0, (n,), "x0")
        _need(dirs, (dirs.shape[0], n), "dirs")
        eps = np.ascontiguousarray(np.broadcast_to(eps, (dirs.shape[0],)), np.float64)
        out = np.empty_like(dirs)
        y0a = np.ascontiguousarray(y0, np.float64) if y0 is not None else None
        if y0a is not None:
            _need(y0a, (n,), "y0")
        self._check(self.lib.snf_jvp_batch(self.h, _ptr(x0), _ptr(y0a) if y0a is not None else None,
                                           _ptr(dirs), dirs.shape[0], _ptr(eps), t0, xoffb0,
                                           xofft0, T, _ptr(out)))
        return out

    def state_len(self):
        """doubles in a state vector [vor | phi | u00 | w00]"""
        n = ctypes.c_int64()
        self._check(self.lib.snf_state_len(self.h, ctypes.byref(n)))
        return n.value

    def set_timing(self, on):
        self._check(self.lib.snf_set_timing(self.h, 1 if on else 0))

    def get_timing(self):
        """{class: (total_ms, launches)} from HIP events around each launch"""
        n = len(KERNEL_CLASSES)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        self._check(self.lib.snf_get_timing(self.h, _ptr(ms),
                                            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
        return {k: (float(ms[q]), int(cnt[q])) for q, k in enumerate(KERNEL_CLASSES)}

    # ------------------------------------------------------------ operators
    def derivyc(self, u, shp, shm, order):
        """u: (my, nl) complex lines; returns d^order u / dy^order"""
        u = _c128(u)
        _need(u, (self.my, u.shape[1] if u.ndim == 2 else -1), "u")
        nl = u.shape[1]
        out = np.empty_like(u)
        shp, shm = _c128(shp), _c128(shm)
        _need(shp, (nl,), "shp")
        _need(shm, (nl,), "shm")
        self._check(self.lib.snf_op_derivy(self.h, _ptr(u.view(np.float64)),
                                           _ptr(out.view(np.float64)), nl,
                                           _ptr(shp.view(np.float64)), _ptr(shm.view(np.float64)),
                                           int(order)))
        return out

    def lapcdy(self, u, k2, shp, shm):
        u = _c128(u)
        _need(u, (self.my, u.shape[1] if u.ndim == 2 else -1), "u")
        nl = u.shape[1]
        out = np.empty_like(u)
        k2 = np.ascontiguousarray(k2, dtype=np.float64)
        shp, shm = _c128(shp), _c128(shm)
        for a, nm in ((k2, "k2"), (shp, "shp"), (shm, "shm")):
            _need(a, (nl,), nm)
        self._check(self.lib.snf_op_helmholtz(self.h, _ptr(u.view(np.float64)),
                                              _ptr(out.view(np.float64)), nl, _ptr(k2),
                                              _ptr(shp.view(np.float64)),
                                              _ptr(shm.view(np.float64))))
        return out

    def fourxz(self, a, direction):
        """direction=+1: (np, kxn, mz) complex -> (np, mgalx, mgalz) real;
        direction=-1: the reverse."""
        if direction == 1:
            a = _c128(a)
            _need(a, (a.shape[0] if a.ndim == 3 else -1, self.kxn, self.mz), "a")
            out = np.empty((a.shape[0], self.mgalx, self.mgalz))
            self._check(self.lib.snf_op_fourxz(self.h, _ptr(a.view(np.float64)), _ptr(out),
                                               a.shape[0], 1))
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            _need(a, (a.shape[0] if a.ndim == 3 else -1, self.mgalx, self.mgalz), "a")
            out = np.empty((a.shape[0], self.kxn, self.mz), dtype=np.complex128)
            self._check(self.lib.snf_op_fourxz(self.h, _ptr(a), _ptr(out.view(np.float64)),
                                               a.shape[0], -1))
        return out

    def getv(self, f, iopt):
        """getv (operators.f90:735-793) on a (my, kxn, mz) field with the twist
        of the current offsets: iopt=+1 phi -> v, iopt=-1 v -> phi.  The
        (0,0) line of the result is zero."""
        f = _c128(f)
        _need(f, (self.my, self.kxn, self.mz), "f")
        out = np.zeros_like(f)
        self._check(self.lib.snf_op_getv(self.h, _ptr(f.view(np.float64)),
                                         _ptr(out.view(np.float64)), int(iopt)))
        return out

    def energy(self):
        """energy (operators.f90:870-924) of the device state: sums of
        |vor|^2, |phi|^2 (kx > 0 counted twice), U^2, W^2."""
        out = np.empty(4)
        self._check(self.lib.snf_energy(self.h, _ptr(out)))
        return out

    def energy_noalloc(self):
        """energy() through snf_energy_noalloc: no device allocation, the
        partial sums in the library's static scratch"""
        out = np.empty(4)
        self._check(self.lib.snf_energy_noalloc(self.h, _ptr(out)))
        return out

    def hvhg(self, v, dvdy, dvordy):
        """Nonlinear terms of the uploaded state given v, dv/dy, d(omega_y)/dy."""
        v, dvdy, dvordy = _c128(v), _c128(dvdy), _c128(dvordy)
        for a, nm in ((v, "v"), (dvdy, "dvdy"), (dvordy, "dvordy")):
            _need(a, (self.my, self.kxn, self.mz), nm)
        o1 = np.empty_like(v)
        o2 = np.empty_like(v)
        o3 = np.empty_like(v)
        rf0u = np.empty(self.my)
        rf0w = np.empty(self.my)
        f = lambda a: _ptr(a.view(np.float64))
        self._check(self.lib.snf_op_hvhg(self.h, f(v), f(dvdy), f(dvordy), f(o1), f(o2), f(o3),
                                         _ptr(rf0u), _ptr(rf0w)))
        return o1, o2, o3, rf0u, rf0w
