"""The reference's own compact y-operators and filter, compiled from its sources.

TEST INFRASTRUCTURE ONLY.  ``oracle/Makefile`` builds ``oracle/_ref/libshear_ref.so``
from mod.f90 and diffy_shear_hepta.f90 as they lie in the reference tree
(ROCm flang, MKL for BLAS ``dcopy``).  The CPU tests call its ``derivyc``
(diffy_shear_hepta.f90:333-450), ``lapcdy`` (diffy_shear_hepta.f90:466-525) and
``filterc`` (filter_shear_hepta.f90:128-165) to pin the numpy restatement in
``oracle/compact.py``.  The library keeps its
grid in Fortran module variables that ``derivadas`` allocates once
(diffy_shear_hepta.f90:535-560), so every grid size runs in a fresh child
process.
"""
import ctypes
import multiprocessing as mp
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libshear_ref.so")


def available():
    return os.path.exists(LIB)


def _cplx(z):
    return (ctypes.c_double * 2)(float(np.real(z)), float(np.imag(z)))


def _child(my, dy, jobs, q):
    try:
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        lib = ctypes.CDLL(LIB)
        n = ctypes.c_int(my)
        h = ctypes.c_double(dy)
        lib.derivadas_(ctypes.byref(n), ctypes.byref(h))
        lib.update_dy_(ctypes.byref(h))
        if any(job[0].startswith("filter") for job in jobs):
            lib.filteras_(ctypes.byref(n))
        out = []
        for job in jobs:
            kind = job[0]
            real = kind.endswith("_real")
            # complex lines are real(2, my); the 00 lines are real(1, my)
            m = ctypes.c_int(1 if real else 2)
            u = np.asarray(job[1], dtype=np.float64 if real else np.complex128)
            res = np.empty_like(u)
            for l in range(u.shape[1]):
                ui = np.ascontiguousarray(u[:, l]).view(np.float64).copy()
                ro = np.zeros_like(ui)
                if kind == "filter_real":
                    lib.filterc_(ui.ctypes.data_as(ctypes.c_void_p), ro.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(1.0), _cplx(1.0), ctypes.byref(m), ctypes.byref(ctypes.c_int(0)))
                elif kind == "filter":
                    lib.filterc_(ui.ctypes.data_as(ctypes.c_void_p), ro.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(job[2][l]), _cplx(job[3][l]), ctypes.byref(m),
                                 ctypes.byref(ctypes.c_int(1)))
                elif kind == "deriv_real":
                    lib.derivyc_(ui.ctypes.data_as(ctypes.c_void_p), ro.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(1.0), _cplx(1.0), ctypes.byref(ctypes.c_int(job[2])),
                                 ctypes.byref(m), ctypes.byref(ctypes.c_int(1)))
                elif kind == "lap_real":
                    lib.lapcdy_(ui.ctypes.data_as(ctypes.c_void_p),
                                ctypes.byref(ctypes.c_double(job[2][l])),
                                ro.ctypes.data_as(ctypes.c_void_p), _cplx(1.0), _cplx(1.0),
                                ctypes.byref(m))
                elif kind == "deriv_reuse":
                    # cross.f90:440-442: a call with ias=1 factors the twisted
                    # operator; the next call, ias=0, reuses that factor
                    ua, shp_a, shp_b, shm_b, order = job[2], job[3][l], job[4][l], job[5][l], job[6]
                    ai = np.ascontiguousarray(ua[:, l]).view(np.float64).copy()
                    ao = np.zeros_like(ai)
                    lib.derivyc_(ai.ctypes.data_as(ctypes.c_void_p), ao.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(shp_a), _cplx(np.conj(shp_a)), ctypes.byref(ctypes.c_int(order)),
                                 ctypes.byref(m), ctypes.byref(ctypes.c_int(1)))
                    lib.derivyc_(ui.ctypes.data_as(ctypes.c_void_p), ro.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(shp_b), _cplx(shm_b), ctypes.byref(ctypes.c_int(order)),
                                 ctypes.byref(m), ctypes.byref(ctypes.c_int(0)))
                elif kind == "deriv":
                    shp, shm, order = job[2][l], job[3][l], job[4]
                    lib.derivyc_(ui.ctypes.data_as(ctypes.c_void_p), ro.ctypes.data_as(ctypes.c_void_p),
                                 _cplx(shp), _cplx(shm), ctypes.byref(ctypes.c_int(order)),
                                 ctypes.byref(m), ctypes.byref(ctypes.c_int(1)))
                else:
                    k2, shp, shm = job[2][l], job[3][l], job[4][l]
                    lib.lapcdy_(ui.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_double(k2)),
                                ro.ctypes.data_as(ctypes.c_void_p), _cplx(shp), _cplx(shm),
                                ctypes.byref(m))
                res[:, l] = ro if real else ro.view(np.complex128)
            out.append(res)
        q.put(("ok", out))
    except Exception as e:  # reported to the parent
        q.put(("err", repr(e)))


def run(my, dy, jobs, timeout=120):
    """Run jobs on lines of length my in a child process.

    jobs: ("deriv", u[my, nl], shp[nl], shm[nl], order),
          ("lap", u[my, nl], k2[nl], shp[nl], shm[nl]),
          ("deriv_real", u[my, nl] real, order), ("lap_real", u[my, nl] real, k2[nl]),
          ("filter", u[my, nl], shp[nl], shm[nl]) or ("filter_real", u[my, nl] real),
          ("deriv_reuse", u[my, nl], ua[my, nl], shp_a[nl], shp_b[nl], shm_b[nl], order):
          per line, derivyc of ua with ias=1 at shp_a, then of u with ias=0 at
          shp_b, shm_b (the second result is returned).
    Returns one array [my, nl] per job.
    """
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=_child, args=(my, dy, jobs, q))
    p.start()
    try:
        status, val = q.get(timeout=timeout)
    finally:
        p.join(timeout=10)
        if p.is_alive():
            p.kill()
    if status != "ok":
        raise RuntimeError(val)
    return val
