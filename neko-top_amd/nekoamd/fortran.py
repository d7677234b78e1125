"""ctypes binding of the Fortran host library libnekoamd_f.so
(include/nekoamd_fortran.h)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import PKG_DIR, NekoAmdError, check, lib

_FLIB = None


def flib() -> C.CDLL:
    global _FLIB
    if _FLIB is None:
        path = os.path.join(PKG_DIR, "libnekoamd_f.so")
        if not os.path.exists(path):
            raise NekoAmdError(f"{path} is not built; run __graft_entry__.build()")
        lib()
        _FLIB = C.CDLL(path)
    return _FLIB


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def pcg_solve(nx, ny, nz, lx, f, kind=0, jitter=0.05, seed=1234, re=100.0, dt=1e-3,
              abstol=1e-8, max_iter=800):
    """Jacobi-PCG through the Fortran ksp_t/pc_t/ax_t surface."""
    n = nx * ny * nz * lx ** 3
    f = np.ascontiguousarray(f, dtype=np.float64)
    assert f.size == n
    x = np.zeros(n)
    it = C.c_int(0)
    res = np.zeros(2)
    rc = flib().nekoamd_f_pcg_solve(_i(nx), _i(ny), _i(nz), _i(lx), _d(jitter),
                                    C.byref(C.c_int64(seed)), _i(kind), _d(re), _d(dt),
                                    f.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                    _d(abstol), _i(max_iter), C.byref(it),
                                    res.ctypes.data_as(C.c_void_p))
    check(rc, "f_pcg_solve")
    return x, it.value, res[0], res[1]
