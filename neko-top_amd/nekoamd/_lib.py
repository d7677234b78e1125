"""ctypes binding of libnekoamd.so (the C-ABI in include/nekoamd.h).

Fails loudly: if the HIP library is not built, or the device cannot be
initialised, every entry raises; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libnekoamd.so")
_LIB = None

HOST_TO_DEVICE, DEVICE_TO_HOST, DEVICE_TO_DEVICE = 1, 2, 3
GS_ADD, GS_MUL, GS_MIN, GS_MAX = 1, 2, 3, 4
FACE_XMIN, FACE_XMAX, FACE_YMIN, FACE_YMAX, FACE_ZMIN, FACE_ZMAX = 1, 2, 4, 8, 16, 32


class NekoAmdError(RuntimeError):
    pass


class CgMonitor(C.Structure):
    _fields_ = [("iter", C.c_int), ("done", C.c_int),
                ("res_start", C.c_double), ("res_final", C.c_double),
                ("nupdate", C.c_int), ("pad", C.c_int)]


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NekoAmdError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.nekoamd_error_string.restype = C.c_char_p
        L.nekoamd_stream.restype = C.c_void_p
        L.nekoamd_rng_u64.restype = C.c_uint64
        _LIB = L
    return _LIB


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().nekoamd_error_string().decode(errors="replace")
        raise NekoAmdError(f"{what}: status {rc}: {msg}")


def ri(v):
    return C.byref(C.c_int(int(v)))


def ri64(v):
    return C.byref(C.c_int64(int(v)))


def rd(v):
    return C.byref(C.c_double(float(v)))


def ptr(a) -> C.c_void_p:
    """Device pointer of a DeviceArray, or a host numpy array's data."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(a.ptr)
