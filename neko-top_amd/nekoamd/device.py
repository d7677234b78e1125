"""Device memory and timing helpers over the C-ABI (no torch needed)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (DEVICE_TO_DEVICE, DEVICE_TO_HOST, HOST_TO_DEVICE, check, lib,
                   ptr, ri, ri64)

_INITIALISED = False


def init(device: int = 0):
    global _INITIALISED
    check(lib().nekoamd_device_init(ri(device)), "device_init")
    _INITIALISED = True


def ensure():
    if not _INITIALISED:
        init(0)


def sync():
    check(lib().nekoamd_device_sync(), "device_sync")


def name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().nekoamd_device_name(buf, ri(256)), "device_name")
    return buf.value.decode()


class DeviceArray:
    """A device allocation owned by Python (freed on deletion)."""

    def __init__(self, n: int, dtype=np.float64):
        ensure()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p(0)
        check(lib().nekoamd_device_alloc(C.byref(p), ri64(self.nbytes)), "alloc")
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.n:
            check(lib().nekoamd_device_memcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data),
                                              ri64(self.nbytes), ri(HOST_TO_DEVICE), ri(1)),
                  "memcpy h2d")

    def get(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(lib().nekoamd_device_memcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr),
                                              ri64(self.nbytes), ri(DEVICE_TO_HOST), ri(1)),
                  "memcpy d2h")
        return out

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        if self.n:
            check(lib().nekoamd_device_memcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                              ri64(self.nbytes), ri(DEVICE_TO_DEVICE), ri(0)),
                  "memcpy d2d")

    def __del__(self):
        try:
            if self.ptr:
                lib().nekoamd_device_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


def zeros(n, dtype=np.float64) -> DeviceArray:
    d = DeviceArray(n, dtype)
    if d.n:
        if d.dtype == np.float64:
            check(lib().nekoamd_rzero(ptr(d), ri64(n)), "rzero")
        else:
            d.upload(np.zeros(n, dtype))
    return d


class Event:
    def __init__(self):
        p = C.c_void_p(0)
        check(lib().nekoamd_event_create(C.byref(p)), "event_create")
        self.h = p.value

    def record(self):
        check(lib().nekoamd_event_record(C.c_void_p(self.h)), "event_record")

    def sync(self):
        check(lib().nekoamd_event_sync(C.c_void_p(self.h)), "event_sync")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float(0)
        check(lib().nekoamd_event_elapsed(C.c_void_p(self.h), C.c_void_p(later.h),
                                          C.byref(ms)), "event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            lib().nekoamd_event_destroy(C.c_void_p(self.h))
        except Exception:
            pass
