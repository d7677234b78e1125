"""The learned solver aid (nekoamd_mlp_*): element-local network, training
and the Galerkin-scaled initial-guess direction."""
import ctypes as C

import numpy as np

from . import device as dev
from ._lib import check, lib, ptr, ri, ri64

NPARAM = 330240
KIN = 544
NP = 512
HID = 256


class MLP:
    def __init__(self, nx, ny, nz, lx=8, capacity=0, batch=4096, seed=1):
        h = C.c_void_p(0)
        check(lib().nekoamd_mlp_create(C.byref(h), ri(nx), ri(ny), ri(nz), ri(lx), ri(capacity),
                                       ri(batch), C.byref(C.c_int64(seed))), "mlp_create")
        self.h = h.value

    def get_params(self):
        out = np.zeros(NPARAM, dtype=np.float32)
        check(lib().nekoamd_mlp_get_params(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p)),
              "mlp_get_params")
        return out

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == NPARAM
        check(lib().nekoamd_mlp_set_params(C.c_void_p(self.h), th.ctypes.data_as(C.c_void_p)),
              "mlp_set_params")

    def forward(self, X, M, hidden=False):
        """X: device uint16 (bf16 bits) [M, 544]; returns Y [M, 512] (float32)
        and, with hidden, H1, H2 as bf16 bit arrays."""
        Y = dev.DeviceArray(M * NP, np.float32)
        H1 = dev.DeviceArray(M * HID, np.uint16) if hidden else None
        H2 = dev.DeviceArray(M * HID, np.uint16) if hidden else None
        check(lib().nekoamd_mlp_forward(C.c_void_p(self.h), ptr(X), ri(M), ptr(Y), ptr(H1),
                                        ptr(H2)), "mlp_forward")
        out = Y.get().reshape(M, NP)
        if not hidden:
            return out
        return out, H1.get().reshape(M, HID), H2.get().reshape(M, HID)

    def loss_grad(self, X, T, M):
        g = dev.DeviceArray(NPARAM, np.float32)
        loss = C.c_double(0)
        check(lib().nekoamd_mlp_loss_grad(C.c_void_p(self.h), ptr(X), ptr(T), ri(M),
                                          C.byref(loss), ptr(g)), "mlp_loss_grad")
        return loss.value, g.get()

    def harvest(self, b, x):
        check(lib().nekoamd_mlp_harvest(C.c_void_p(self.h), ptr(b), ptr(x)), "mlp_harvest")

    @property
    def pairs(self):
        n = C.c_int(0)
        check(lib().nekoamd_mlp_pairs(C.c_void_p(self.h), C.byref(n)), "mlp_pairs")
        return n.value

    def train(self, epochs, batch, lr=1e-3, seed=7):
        l0, l1 = C.c_double(0), C.c_double(0)
        check(lib().nekoamd_mlp_train(C.c_void_p(self.h), ri(epochs), ri(batch),
                                      C.byref(C.c_float(lr)), C.byref(C.c_int64(seed)),
                                      C.byref(l0), C.byref(l1)), "mlp_train")
        return l0.value, l1.value

    def pre(self, case, b, mask, h1, h2=None, ifh2=False, apply=True, save=False):
        G = case.G
        g = C.c_double(0)
        check(lib().nekoamd_mlp_pre(
            C.c_void_p(self.h), ptr(b), ptr(case.mult), C.c_void_p(case.gs.h), ptr(mask),
            ri64(getattr(mask, "count", 0)), ptr(case.dx), ptr(h1), ptr(h2), ptr(G["G11"]),
            ptr(G["G12"]), ptr(G["G13"]), ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]),
            ptr(case.B), ri(case.lx), ri(1 if ifh2 else 0), ri(1 if apply else 0),
            ri(1 if save else 0), C.byref(g)), "mlp_pre")
        return g.value

    def post(self, x, harvest=False):
        check(lib().nekoamd_mlp_post(C.c_void_p(self.h), ptr(x), ri(1 if harvest else 0)),
              "mlp_post")

    def __del__(self):
        try:
            lib().nekoamd_mlp_free(C.c_void_p(self.h))
        except Exception:
            pass


def to_bf16_bits(a):
    """float -> bf16 bit patterns (nearest, ties to even), as uint16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)
