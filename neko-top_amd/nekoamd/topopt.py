"""ctypes wrappers of the topology-optimisation and MMA entry points."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as dev
from ._lib import check, lib, ptr, rd, ri, ri64


def brinkman(fx, fy, fz, u, v, w, chi):
    check(lib().nekoamd_brinkman(ptr(fx), ptr(fy), ptr(fz), ptr(u), ptr(v), ptr(w), ptr(chi),
                                 ri64(fx.n)), "brinkman")


def update_permeability(res, rho, perm_0=1000.0, perm_1=0.0, penalty=1.0):
    check(lib().nekoamd_update_permeability(ptr(res), ptr(rho), rd(perm_0), rd(perm_1),
                                            rd(penalty), ri64(res.n)), "update_permeability")


def ramp_fwd(out, x, f_min, f_max, q, convex_up=True):
    check(lib().nekoamd_ramp_fwd(ptr(out), ptr(x), rd(f_min), rd(f_max), rd(q),
                                 ri(int(convex_up)), ri64(out.n)), "ramp_fwd")


def ramp_bwd(df_din, df_dout, x, f_min, f_max, q, convex_up=True):
    check(lib().nekoamd_ramp_bwd(ptr(df_din), ptr(df_dout), ptr(x), rd(f_min), rd(f_max), rd(q),
                                 ri(int(convex_up)), ri64(df_din.n)), "ramp_bwd")


def linear_fwd(out, x, f_min, f_max):
    check(lib().nekoamd_linear_fwd(ptr(out), ptr(x), rd(f_min), rd(f_max), ri64(out.n)),
          "linear_fwd")


def linear_bwd(df_din, df_dout, f_min, f_max):
    check(lib().nekoamd_linear_bwd(ptr(df_din), ptr(df_dout), rd(f_min), rd(f_max),
                                   ri64(df_din.n)), "linear_bwd")


def dissipation_sensitivity(s, u, v, w, ua, va, wa, K=1.0, obj_scale=1.0, if_lube=True):
    check(lib().nekoamd_dissipation_sensitivity(ptr(s), ptr(u), ptr(v), ptr(w), ptr(ua), ptr(va),
                                                ptr(wa), rd(K), rd(obj_scale), ri(int(if_lube)),
                                                ri64(s.n)), "dissipation_sensitivity")


def adjoint_lube(fx, fy, fz, u, v, w, chi, K=1.0):
    check(lib().nekoamd_adjoint_lube(ptr(fx), ptr(fy), ptr(fz), ptr(u), ptr(v), ptr(w), ptr(chi),
                                     rd(K), ri64(fx.n)), "adjoint_lube")


def grad(case, u):
    gx, gy, gz = case.field(), case.field(), case.field()
    check(lib().nekoamd_grad(ptr(gx), ptr(gy), ptr(gz), ptr(u), ptr(case.x), ptr(case.y),
                             ptr(case.zc), ptr(case.dx), ri(case.lx), ri(case.nelv)), "grad")
    return gx, gy, gz


def dissipation_objective(case, u, v, w, chi, K=1.0, obj_scale=1.0, if_lube=True):
    out = np.zeros(3)
    check(lib().nekoamd_dissipation_objective(ptr(out), ptr(u), ptr(v), ptr(w), ptr(chi),
                                              ptr(case.B), ptr(case.x), ptr(case.y), ptr(case.zc),
                                              ptr(case.dx), ri(case.lx), ri(case.nelv), rd(K),
                                              rd(obj_scale), ri(int(if_lube))),
          "dissipation_objective")
    return out


def dissipation_objective_noalloc(case, u, v, w, chi, K=1.0, obj_scale=1.0, if_lube=True,
                                  nelv=None):
    """The allocation-free entry point (partials in the library's static
    scratch); nelv overrides the element count (the size check)."""
    out = np.zeros(3)
    ne = case.nelv if nelv is None else nelv
    check(lib().nekoamd_dissipation_objective_noalloc(
        ptr(out), ptr(u), ptr(v), ptr(w), ptr(chi), ptr(case.B), ptr(case.x), ptr(case.y),
        ptr(case.zc), ptr(case.dx), ri(case.lx), ri(ne), rd(K), rd(obj_scale),
        ri(int(if_lube))), "dissipation_objective_noalloc")
    return out


def volume_constraint(rho, B, volume_domain, v_max, is_max=False):
    out = np.zeros(2)
    check(lib().nekoamd_volume_constraint(ptr(out), ptr(rho), ptr(B), rd(volume_domain), rd(v_max),
                                          ri(int(is_max)), ri64(rho.n)), "volume_constraint")
    return out


def volume_sensitivity(s, volume_domain, is_max=False):
    check(lib().nekoamd_volume_sensitivity(ptr(s), rd(volume_domain), ri(int(is_max)),
                                           ri64(s.n)), "volume_sensitivity")


class MMA:
    """Device MMA (mma_t): x, df0dx, dfdx are DeviceArrays."""

    def __init__(self, x, m, a0, a, c, d, xmin, xmax, max_iter=100, epsimin=0.0,
                 asyinit=0.5, asyincr=1.2, asydecr=0.7):
        dev.ensure()
        self.n, self.m = x.n, m
        h = C.c_void_p(0)
        a_, c_, d_ = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, c, d))
        self._keep = (a_, c_, d_)
        check(lib().nekoamd_mma_create(C.byref(h), ri(self.n), ri(m), ptr(x), rd(a0), ptr(a_),
                                       ptr(c_), ptr(d_), ptr(xmin), ptr(xmax), ri(max_iter),
                                       rd(epsimin), rd(asyinit), rd(asyincr), rd(asydecr)),
              "mma_create")
        self.h = h.value

    def update(self, it, x, df0dx, fval, dfdx):
        f = np.ascontiguousarray(fval, dtype=np.float64)
        check(lib().nekoamd_mma_update(C.c_void_p(self.h), ri(it), ptr(x), ptr(df0dx), ptr(f),
                                       ptr(dfdx)), "mma_update")

    def kkt(self, x, df0dx, fval, dfdx):
        f = np.ascontiguousarray(fval, dtype=np.float64)
        check(lib().nekoamd_mma_kkt(C.c_void_p(self.h), ptr(x), ptr(df0dx), ptr(f), ptr(dfdx)),
              "mma_kkt")
        r = self.residuals()
        return r[0], r[1]

    def residuals(self):
        r = np.zeros(3)
        check(lib().nekoamd_mma_residuals(C.c_void_p(self.h), ptr(r)), "mma_residuals")
        return r

    def duals(self):
        m = self.m
        y, lam, s, mu, zs = np.zeros(m), np.zeros(m), np.zeros(m), np.zeros(m), np.zeros(2)
        check(lib().nekoamd_mma_duals(C.c_void_p(self.h), ptr(y), ptr(lam), ptr(s), ptr(mu),
                                      ptr(zs)), "mma_duals")
        return dict(y=y, lam=lam, s=s, mu=mu, z=zs[0], zeta=zs[1])

    def __del__(self):
        try:
            lib().nekoamd_mma_free(C.c_void_p(self.h))
        except Exception:
            pass
