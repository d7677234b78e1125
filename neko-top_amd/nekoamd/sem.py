"""The SEM discretisation on the device, built through the C-ABI only.

`BoxCase` is the synthetic channel of the workloads: an nx*ny*nz hex box
[0,nx/ny]x[0,1]x[0,nz/ny] with seeded vertex jitter, GLL order lx-1.  It
holds the device fields the operator core consumes (Neko's coef_t G11..G33,
B, mult; the GLL derivative matrix; the gs map; Dirichlet masks) and thin
wrappers for ax_helm, gs_op, Jacobi and CG.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as dev
from ._lib import (FACE_XMAX, FACE_XMIN, FACE_YMAX, FACE_YMIN, FACE_ZMAX, FACE_ZMIN,
                   GS_ADD, CgMonitor, check, lib, ptr, rd, ri, ri64)

VEL_DIRICHLET = FACE_XMIN | FACE_YMIN | FACE_YMAX | FACE_ZMIN | FACE_ZMAX
PRS_DIRICHLET = FACE_XMAX


def gll(lx: int):
    z, w, D = np.zeros(lx), np.zeros(lx), np.zeros(lx * lx)
    check(lib().nekoamd_gll(ri(lx), ptr(z), ptr(w), ptr(D)), "gll")
    return z, w, D.reshape(lx, lx)


def box_vertices(nx, ny, nz, jitter, seed, L=None):
    Lxyz = np.array(L if L is not None else (nx / ny, 1.0, nz / ny), dtype=np.float64)
    v = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    check(lib().nekoamd_box_vertices(ri(nx), ri(ny), ri(nz), ptr(Lxyz), rd(jitter),
                                      C.byref(C.c_uint64(seed)), ptr(v)), "box_vertices")
    return v


def box_gids(nx, ny, nz, lx):
    g = np.zeros(nx * ny * nz * lx ** 3, dtype=np.int64)
    check(lib().nekoamd_box_gids(ri(nx), ri(ny), ri(nz), ri(lx), ptr(g)), "box_gids")
    return g


def face_mask(nx, ny, nz, lx, faces):
    cnt = C.c_int64(0)
    check(lib().nekoamd_box_face_mask(ri(nx), ri(ny), ri(nz), ri(lx), ri(faces), None,
                                      C.byref(cnt)), "face_mask")
    out = np.zeros(max(cnt.value, 1), dtype=np.int32)
    check(lib().nekoamd_box_face_mask(ri(nx), ri(ny), ri(nz), ri(lx), ri(faces), ptr(out),
                                      C.byref(cnt)), "face_mask")
    return out[:cnt.value]


class GatherScatter:
    def __init__(self, gid: np.ndarray):
        g = np.ascontiguousarray(gid, dtype=np.int64)
        h = C.c_void_p(0)
        check(lib().nekoamd_gs_init(C.byref(h), ptr(g), ri64(g.size)), "gs_init")
        self.h = h.value

    def op(self, u: dev.DeviceArray, op=GS_ADD):
        fn = lib().nekoamd_gs_op_f64 if u.dtype == np.float64 else lib().nekoamd_gs_op_f32
        check(fn(C.c_void_p(self.h), ptr(u), ri(op)), "gs_op")

    def maps(self):
        nb = C.c_int(0)
        nocc = C.c_int64(0)
        bm = np.zeros(64, dtype=np.intc)
        bc = np.zeros(64, dtype=np.int64)
        check(lib().nekoamd_gs_map_info(C.c_void_p(self.h), C.byref(nb), ptr(bm), ptr(bc),
                                        C.byref(nocc)), "gs_map_info")
        idx = np.zeros(max(nocc.value, 1), dtype=np.int32)
        check(lib().nekoamd_gs_map_copy(C.c_void_p(self.h), ptr(idx)), "gs_map_copy")
        return idx[:nocc.value], bm[:nb.value].copy(), bc[:nb.value].copy()

    def __del__(self):
        try:
            lib().nekoamd_gs_free(C.c_void_p(self.h))
        except Exception:
            pass


class BoxCase:
    def __init__(self, nx, ny, nz, lx, jitter=0.05, seed=1234, L=None, keep_xyz=True):
        dev.ensure()
        self.nx, self.ny, self.nz, self.lx = nx, ny, nz, lx
        self.nelv = nx * ny * nz
        self.n = self.nelv * lx ** 3
        n = self.n
        z, w, D = gll(lx)
        self.zgll, self.wgll, self.D = z, w, D
        self.dx = dev.DeviceArray.from_host(D.ravel())
        zd = dev.DeviceArray.from_host(z)
        wd = dev.DeviceArray.from_host(w)
        verts = box_vertices(nx, ny, nz, jitter, seed, L)
        corners = np.zeros(self.nelv * 24)
        check(lib().nekoamd_box_corners(ri(nx), ri(ny), ri(nz), ptr(verts), ptr(corners)),
              "box_corners")
        self.verts = verts
        self.corners = corners
        cd = dev.DeviceArray.from_host(corners)
        self.G = {k: dev.DeviceArray(n) for k in ("G11", "G12", "G13", "G22", "G23", "G33")}
        self.B = dev.DeviceArray(n)
        self.jac = dev.DeviceArray(n)
        if keep_xyz:
            self.x, self.y, self.zc = dev.DeviceArray(n), dev.DeviceArray(n), dev.DeviceArray(n)
        else:
            self.x = self.y = self.zc = None
        G = self.G
        check(lib().nekoamd_coef_geo(ptr(self.x), ptr(self.y), ptr(self.zc), ptr(G["G11"]),
                                     ptr(G["G12"]), ptr(G["G13"]), ptr(G["G22"]), ptr(G["G23"]),
                                     ptr(G["G33"]), ptr(self.B), ptr(self.jac), ptr(cd), ptr(zd),
                                     ptr(wd), ptr(self.dx), ri(lx), ri(self.nelv)), "coef_geo")
        del cd, zd, wd
        self.gid = box_gids(nx, ny, nz, lx)
        self.gs = GatherScatter(self.gid)
        # mult = 1 / gs_ADD(1)
        self.mult = dev.DeviceArray(n)
        check(lib().nekoamd_cfill(ptr(self.mult), rd(1.0), ri64(n)), "cfill")
        self.gs.op(self.mult)
        check(lib().nekoamd_invcol1(ptr(self.mult), ri64(n)), "invcol1")
        ones = dev.DeviceArray(n)
        check(lib().nekoamd_cfill(ptr(ones), rd(1.0), ri64(n)), "cfill")
        self.volume = self.glsc2(self.B, ones)
        del ones

    # ---- field helpers ----------------------------------------------------
    def field(self, values=None, dtype=np.float64) -> dev.DeviceArray:
        if values is None:
            return dev.zeros(self.n, dtype)
        return dev.DeviceArray.from_host(np.broadcast_to(values, (self.n,)).astype(dtype))

    def mask(self, faces) -> dev.DeviceArray:
        m = face_mask(self.nx, self.ny, self.nz, self.lx, faces)
        d = dev.DeviceArray.from_host(m) if m.size else dev.DeviceArray(0, np.int32)
        d.count = m.size
        return d

    def glsc2(self, a, b) -> float:
        r = C.c_double(0)
        check(lib().nekoamd_glsc2(C.byref(r), ptr(a), ptr(b), ri64(self.n)), "glsc2")
        return r.value

    def glsc3(self, a, b, c) -> float:
        r = C.c_double(0)
        check(lib().nekoamd_glsc3(C.byref(r), ptr(a), ptr(b), ptr(c), ri64(self.n)), "glsc3")
        return r.value

    # ---- operators --------------------------------------------------------
    def ax(self, w, u, h1, h2=None, ifh2=False, variant="f64"):
        G = self.G
        fn = {"f64": lib().nekoamd_ax_helm_f64, "f32": lib().nekoamd_ax_helm_f32,
              "f32g32": lib().nekoamd_ax_helm_f32g32}[variant]
        Gs = G if variant != "f32g32" else self.G32
        B = self.B if variant != "f32g32" else self.B32
        check(fn(ptr(w), ptr(u), ptr(self.dx), ptr(h1), ptr(h2), ptr(Gs["G11"]), ptr(Gs["G12"]),
                 ptr(Gs["G13"]), ptr(Gs["G22"]), ptr(Gs["G23"]), ptr(Gs["G33"]), ptr(B),
                 ri(self.lx), ri(self.nelv), ri(1 if ifh2 else 0)), "ax_helm")

    def ax_vector(self, w3, u3, h1, h2=None, ifh2=False):
        """w3[c] = A u3[c] for three FP64 fields in one launch
        (nekoamd_ax_helm_vector_f64, Neko's ax_t%compute_vector)."""
        G = self.G
        check(lib().nekoamd_ax_helm_vector_f64(
            ptr(w3[0]), ptr(w3[1]), ptr(w3[2]), ptr(u3[0]), ptr(u3[1]), ptr(u3[2]),
            ptr(self.dx), ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]),
            ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]), ptr(self.B), ri(self.lx),
            ri(self.nelv), ri(1 if ifh2 else 0)), "ax_helm_vector")

    def make_f32_geometry(self):
        self.G32 = {}
        for k, v in self.G.items():
            a = dev.DeviceArray(self.n, np.float32)
            check(lib().nekoamd_f64_to_f32(ptr(a), ptr(v), ri64(self.n)), "f64_to_f32")
            self.G32[k] = a
        self.B32 = dev.DeviceArray(self.n, np.float32)
        check(lib().nekoamd_f64_to_f32(ptr(self.B32), ptr(self.B), ri64(self.n)), "f64_to_f32")

    def jacobi(self, d, h1, h2=None, ifh2=False):
        G = self.G
        check(lib().nekoamd_jacobi_update(ptr(d), C.c_void_p(self.gs.h), ptr(self.dx),
                                          ptr(G["G11"]), ptr(G["G22"]), ptr(G["G33"]), ptr(h1),
                                          ptr(h2), ptr(self.B), ri(self.lx), ri(self.nelv),
                                          ri(1 if ifh2 else 0)), "jacobi_update")


class Dealias:
    """Dealiased advection on lxd = 3 lx / 2 Gauss-Legendre points per
    direction: the state of adv_lin_dealias_t
    (sources/adjoint/adv_adjoint_dealias.f90:138-213) and its operators.
    Every operator subtracts from f (forward: f = beta f - ...)."""

    def __init__(self, case: BoxCase):
        if case.x is None:
            raise ValueError("dealiasing needs a case that keeps its coordinates")
        self.case = case
        self.lxd = 3 * case.lx // 2
        h = C.c_void_p()
        check(lib().nekoamd_dealias_create(C.byref(h), ptr(case.x), ptr(case.y), ptr(case.zc),
                                           ptr(case.dx), ri(case.lx), ri(self.lxd),
                                           ri(case.nelv)), "dealias_create")
        self.h = h.value

    def metric(self) -> np.ndarray:
        out = dev.DeviceArray(9 * self.case.nelv * self.lxd ** 3)
        check(lib().nekoamd_dealias_metric(C.c_void_p(self.h), ptr(out)), "dealias_metric")
        return out.get().reshape(9, -1)

    def advection(self, f, u, beta=0.0):
        check(lib().nekoamd_dealias_advection(C.c_void_p(self.h), *map(ptr, f), *map(ptr, u),
                                              rd(beta)), "dealias_advection")

    def linear(self, f, up, ub):
        check(lib().nekoamd_dealias_linear(C.c_void_p(self.h), *map(ptr, f), *map(ptr, up),
                                           *map(ptr, ub)), "dealias_linear")

    def adjoint(self, f, ua, ub):
        check(lib().nekoamd_dealias_adjoint(C.c_void_p(self.h), *map(ptr, f), *map(ptr, ua),
                                            *map(ptr, ub)), "dealias_adjoint")

    def __del__(self):
        try:
            lib().nekoamd_dealias_free(C.c_void_p(self.h))
        except Exception:
            pass


class CG:
    """Jacobi-PCG on the device; the library state owns the work vectors.
    mixed: FP32 iterations with FP64 reliable updates."""

    FP64, MIXED = 0, 1

    def __init__(self, case: BoxCase, mixed=False):
        self.case = case
        h = C.c_void_p(0)
        check(lib().nekoamd_cg_create(C.byref(h), ri64(case.n), ri(case.nelv),
                                      ri(self.MIXED if mixed else self.FP64)), "cg_create")
        self.h = h.value

    def solve(self, x, f, d, mask, h1, h2=None, ifh2=False, abstol=1e-8, max_iter=800,
              poll=8) -> CgMonitor:
        c = self.case
        G = c.G
        mon = CgMonitor()
        check(lib().nekoamd_cg_solve(
            C.c_void_p(self.h), ptr(x), ptr(f), ptr(d), ptr(c.mult), rd(c.volume), rd(abstol),
            ri(max_iter), ri(poll), C.c_void_p(c.gs.h),
            ptr(mask), ri64(getattr(mask, "count", mask.n if mask is not None else 0)),
            ptr(c.dx), ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]),
            ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]), ptr(c.B), ri(c.lx),
            ri(1 if ifh2 else 0), C.byref(mon)), "cg_solve")
        return mon

    @staticmethod
    def solve3(cgs, xs, fs, d, mask, h1, h2=None, ifh2=False, abstol=1e-8, max_iter=800,
               poll=8):
        """The coupled solve (nekoamd_cg_solve3) of three systems with one
        operator on three CG states; returns the three monitors."""
        c = cgs[0].case
        G = c.G
        mons = (CgMonitor * 3)()
        P3 = C.c_void_p * 3
        check(lib().nekoamd_cg_solve3(
            P3(*[g.h for g in cgs]), P3(*[ptr(x) for x in xs]), P3(*[ptr(f) for f in fs]),
            ptr(d), ptr(c.mult), rd(c.volume), rd(abstol), ri(max_iter), ri(poll),
            C.c_void_p(c.gs.h), ptr(mask),
            ri64(getattr(mask, "count", mask.n if mask is not None else 0)),
            ptr(c.dx), ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]),
            ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]), ptr(c.B), ri(c.lx),
            ri(1 if ifh2 else 0), mons), "cg_solve3")
        return list(mons)

    def __del__(self):
        try:
            lib().nekoamd_cg_free(C.c_void_p(self.h))
        except Exception:
            pass


class FCG:
    """Flexible PCG (Polak-Ribiere beta) on the device: Jacobi, plus the
    element network's Galerkin-weighted correction every `every`-th
    iteration when an MLP is given (nekoamd_fcg_solve)."""

    def __init__(self, case: BoxCase):
        self.case = case
        h = C.c_void_p(0)
        check(lib().nekoamd_fcg_create(C.byref(h), ri64(case.n), ri(case.nelv)), "fcg_create")
        self.h = h.value

    def solve(self, x, f, d, mask, h1, h2=None, ifh2=False, abstol=1e-8, max_iter=800,
              poll=8, mlp=None, every=1) -> CgMonitor:
        c = self.case
        G = c.G
        mon = CgMonitor()
        check(lib().nekoamd_fcg_solve(
            C.c_void_p(self.h), ptr(x), ptr(f), ptr(d), ptr(c.mult), rd(c.volume), rd(abstol),
            ri(max_iter), ri(poll), C.c_void_p(c.gs.h),
            ptr(mask), ri64(getattr(mask, "count", mask.n if mask is not None else 0)),
            ptr(c.dx), ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]),
            ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]), ptr(c.B), ri(c.lx),
            ri(1 if ifh2 else 0), C.c_void_p(mlp.h if mlp is not None else 0), ri(every),
            C.byref(mon)), "fcg_solve")
        return mon

    def __del__(self):
        try:
            lib().nekoamd_fcg_free(C.c_void_p(self.h))
        except Exception:
            pass


class GMRES:
    """Restarted flexible GMRES(m) on the device (Neko gmres_t).  The
    preconditioner is None (identity), a device array (Jacobi inverse
    diagonal) or an HSMG object."""

    IDENT, JACOBI, HSMG = 0, 1, 2

    def __init__(self, case: BoxCase, m=30):
        self.case = case
        h = C.c_void_p(0)
        check(lib().nekoamd_gmres_create(C.byref(h), ri64(case.n), ri(m)), "gmres_create")
        self.h = h.value

    def solve(self, x, f, pc, mask, h1, h2=None, ifh2=False, abstol=1e-8,
              max_iter=800) -> CgMonitor:
        c = self.case
        G = c.G
        if pc is None:
            kind, p = self.IDENT, None
        elif isinstance(pc, HSMG):
            kind, p = self.HSMG, C.c_void_p(pc.h)
        else:
            kind, p = self.JACOBI, ptr(pc)
        mon = CgMonitor()
        check(lib().nekoamd_gmres_solve(
            C.c_void_p(self.h), ptr(x), ptr(f), ri(kind), p, ptr(c.mult), rd(c.volume),
            rd(abstol), ri(max_iter), C.c_void_p(c.gs.h),
            ptr(mask), ri64(getattr(mask, "count", mask.n if mask is not None else 0)),
            ptr(c.dx), ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]),
            ptr(G["G22"]), ptr(G["G23"]), ptr(G["G33"]), ptr(c.B), ri(c.lx),
            ri(1 if ifh2 else 0), C.byref(mon)), "gmres_solve")
        return mon

    def __del__(self):
        try:
            lib().nekoamd_gmres_free(C.c_void_p(self.h))
        except Exception:
            pass


class HSMG:
    """Hybrid Schwarz multigrid preconditioner (nekoamd_hsmg_*) of a BoxCase:
    levels lx, 4, 2; dirichlet_faces of the preconditioned operator."""

    def __init__(self, case: BoxCase, dirichlet_faces=PRS_DIRICHLET, crs_iter=10, mixed=False):
        """mixed: FP32 data path (the coarse PCG stays FP64)."""
        self.case = case
        h = C.c_void_p(0)
        corners = np.ascontiguousarray(case.corners, dtype=np.float64)
        check(lib().nekoamd_hsmg_create(C.byref(h), ri(case.nx), ri(case.ny), ri(case.nz),
                                        ri(case.lx), ptr(corners), ri(dirichlet_faces),
                                        ri(crs_iter), ri(1 if mixed else 0)), "hsmg_create")
        self.h = h.value

    def apply(self, z, r):
        check(lib().nekoamd_hsmg_apply(C.c_void_p(self.h), ptr(z), ptr(r)), "hsmg_apply")

    def __del__(self):
        try:
            lib().nekoamd_hsmg_free(C.c_void_p(self.h))
        except Exception:
            pass


class Projection:
    """Previous-solution projection (nekoamd_proj_*) on a BoxCase."""

    def __init__(self, case: BoxCase, L: int, keep: int = 0, fp32: bool = False):
        self.case = case
        h = C.c_void_p(0)
        check(lib().nekoamd_proj_create(C.byref(h), ri64(case.n), ri(L)), "proj_create")
        self.h = h.value
        if keep:
            check(lib().nekoamd_proj_set_keep(C.c_void_p(self.h), ri(keep)), "proj_set_keep")
        if fp32:
            check(lib().nekoamd_proj_set_fp32(C.c_void_p(self.h), ri(1)), "proj_set_fp32")

    @property
    def m(self):
        m = C.c_int(0)
        check(lib().nekoamd_proj_size(C.c_void_p(self.h), C.byref(m)), "proj_size")
        return m.value

    def clear(self):
        check(lib().nekoamd_proj_clear(C.c_void_p(self.h)), "proj_clear")

    def pre(self, b):
        check(lib().nekoamd_proj_pre(C.c_void_p(self.h), ptr(b), ptr(self.case.mult)),
              "proj_pre")

    def augment(self, b, q, aq):
        g = C.c_double(0.0)
        check(lib().nekoamd_proj_augment(C.c_void_p(self.h), ptr(b), ptr(q), ptr(aq),
                                         ptr(self.case.mult), C.byref(g)), "proj_augment")
        return g.value

    def post(self, x, mask, h1, h2=None, ifh2=False):
        c = self.case
        G = c.G
        check(lib().nekoamd_proj_post(
            C.c_void_p(self.h), ptr(x), ptr(c.mult), C.c_void_p(c.gs.h), ptr(mask),
            ri64(getattr(mask, "count", mask.n if mask is not None else 0)), ptr(c.dx),
            ptr(h1), ptr(h2), ptr(G["G11"]), ptr(G["G12"]), ptr(G["G13"]), ptr(G["G22"]),
            ptr(G["G23"]), ptr(G["G33"]), ptr(c.B), ri(c.lx), ri(c.nelv),
            ri(1 if ifh2 else 0)), "proj_post")

    def __del__(self):
        try:
            lib().nekoamd_proj_free(C.c_void_p(self.h))
        except Exception:
            pass
