"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only tests/, the smoke
check and the CPU-baseline leg of bench.py import this module; the product
path (neko-top_amd/) never does.

`build_case` assembles the same synthetic box case the product builds (box
mesh with seeded jitter, GLL space, dofmap ids, geometric factors, gs maps,
multiplicity, Dirichlet masks), entirely on the CPU, so parity tests can
compare every intermediate.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GS_ADD, GS_MUL, GS_MIN, GS_MAX = 1, 2, 3, 4
# face bits: x=0, x=L, y=0, y=L, z=0, z=L
F_XMIN, F_XMAX, F_YMIN, F_YMAX, F_ZMIN, F_ZMAX = 1, 2, 4, 8, 16, 32
VEL_DIRICHLET = F_XMIN | F_YMIN | F_YMAX | F_ZMIN | F_ZMAX  # inflow + walls
PRS_DIRICHLET = F_XMAX  # outflow


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_intp = np.ctypeslib.ndpointer(dtype=np.intc, flags="C_CONTIGUOUS")


class SemT(C.Structure):
    _fields_ = [
        ("lx", C.c_int), ("nelv", C.c_int), ("n", C.c_int64),
        ("D", C.c_void_p), ("G11", C.c_void_p), ("G12", C.c_void_p),
        ("G13", C.c_void_p), ("G22", C.c_void_p), ("G23", C.c_void_p),
        ("G33", C.c_void_p), ("B", C.c_void_p), ("mult", C.c_void_p),
        ("gs_nb", C.c_int), ("gs_m", C.c_void_p), ("gs_cnt", C.c_void_p),
        ("gs_idx", C.c_void_p), ("nmask", C.c_int64), ("mask", C.c_void_p),
        ("volume", C.c_double),
    ]


def _declare(L):
    L.or_splitmix64.restype = C.c_uint64
    L.or_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
    L.or_u01.restype = C.c_double
    L.or_u01.argtypes = [C.c_uint64, C.c_uint64]
    L.or_set_threads.argtypes = [C.c_int]
    L.or_get_threads.restype = C.c_int
    L.or_gll.argtypes = [C.c_int, _dp, _dp, _dp]
    L.or_box_vertices.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_uint64, _dp]
    L.or_element_corners.argtypes = [C.c_int] * 3 + [_dp, _dp]
    L.or_dofmap_gids.argtypes = [C.c_int] * 4 + [_i64p]
    L.or_box_rng_uniform.argtypes = [C.c_int] * 4 + [C.c_uint64, _dp]
    L.or_face_mask.restype = C.c_int64
    L.or_face_mask.argtypes = [C.c_int] * 5 + [C.c_void_p]
    L.or_coef.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp] + [_dp] * 10 + [C.c_void_p]
    L.or_ax_helm.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp] + [_dp] * 7 + [C.c_int]
    L.or_gs_build.restype = C.c_int64
    L.or_gs_build.argtypes = [C.c_int64, _i64p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_gs_op.argtypes = [_dp, C.c_int, _intp, _i64p, _i32p, C.c_int]
    L.or_jacobi_update.argtypes = [C.POINTER(SemT), _dp, _dp, C.c_int, _dp]
    L.or_glsc3.restype = C.c_double
    L.or_glsc3.argtypes = [C.c_int64, _dp, _dp, _dp]
    L.or_glsc2.restype = C.c_double
    L.or_glsc2.argtypes = [C.c_int64, _dp, _dp]
    L.or_cg_solve.argtypes = [C.POINTER(SemT), _dp, _dp, C.c_int, _dp, _dp, _dp,
                              C.c_double, C.c_int, C.POINTER(C.c_int),
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _declare_topopt(L)


def _declare_topopt(L):
    """Signatures of the topology-optimisation and MMA restatements, when built."""
    try:
        from . import _topopt_sigs  # noqa: F401
        _topopt_sigs.declare(L, _dp, _i32p)
        _topopt_sigs.declare_pnpn(L, _dp)
    except ImportError:
        pass


def set_threads(t: int):
    lib().or_set_threads(int(t))


def u01(seed: int, idx: int) -> float:
    return lib().or_u01(seed, idx)


def gll(lx: int):
    z = np.zeros(lx); w = np.zeros(lx); D = np.zeros(lx * lx)
    lib().or_gll(lx, z, w, D)
    return z, w, D.reshape(lx, lx)


def box_vertices(nx, ny, nz, L=None, jitter=0.05, seed=1234):
    Lx, Ly, Lz = L if L is not None else (nx / ny, 1.0, nz / ny)
    v = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    lib().or_box_vertices(nx, ny, nz, Lx, Ly, Lz, jitter, seed, v)
    return v


def face_mask(nx, ny, nz, lx, faces):
    L = lib()
    cnt = L.or_face_mask(nx, ny, nz, lx, faces, None)
    out = np.zeros(max(cnt, 1), dtype=np.int32)
    L.or_face_mask(nx, ny, nz, lx, faces, out.ctypes.data)
    return out[:cnt]


def gs_build(gid):
    L = lib()
    n = gid.size
    tot = L.or_gs_build(n, gid, None, None, None, None)
    idx = np.zeros(max(tot, 1), dtype=np.int32)
    nb = C.c_int(0)
    bm = np.zeros(64, dtype=np.intc)
    bc = np.zeros(64, dtype=np.int64)
    L.or_gs_build(n, gid, idx.ctypes.data, C.addressof(nb), bm.ctypes.data, bc.ctypes.data)
    k = nb.value
    return idx[:tot].copy(), bm[:k].copy(), bc[:k].copy()


@dataclass
class Case:
    nx: int
    ny: int
    nz: int
    lx: int
    jitter: float
    seed: int
    z: np.ndarray = None
    w: np.ndarray = None
    D: np.ndarray = None
    verts: np.ndarray = None
    corners: np.ndarray = None
    gid: np.ndarray = None
    x: np.ndarray = None
    y: np.ndarray = None
    zc: np.ndarray = None
    G: dict = field(default_factory=dict)
    B: np.ndarray = None
    jac: np.ndarray = None
    gs_idx: np.ndarray = None
    gs_m: np.ndarray = None
    gs_cnt: np.ndarray = None
    mult: np.ndarray = None
    volume: float = 0.0

    @property
    def nelv(self):
        return self.nx * self.ny * self.nz

    @property
    def n(self):
        return self.nelv * self.lx ** 3

    def gs(self, u, op=GS_ADD):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        lib().or_gs_op(u, len(self.gs_m), self.gs_m, self.gs_cnt, self.gs_idx, op)
        return u

    def ax(self, u, h1=None, h2=None, ifh2=False):
        n = self.n
        h1 = np.ones(n) if h1 is None else np.broadcast_to(h1, (n,)).astype(np.float64).copy()
        h2 = np.zeros(n) if h2 is None else np.broadcast_to(h2, (n,)).astype(np.float64).copy()
        w = np.zeros(n)
        G = self.G
        lib().or_ax_helm(self.lx, self.nelv, w, np.ascontiguousarray(u, dtype=np.float64),
                         np.ascontiguousarray(self.D.ravel()), h1, h2, G["G11"], G["G12"],
                         G["G13"], G["G22"], G["G23"], G["G33"], self.B, int(ifh2))
        return w

    def sem(self, mask=None):
        """or_sem_t view; keeps references alive on the returned object."""
        s = SemT()
        s.lx, s.nelv, s.n = self.lx, self.nelv, self.n
        keep = [np.ascontiguousarray(self.D.ravel())]
        s.D = keep[0].ctypes.data
        for k in ("G11", "G12", "G13", "G22", "G23", "G33"):
            setattr(s, k, self.G[k].ctypes.data)
        s.B = self.B.ctypes.data
        s.mult = self.mult.ctypes.data
        s.gs_nb = len(self.gs_m)
        s.gs_m = self.gs_m.ctypes.data
        s.gs_cnt = self.gs_cnt.ctypes.data
        s.gs_idx = self.gs_idx.ctypes.data
        m = np.zeros(1, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        keep.append(m)
        s.nmask = 0 if mask is None else m.size
        s.mask = m.ctypes.data
        s.volume = self.volume
        s._keep = keep
        return s

    def jacobi(self, h1=None, h2=None, ifh2=False):
        n = self.n
        h1 = np.ones(n) if h1 is None else np.broadcast_to(h1, (n,)).astype(np.float64).copy()
        h2 = np.zeros(n) if h2 is None else np.broadcast_to(h2, (n,)).astype(np.float64).copy()
        d = np.zeros(n)
        s = self.sem()
        lib().or_jacobi_update(C.byref(s), h1, h2, int(ifh2), d)
        return d

    def cg(self, f, mask, h1=None, h2=None, ifh2=False, abstol=1e-8, max_iter=800, dinv=None):
        n = self.n
        h1 = np.ones(n) if h1 is None else np.broadcast_to(h1, (n,)).astype(np.float64).copy()
        h2 = np.zeros(n) if h2 is None else np.broadcast_to(h2, (n,)).astype(np.float64).copy()
        if dinv is None:
            dinv = self.jacobi(h1, h2, ifh2)
        s = self.sem(mask)
        x = np.zeros(n)
        it = C.c_int(0); r0 = C.c_double(0); r1 = C.c_double(0)
        lib().or_cg_solve(C.byref(s), h1, h2, int(ifh2), dinv, x,
                          np.ascontiguousarray(f, dtype=np.float64), abstol, max_iter,
                          C.byref(it), C.byref(r0), C.byref(r1))
        return x, it.value, r0.value, r1.value


def build_case(nx, ny, nz, lx, jitter=0.05, seed=1234, L=None) -> Case:
    Lb = lib()
    c = Case(nx, ny, nz, lx, jitter, seed)
    c.z, c.w, c.D = gll(lx)
    c.verts = box_vertices(nx, ny, nz, L, jitter, seed)
    nelv = c.nelv
    c.corners = np.zeros(nelv * 24)
    Lb.or_element_corners(nx, ny, nz, c.verts, c.corners)
    n = c.n
    c.gid = np.zeros(n, dtype=np.int64)
    Lb.or_dofmap_gids(nx, ny, nz, lx, c.gid)
    arrs = {k: np.zeros(n) for k in ("x", "y", "zc", "G11", "G12", "G13", "G22", "G23", "G33", "B", "jac")}
    Lb.or_coef(lx, nelv, c.corners, c.z, c.w, np.ascontiguousarray(c.D.ravel()),
               arrs["x"], arrs["y"], arrs["zc"], arrs["G11"], arrs["G12"], arrs["G13"],
               arrs["G22"], arrs["G23"], arrs["G33"], arrs["B"], arrs["jac"].ctypes.data)
    c.x, c.y, c.zc, c.B, c.jac = arrs["x"], arrs["y"], arrs["zc"], arrs["B"], arrs["jac"]
    c.G = {k: arrs[k] for k in ("G11", "G12", "G13", "G22", "G23", "G33")}
    c.gs_idx, c.gs_m, c.gs_cnt = gs_build(c.gid)
    c.mult = 1.0 / c.gs(np.ones(n))
    c.volume = float(Lb.or_glsc2(n, c.B, np.ones(n)))
    return c


# ---------------------------------------------------------------------------
# topology optimisation and MMA (oracle/topopt.c, oracle/mma.c)

def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def brinkman(fx, fy, fz, u, v, w, chi):
    fx, fy, fz = _f(fx).copy(), _f(fy).copy(), _f(fz).copy()
    lib().or_brinkman(fx.size, fx, fy, fz, _f(u), _f(v), _f(w), _f(chi))
    return fx, fy, fz


def update_permeability(rho, perm_0=1000.0, perm_1=0.0, penalty=1.0):
    out = np.zeros(_f(rho).size)
    lib().or_update_permeability(out.size, out, _f(rho), perm_0, perm_1, penalty)
    return out


def ramp_fwd(x, f_min, f_max, q, convex_up=True):
    x = _f(x)
    out = np.zeros(x.size)
    lib().or_ramp_fwd(x.size, out, x, f_min, f_max, q, int(convex_up))
    return out


def ramp_bwd(df_dout, x, f_min, f_max, q, convex_up=True):
    x = _f(x)
    out = np.zeros(x.size)
    lib().or_ramp_bwd(x.size, out, _f(df_dout), x, f_min, f_max, q, int(convex_up))
    return out


def linear_fwd(x, f_min, f_max):
    x = _f(x)
    out = np.zeros(x.size)
    lib().or_linear_fwd(x.size, out, x, f_min, f_max)
    return out


def linear_bwd(df_dout, f_min, f_max):
    d = _f(df_dout)
    out = np.zeros(d.size)
    lib().or_linear_bwd(d.size, out, d, f_min, f_max)
    return out


def grad(case, u):
    n = case.n
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().or_grad(case.lx, case.nelv, np.ascontiguousarray(case.D.ravel()), case.x, case.y,
                  case.zc, _f(u), ux, uy, uz)
    return ux, uy, uz


def dissipation_objective(case, u, v, w, chi, K=1.0, obj_scale=1.0, if_lube=True):
    out = np.zeros(3)
    lib().or_dissipation_objective(case.lx, case.nelv, np.ascontiguousarray(case.D.ravel()),
                                   case.x, case.y, case.zc, case.B, _f(u), _f(v), _f(w),
                                   _f(chi), K, obj_scale, int(if_lube), out)
    return out


def dissipation_sensitivity(u, v, w, ua, va, wa, K=1.0, obj_scale=1.0, if_lube=True):
    u = _f(u)
    s = np.zeros(u.size)
    lib().or_dissipation_sensitivity(u.size, s, u, _f(v), _f(w), _f(ua), _f(va), _f(wa),
                                     K, obj_scale, int(if_lube))
    return s


def adjoint_lube(fx, fy, fz, u, v, w, chi, K=1.0):
    fx, fy, fz = _f(fx).copy(), _f(fy).copy(), _f(fz).copy()
    lib().or_adjoint_lube(fx.size, fx, fy, fz, _f(u), _f(v), _f(w), _f(chi), K)
    return fx, fy, fz


def volume_constraint(rho, B, volume_domain, v_max=float(np.float32(0.2)), is_max=False):
    out = np.zeros(2)
    lib().or_volume_constraint(_f(rho).size, _f(rho), _f(B), volume_domain, v_max,
                               int(is_max), out)
    return out


def volume_sensitivity(n, volume_domain, is_max=False):
    s = np.zeros(n)
    lib().or_volume_sensitivity(n, s, volume_domain, int(is_max))
    return s


FILTER_R = float(np.float32(0.01))  # the reference's default-real literal


def filter_apply(case, x, r=FILTER_R, abstol=1e-10, max_iter=200):
    s = case.sem()
    out = np.zeros(case.n)
    it = lib().or_filter_apply(C.addressof(s), r, abstol, max_iter, _f(x), out)
    return out, it


class MMA:
    """Handle on the oracle's MMA state (sources/mma/mma.f90 mma_t)."""

    def __init__(self, x, m, a0, a, c, d, xmin, xmax, max_iter=100, epsimin=0.0,
                 asyinit=0.5, asyincr=1.2, asydecr=0.7):
        x = _f(x)
        self.n, self.m = x.size, m
        self.h = lib().or_mma_create(self.n, m, x, a0, _f(a), _f(c), _f(d), _f(xmin), _f(xmax),
                                     max_iter, epsimin, asyinit, asyincr, asydecr)

    def update(self, it, x, df0dx, fval, dfdx):
        """dfdx: (m, n) array; x updated in place (returned)."""
        x = _f(x).copy()
        dfdx = np.asfortranarray(np.asarray(dfdx, dtype=np.float64).reshape(self.m, self.n))
        rc = lib().or_mma_update(self.h, it, x, _f(df0dx), _f(fval), _f(dfdx.T.ravel()))
        if rc:
            raise RuntimeError(f"MMA subsolve: DGESV failed ({rc})")
        return x

    def kkt(self, x, df0dx, fval, dfdx):
        dfdx = np.asarray(dfdx, dtype=np.float64).reshape(self.m, self.n)
        lib().or_mma_kkt(self.h, _f(x), _f(df0dx), _f(fval), _f(dfdx.T.ravel()))
        return lib().or_mma_residumax(self.h), lib().or_mma_residunorm(self.h)

    def subproblem(self):
        n, m = self.n, self.m
        arrs = [np.zeros(n) for _ in range(6)] + [np.zeros(m * n), np.zeros(m * n), np.zeros(m)]
        lib().or_mma_get_subproblem(self.h, *arrs)
        return dict(zip(["low", "upp", "alpha", "beta", "p0j", "q0j", "pij", "qij", "bi"], arrs))

    def duals(self):
        n, m = self.n, self.m
        arrs = [np.zeros(m) for _ in range(4)] + [np.zeros(n), np.zeros(n), np.zeros(2)]
        lib().or_mma_get_duals(self.h, *arrs)
        return dict(zip(["y", "lambda", "s", "mu", "xsi", "eta", "zzeta"], arrs))

    def __del__(self):
        try:
            lib().or_mma_free(self.h)
        except Exception:
            pass


def dgesv(A, b):
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    N = A.shape[0]
    a = A.T.ravel().copy() if False else np.ascontiguousarray(A.ravel(order="F"))
    bb = _f(b).copy()
    info = lib().or_dgesv(N, a, bb)
    return bb, info


# ---------------------------------------------------------------------------
# Pn/Pn step operators and the step (oracle/pnpn.c)

def _Dflat(c):
    return np.ascontiguousarray(c.D.ravel())


def advection(c, u, v, w):
    out = [np.zeros(c.n) for _ in range(3)]
    lib().or_advection(c.lx, c.nelv, _Dflat(c), c.x, c.y, c.zc, c.B, _f(u), _f(v), _f(w), *out)
    return out


def curl_local(c, u, v, w):
    out = [np.zeros(c.n) for _ in range(3)]
    lib().or_curl_local(c.lx, c.nelv, _Dflat(c), c.x, c.y, c.zc, _f(u), _f(v), _f(w), *out)
    return out


def opgrad(c, p):
    out = [np.zeros(c.n) for _ in range(3)]
    lib().or_opgrad(c.lx, c.nelv, _Dflat(c), c.w, c.x, c.y, c.zc, _f(p), *out)
    return out


def cdtp(c, tx, ty, tz):
    out = np.zeros(c.n)
    lib().or_cdtp(c.lx, c.nelv, _Dflat(c), c.w, c.x, c.y, c.zc, _f(tx), _f(ty), _f(tz), out)
    return out


def surface_flux(c, u, v, w, faces):
    out = np.zeros(c.n)
    lib().or_surface_flux(c.nx, c.ny, c.nz, c.lx, _Dflat(c), c.w, c.x, c.y, c.zc, _f(u), _f(v),
                          _f(w), faces, out)
    return out


def pnpn_coeffs(tstep, order=2):
    """BDF / EXT coefficients of step tstep for time_order 2 (bd[3], ext[2])
    or 3 (bd[4], ext[3])."""
    if order == 2:
        bd, ext = np.zeros(3), np.zeros(2)
        lib().or_pnpn_coeffs(tstep, bd, ext)
        return bd, ext
    bd, ext = np.zeros(4), np.zeros(3)
    lib().or_pnpn_coeffs3(tstep, order, bd, ext)
    return bd, ext


class PnpnT(C.Structure):
    _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lx", C.c_int),
                 ("nelv", C.c_int), ("n", C.c_int64)]
                + [(k, C.c_void_p) for k in ("D", "wq", "x", "y", "z", "B", "Binv", "mult", "chi",
                                             "G11", "G12", "G13", "G22", "G23", "G33")]
                + [("gs_nb", C.c_int), ("gs_m", C.c_void_p), ("gs_cnt", C.c_void_p),
                   ("gs_idx", C.c_void_p), ("volume", C.c_double), ("nmask_v", C.c_int64),
                   ("nmask_p", C.c_int64), ("mask_v", C.c_void_p), ("mask_p", C.c_void_p),
                   ("faces_v", C.c_int), ("Re", C.c_double), ("dt", C.c_double),
                   ("rho", C.c_double), ("abstol_p", C.c_double), ("abstol_v", C.c_double),
                   ("max_iter", C.c_int)]
                + [(k, C.c_void_p) for k in ("u", "v", "w", "p", "ulag", "vlag", "wlag",
                                             "fxlag", "fylag", "fzlag")]
                + [("tstep", C.c_int), ("iters", C.c_int * 4), ("res", C.c_double * 4),
                   ("adjoint", C.c_int), ("ub", C.c_void_p), ("vb", C.c_void_p),
                   ("wb", C.c_void_p), ("K", C.c_double), ("obj_scale", C.c_double),
                   ("dealias", C.c_int), ("lxd", C.c_int)]
                + [(k, C.c_void_p) for k in ("Jm", "Dg", "wg", "cof")]
                + [("plant", C.c_int)]
                + [(k, C.c_void_p) for k in ("ubc", "vbc", "wbc")]
                + [("brinkman_explicit", C.c_int), ("time_order", C.c_int)]
                + [(k, C.c_void_p) for k in ("ulag2", "vlag2", "wlag2", "fxlag2", "fylag2",
                                             "fzlag2")])


# planted mistakes of the step (oracle/pnpn.c OR_PLANT_*), for the tests that
# show the step's checks fail on them
PLANT_MU2, PLANT_CURLCURL_SIGN, PLANT_SWAP_BDF_LAGS, PLANT_SURF_DT = 1, 2, 4, 8
PLANT_ADV_SIGN, PLANT_CHI_PRS_OFF, PLANT_EXT_SWAP, PLANT_CHI_VEL_OFF = 16, 32, 64, 128
PLANT_BDF1_ONLY = 256


WALLS = F_YMIN | F_YMAX | F_ZMIN | F_ZMAX


def channel_initial_velocity(c):
    """u = (1, 0, 0) with u = 0 on the walls (uniform inflow, no-slip)."""
    u = np.ones(c.n)
    u[face_mask(c.nx, c.ny, c.nz, c.lx, WALLS)] = 0.0
    return u, np.zeros(c.n), np.zeros(c.n)


class Channel:
    """Oracle Brinkman channel: state + or_pnpn_step."""

    def __init__(self, c, chi, Re=100.0, dt=1e-3, abstol_p=1e-8, abstol_v=1e-8, max_iter=800,
                 dealias=False, vel_faces=None, prs_faces=None, plant=0,
                 brinkman_explicit=False, time_order=2):
        """vel_faces / prs_faces: the box faces (F_* bits) where the velocity
        / pressure are Dirichlet (default: the channel's inflow + walls /
        outflow).  Dirichlet values are whatever u, v, w, p hold on those
        faces: the solves keep them.  plant: OR_PLANT_* bits (0 = as
        written).  brinkman_explicit: the reference's explicit Brinkman
        term instead of the semi-implicit one.  time_order: 2 (BDF2/EXT2)
        or 3 (BDF3/EXT3 from the third step on)."""
        self.c = c
        n = c.n
        vel_faces = VEL_DIRICHLET if vel_faces is None else vel_faces
        prs_faces = PRS_DIRICHLET if prs_faces is None else prs_faces
        self.Binv = 1.0 / c.gs(c.B)
        self.chi = _f(chi).copy()
        self.mask_v = face_mask(c.nx, c.ny, c.nz, c.lx, vel_faces)
        self.mask_p = face_mask(c.nx, c.ny, c.nz, c.lx, prs_faces)
        self.u, self.v, self.w = channel_initial_velocity(c)
        self.p = np.zeros(n)
        self.lags = [np.zeros(n) for _ in range(6)]
        self.lags[0][:] = self.u
        self.D = _Dflat(c)
        s = PnpnT()
        s.nx, s.ny, s.nz, s.lx, s.nelv, s.n = c.nx, c.ny, c.nz, c.lx, c.nelv, n
        for k, a in (("D", self.D), ("wq", c.w), ("x", c.x), ("y", c.y), ("z", c.zc), ("B", c.B),
                     ("Binv", self.Binv), ("mult", c.mult), ("chi", self.chi)):
            setattr(s, k, a.ctypes.data)
        for k in ("G11", "G12", "G13", "G22", "G23", "G33"):
            setattr(s, k, c.G[k].ctypes.data)
        s.gs_nb, s.gs_m, s.gs_cnt, s.gs_idx = (len(c.gs_m), c.gs_m.ctypes.data,
                                               c.gs_cnt.ctypes.data, c.gs_idx.ctypes.data)
        s.volume = c.volume
        s.nmask_v, s.mask_v = self.mask_v.size, self.mask_v.ctypes.data
        s.nmask_p, s.mask_p = self.mask_p.size, self.mask_p.ctypes.data
        s.faces_v = vel_faces
        s.plant = plant
        s.brinkman_explicit = int(brinkman_explicit)
        s.time_order = int(time_order)
        self.lags2 = [np.zeros(n) for _ in range(6)] if time_order == 3 else None
        if self.lags2 is not None:
            self.lags2[0][:] = self.u
            for k, a in zip(("ulag2", "vlag2", "wlag2", "fxlag2", "fylag2", "fzlag2"),
                            self.lags2):
                setattr(s, k, a.ctypes.data)
        s.Re, s.dt, s.rho = Re, dt, 1.0
        s.abstol_p, s.abstol_v, s.max_iter = abstol_p, abstol_v, max_iter
        for k, a in zip(("u", "v", "w", "p"), (self.u, self.v, self.w, self.p)):
            setattr(s, k, a.ctypes.data)
        for k, a in zip(("ulag", "vlag", "wlag", "fxlag", "fylag", "fzlag"), self.lags):
            setattr(s, k, a.ctypes.data)
        s.tstep = 0
        if dealias:
            dl = dealias_setup(c)
            self._dl = (np.ascontiguousarray(dl.Jm.ravel()), np.ascontiguousarray(dl.Dg.ravel()),
                        dl.wg, dl.cof)
            s.dealias, s.lxd = 1, dl.lxd
            s.Jm, s.Dg, s.wg, s.cof = (a.ctypes.data for a in self._dl)
        self.s = s

    def make_adjoint(self, ub, vb, wb, K=1.0, obj_scale=1.0):
        """Turn this case into the adjoint solve about the frozen base flow
        (ub, vb, wb): adjoint velocity and pressure start at zero."""
        self.base = [_f(a).copy() for a in (ub, vb, wb)]
        for a in (self.u, self.v, self.w, self.p, *self.lags, *(self.lags2 or [])):
            a[:] = 0.0
        s = self.s
        s.adjoint = 1
        s.ub, s.vb, s.wb = (a.ctypes.data for a in self.base)
        s.K, s.obj_scale = K, obj_scale
        s.tstep = 0

    def set_velocity_bc(self, ub, vb, wb):
        """Dirichlet velocity values for the next step's new time level
        (read on the velocity mask only); None keeps the held values."""
        s = self.s
        if ub is None:
            s.ubc = s.vbc = s.wbc = None
            self._bc = None
            return
        self._bc = [_f(a).copy() for a in (ub, vb, wb)]
        s.ubc, s.vbc, s.wbc = (a.ctypes.data for a in self._bc)

    def step(self):
        lib().or_pnpn_step(C.addressof(self.s))
        return list(self.s.iters), list(self.s.res)

    @property
    def tstep(self):
        return self.s.tstep


def adjoint_advection(c, ua, va, wa, ub, vb, wb):
    """(fx, fy, fz) = the adjoint advection of (ua, va, wa) about (ub, vb, wb)
    added to zero forcing."""
    n = c.n
    f = [np.zeros(n) for _ in range(3)]
    work = np.zeros(12 * n)
    lib().or_adjoint_advection(c.lx, c.nelv, _Dflat(c), c.w, c.x, c.y, c.zc, _f(ua), _f(va),
                               _f(wa), _f(ub), _f(vb), _f(wb), f[0], f[1], f[2], work)
    return f


# ---------------------------------------------------------------------------
# Dealiased advection (oracle/dealias.c)

def gauss_legendre(n):
    z = np.zeros(n); w = np.zeros(n)
    lib().or_gauss_legendre(n, z, w)
    return z, w


def lagrange_matrix(zf, zt):
    """J[q, i] = l_i(zt[q]) for the Lagrange basis of the nodes zf."""
    zf = np.ascontiguousarray(zf, dtype=np.float64)
    zt = np.ascontiguousarray(zt, dtype=np.float64)
    J = np.zeros(zt.size * zf.size)
    lib().or_lagrange_matrix(zf.size, zf, zt.size, zt, J)
    return J.reshape(zt.size, zf.size)


def deriv_matrix(z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    D = np.zeros(z.size * z.size)
    lib().or_deriv_matrix(z.size, z, D)
    return D.reshape(z.size, z.size)


def cofactors(c):
    """[9, n]: c_b[a] = J d(r_b)/d(x_a), row m = 3b + a."""
    out = np.zeros(9 * c.n)
    lib().or_cofactors(c.lx, c.nelv, _Dflat(c), c.x, c.y, c.zc, out)
    return out.reshape(9, c.n)


@dataclass
class Dealias:
    """The GL space of one case: lxd points per direction (3 lx / 2 by
    default, advection_adjoint_fctry.f90:66-70), the GLL->GL interpolation
    Jm, the GL derivative matrix Dg, the GL weights wg and the GL metric."""
    lx: int
    lxd: int
    nelv: int
    Jm: np.ndarray
    Dg: np.ndarray
    wg: np.ndarray
    zg: np.ndarray
    cof: np.ndarray


def dealias_setup(c, lxd=None):
    lxd = lxd or 3 * c.lx // 2
    zg, wg = gauss_legendre(lxd)
    Jm = lagrange_matrix(c.z, zg)
    Dg = deriv_matrix(zg)
    cof = np.zeros(9 * c.nelv * lxd ** 3)
    lib().or_dealias_metric(c.lx, lxd, c.nelv, np.ascontiguousarray(Jm.ravel()),
                            np.ascontiguousarray(cofactors(c).ravel()), cof)
    return Dealias(c.lx, lxd, c.nelv, Jm, Dg, wg, zg, cof)


def _dl_args(dl):
    return (dl.lx, dl.lxd, dl.nelv, np.ascontiguousarray(dl.Jm.ravel()),
            np.ascontiguousarray(dl.Dg.ravel()), dl.wg, dl.cof)


def dealias_advection(dl, u, v, w, f=None):
    """f_c -= J^T [ t . (B grad t_c) ] (forward convective term)."""
    f = [np.zeros(u.size) for _ in range(3)] if f is None else [np.array(a, dtype=np.float64) for a in f]
    lib().or_dealias_advection(*_dl_args(dl), _f(u), _f(v), _f(w), *f)
    return f


def dealias_linear(dl, up, vp, wp, ub, vb, wb, f=None):
    f = [np.zeros(up.size) for _ in range(3)] if f is None else [np.array(a, dtype=np.float64) for a in f]
    lib().or_dealias_linear(*_dl_args(dl), _f(up), _f(vp), _f(wp), _f(ub), _f(vb), _f(wb), *f)
    return f


def dealias_adjoint(dl, ua, va, wa, ub, vb, wb, f=None):
    f = [np.zeros(ua.size) for _ in range(3)] if f is None else [np.array(a, dtype=np.float64) for a in f]
    lib().or_dealias_adjoint(*_dl_args(dl), _f(ua), _f(va), _f(wa), _f(ub), _f(vb), _f(wb), *f)
    return f


def box_rng_uniform(c, seed):
    out = np.zeros(c.n)
    lib().or_box_rng_uniform(c.nx, c.ny, c.nz, c.lx, seed, out)
    return out


def seeded_design(c, seed, npass=3, f_min=1e4, f_max=0.0, q=0.1, abstol=1e-10, max_iter=200):
    """rho ~ U[0,1) per global point, npass filter passes, chi = RAMP."""
    rho = box_rng_uniform(c, seed)
    rf = rho
    for _ in range(npass):
        rf, _ = filter_apply(c, rf, abstol=abstol, max_iter=max_iter)
    return rho, rf, ramp_fwd(rf, f_min, f_max, q, True)


class TopOpt:
    """The oracle's optimisation loop (examples/rugby_ball/driver.f90:69-138
    with sources/problems/steady_state_problem.f90:273-353): forward Pn/Pn
    solve, objective and volume, adjoint solve about the forward flow,
    dF/dchi, chain rule through RAMP and the filter, MMA update with the
    constraint scaled by 100, KKT, map_forward.  The design indicator starts
    as seeded noise smoothed by npass-1 filter passes."""

    def __init__(self, c, design_seed=4321, npass=3, f_min=1e4, f_max=0.0, q=0.1, v_max=0.4,
                 is_max=False, K=1.0, obj_scale=1.0, Re=100.0, dt=1e-3, abstol_p=1e-8,
                 abstol_v=1e-8, max_iter=800, filter_tol=1e-10):
        self.c = c
        self.f_min, self.f_max, self.q = f_min, f_max, q
        self.v_max, self.is_max, self.K, self.obj_scale = v_max, is_max, K, obj_scale
        self.kw = dict(Re=Re, dt=dt, abstol_p=abstol_p, abstol_v=abstol_v, max_iter=max_iter)
        self.ftol = filter_tol
        x = box_rng_uniform(c, design_seed)
        for _ in range(npass - 1):
            x, _ = filter_apply(c, x, abstol=filter_tol)
        self.x = x
        self._map_forward()
        self.vol_dom = c.volume
        self.mma = MMA(self.x, 1, 1.0, [0.0], [100.0], [0.0], np.zeros(c.n), np.ones(c.n),
                       max_iter=100)
        self.it = 1

    def _map_forward(self):
        self.rf, _ = filter_apply(self.c, self.x, abstol=self.ftol)
        self.chi = ramp_fwd(self.rf, self.f_min, self.f_max, self.q, True)

    def iterate(self, nfwd, nadj):
        c = self.c
        fwd = Channel(c, self.chi, **self.kw)
        pf = sum(fwd.step()[0][0] for _ in range(nfwd))
        obj = dissipation_objective(c, fwd.u, fwd.v, fwd.w, self.chi, self.K, self.obj_scale,
                                    True)
        vol = volume_constraint(self.x, c.B, self.vol_dom, self.v_max, self.is_max)
        adj = Channel(c, self.chi, **self.kw)
        adj.make_adjoint(fwd.u, fwd.v, fwd.w, self.K, self.obj_scale)
        pa = sum(adj.step()[0][0] for _ in range(nadj))
        dchi = dissipation_sensitivity(fwd.u, fwd.v, fwd.w, adj.u, adj.v, adj.w, self.K,
                                       self.obj_scale, True)
        sens, _ = filter_apply(c, ramp_bwd(dchi, self.rf, self.f_min, self.f_max, self.q, True),
                               abstol=self.ftol)
        dfdx = 100.0 * volume_sensitivity(c.n, self.vol_dom, self.is_max)
        fval = np.array([100.0 * vol[0]])
        self.fields = {"u": fwd.u.copy(), "ua": adj.u.copy(), "dF_dchi": dchi, "dF_drho": sens}
        self.x = self.mma.update(self.it, self.x, sens, fval, dfdx)
        rmax, rnorm = self.mma.kkt(self.x, sens, fval, dfdx)
        self.it += 1
        self._map_forward()
        return {"F": obj[0], "dissipation": obj[1], "lube": obj[2], "volume_constraint": vol[0],
                "volume": vol[1], "residumax": rmax, "residunorm": rnorm, "p_iters_fwd": pf,
                "p_iters_adj": pa}
