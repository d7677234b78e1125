"""The element derivative operators of the Pn/Pn step (csrc/pnpn.hip:
nekoamd_advection, nekoamd_curl, nekoamd_opgrad, nekoamd_cdtp) against the
FP64 CPU oracle (oracle/pnpn.c), on two geometries:

  - the jittered box (every element trilinear: the kernels take the metric
    from the end-to-end differences of the grid lines);
  - the same box with its interior points moved along a sine in each
    direction (curved elements: the kernels take the spectral derivatives
    of the coordinates).

Bar: within 1e-12 of max|ref| in the max norm (FP64 on both sides, the
device sums in its own order).
"""
import copy
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(4, 3, 2, 8), (3, 2, 2, 5)], ids=["lx8", "lx5"])
def pair(request):
    from nekoamd import device, sem
    device.init(0)
    nx, ny, nz, lx = request.param
    print(f"building {nx}x{ny}x{nz} lx={lx}", flush=True)
    return (sem.BoxCase(nx, ny, nz, lx, jitter=0.05, seed=1234),
            oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234))


def _curved(c):
    """The case's coordinates with every point moved by 0.02 sin(pi y) sin(pi z)
    in x (and the like in y and z): no grid line stays straight."""
    cc = copy.copy(c)
    x, y, z = c.x.copy(), c.y.copy(), c.zc.copy()
    cc.x = x + 0.02 * np.sin(np.pi * y) * np.sin(np.pi * z)
    cc.y = y + 0.02 * np.sin(np.pi * z) * np.sin(np.pi * x)
    cc.zc = z + 0.02 * np.sin(np.pi * x) * np.sin(np.pi * y)
    return cc


def _err(got, ref):
    a = np.concatenate([np.ravel(v) for v in got])
    b = np.concatenate([np.ravel(v) for v in ref])
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("geom", ["trilinear", "curved"])
def test_derivative_operators_match_oracle(pair, geom):
    from nekoamd import device as dev
    from nekoamd._lib import check, lib, ptr, ri
    g, c = pair
    cc = c if geom == "trilinear" else _curved(c)
    xyz = [g.field(a) for a in (cc.x, cc.y, cc.zc)]
    pxyz = [ptr(a) for a in xyz]
    wd = dev.DeviceArray.from_host(g.wgll)
    lx, nelv = ri(g.lx), ri(g.nelv)
    rng = np.random.default_rng(7)
    u = [rng.standard_normal(g.n) for _ in range(3)]
    ud = [g.field(a) for a in u]
    B = g.field(c.B)
    out = [g.field(np.full(g.n, np.nan)) for _ in range(3)]
    errs = {}

    check(lib().nekoamd_advection(*(ptr(a) for a in out), *(ptr(a) for a in ud), ptr(B),
                                  *pxyz, ptr(g.dx), lx, nelv), "advection")
    errs["advection"] = _err([a.get() for a in out], oracle.advection(cc, *u))

    check(lib().nekoamd_curl(*(ptr(a) for a in out), *(ptr(a) for a in ud), *pxyz,
                             ptr(g.dx), lx, nelv), "curl")
    errs["curl"] = _err([a.get() for a in out], oracle.curl_local(cc, *u))

    check(lib().nekoamd_opgrad(*(ptr(a) for a in out), ptr(ud[0]), *pxyz, ptr(g.dx),
                               ptr(wd), lx, nelv), "opgrad")
    errs["opgrad"] = _err([a.get() for a in out], oracle.opgrad(cc, u[0]))

    check(lib().nekoamd_cdtp(ptr(out[0]), *(ptr(a) for a in ud), *pxyz, ptr(g.dx), ptr(wd),
                             lx, nelv), "cdtp")
    errs["cdtp"] = _err([out[0].get()], [oracle.cdtp(cc, *u)])

    print(f"lx={g.lx} {geom}: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) < 1e-12, errs
