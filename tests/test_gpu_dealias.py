"""Dealiased advection on the GPU (csrc/dealias.hip) against the CPU oracle
(oracle/dealias.c, pinned in test_oracle_dealias.py).

Bar (FP64 on both sides, the device sums in its own order): the GL metric
within 1e-13 and every operator within 1e-12 of max|f| in the max norm, at
p = 7 (lxd = 12), p = 5 (lxd = 9) and p = 3 (lxd = 6), on jittered boxes
with random fields.  The device's forward operator also honours beta (0:
overwrite, 1: subtract from the forcing it is given).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(4, 3, 2, 8), (3, 2, 2, 6), (2, 2, 2, 4)]


@pytest.fixture(scope="module", params=SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}_lx{s[3]}")
def pair(request):
    from nekoamd import device, sem
    device.init(0)
    nx, ny, nz, lx = request.param
    g = sem.BoxCase(nx, ny, nz, lx, jitter=0.05, seed=1234)
    c = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234)
    return g, c, sem.Dealias(g), oracle.dealias_setup(c)


def _fields(c, seed, k=3):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(c.n) for _ in range(k)]


def _err(got, ref):
    a = np.concatenate(got)
    b = np.concatenate(ref)
    return np.abs(a - b).max() / np.abs(b).max()


def test_gl_metric(pair):
    g, c, d, o = pair
    m = d.metric()
    ref = o.cof.reshape(9, -1)
    err = np.abs(m - ref).max() / np.abs(ref).max()
    print(f"lx={c.lx}: GL metric error {err:.1e}")
    assert err < 1e-13


def test_forward_advection(pair):
    g, c, d, o = pair
    u = _fields(c, 1)
    f0 = _fields(c, 2)
    ud = [g.field(a) for a in u]
    f = [g.field(np.full(c.n, np.nan)) for _ in range(3)]
    d.advection(f, ud, beta=0.0)
    ref = oracle.dealias_advection(o, *u)
    e0 = _err([a.get() for a in f], ref)
    f = [g.field(a) for a in f0]
    d.advection(f, ud, beta=1.0)
    ref1 = oracle.dealias_advection(o, *u, f=f0)
    e1 = _err([a.get() for a in f], ref1)
    print(f"lx={c.lx}: forward (overwrite) {e0:.1e}, (subtract) {e1:.1e}")
    assert e0 < 1e-12 and e1 < 1e-12


def test_linear_advection(pair):
    g, c, d, o = pair
    up, ub, f0 = _fields(c, 3), _fields(c, 4), _fields(c, 5)
    f = [g.field(a) for a in f0]
    d.linear(f, [g.field(a) for a in up], [g.field(a) for a in ub])
    ref = oracle.dealias_linear(o, *up, *ub, f=f0)
    e = _err([a.get() for a in f], ref)
    print(f"lx={c.lx}: linearised {e:.1e}")
    assert e < 1e-12


def test_adjoint_advection(pair):
    g, c, d, o = pair
    ua, ub, f0 = _fields(c, 6), _fields(c, 7), _fields(c, 8)
    f = [g.field(a) for a in f0]
    d.adjoint(f, [g.field(a) for a in ua], [g.field(a) for a in ub])
    ref = oracle.dealias_adjoint(o, *ua, *ub, f=f0)
    e = _err([a.get() for a in f], ref)
    # the adjoint term alone (without the forcing it is subtracted from)
    fz = [g.field() for _ in range(3)]
    d.adjoint(fz, [g.field(a) for a in ua], [g.field(a) for a in ub])
    rz = oracle.dealias_adjoint(o, *ua, *ub)
    ez = _err([a.get() for a in fz], rz)
    print(f"lx={c.lx}: adjoint {e:.1e} (term alone {ez:.1e})")
    assert e < 1e-12 and ez < 1e-12


def test_device_duality(pair):
    """<a, L(u'; U)> = <u', A(a; U)> with the device's own operators."""
    g, c, d, o = pair
    U, up, a = _fields(c, 9), _fields(c, 10), _fields(c, 11)
    L = [g.field() for _ in range(3)]
    A = [g.field() for _ in range(3)]
    d.linear(L, [g.field(x) for x in up], [g.field(x) for x in U])
    d.adjoint(A, [g.field(x) for x in a], [g.field(x) for x in U])
    Lh, Ah = [x.get() for x in L], [x.get() for x in A]
    lhs = sum(a[k] @ Lh[k] for k in range(3))
    rhs = sum(up[k] @ Ah[k] for k in range(3))
    scale = sum(np.abs(a[k]) @ np.abs(Lh[k]) for k in range(3))
    print(f"lx={c.lx}: duality {lhs:.14e} vs {rhs:.14e}")
    assert abs(lhs - rhs) < 1e-13 * scale
