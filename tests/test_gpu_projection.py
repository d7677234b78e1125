"""The HIP solution projection (nekoamd_proj_*) against the oracle's
restatement (oracle/projection.py) on the same inputs, and the channel step
with projection on against the same step with projection off.

Bars: projected right-hand sides and restored solutions within 1e-10
relative of the oracle (both sides sum in a different order over ~1e5
points and orthogonalise twice), basis sizes equal, with Neko's restart
(keep = 0) and with the restart that keeps the span of the newest solutions
(keep = 2: two restarts in the run), and at an odd point count (3 x 3 x 3
elements, lx = 5: the basis passes' one-point-per-thread form); the channel
fields with
projection stay within the 1e-8 solver tolerance of those without, with
fewer pressure iterations once the basis is filled."""
import numpy as np
import pytest

import oracle
from oracle.projection import Projection as OracleProjection

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def sem():
    from nekoamd import device, sem as S
    device.init(0)
    return S


@pytest.mark.parametrize("dims,lx,kind,keep", [((4, 3, 2), 8, "pressure", 0),
                                               ((8, 4, 4), 6, "velocity", 0),
                                               ((4, 3, 2), 8, "pressure", 2),
                                               ((8, 4, 4), 6, "velocity", 2),
                                               ((3, 3, 3), 5, "pressure", 2)])
def test_projection_matches_oracle(sem, dims, lx, kind, keep):
    nx, ny, nz = dims
    g = sem.BoxCase(nx, ny, nz, lx, jitter=0.05, seed=1234)
    o = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234)
    if kind == "pressure":
        faces, h1v, h2v, ifh2 = sem.PRS_DIRICHLET, 1.0, None, False
    else:
        faces, h1v, h2v, ifh2 = sem.VEL_DIRICHLET, 0.01, 1500.0, True
    mask = oracle.face_mask(nx, ny, nz, lx, faces)
    h1 = np.full(o.n, h1v)
    h2 = np.full(o.n, h2v) if ifh2 else None

    def A(x):
        w = o.gs(o.ax(x, h1, h2, ifh2))
        w[mask] = 0.0
        return w

    L = 4
    P = OracleProjection(o.n, L, A, o.mult, keep=keep)
    Q = sem.Projection(g, L, keep=keep)
    dmask = g.mask(faces)
    h1d = g.field(h1)
    h2d = g.field(h2) if ifh2 else None
    rng = np.random.default_rng(9)
    base = o.gs(rng.standard_normal(o.n)) * o.mult
    for s in range(7):          # fills the basis, restarts it once (keep = 2: twice)
        x = base * (1.0 + 0.1 * s) + 0.05 * o.gs(rng.standard_normal(o.n)) * o.mult
        x[mask] = 0.0
        b = A(x)
        bd = g.field(b)
        Q.pre(bd)
        P.pre(b)
        assert rel(bd.get(), b) < 1e-10, s
        dx = x - (P.xbar if P.m > 0 else 0.0)       # the solver's correction
        xd = g.field(dx)
        Q.post(xd, dmask, h1d, h2d, ifh2)
        xo = P.post(dx.copy())
        assert rel(xd.get(), xo) < 1e-10, s
        assert Q.m == P.m, (s, Q.m, P.m)
    assert P.m < L          # the restart happened


@pytest.mark.parametrize("dims,lx,kind,keep", [((4, 3, 2), 8, "pressure", 0),
                                               ((8, 4, 4), 6, "velocity", 2),
                                               ((3, 3, 3), 5, "pressure", 2)])
def test_fp32_basis_projection(sem, dims, lx, kind, keep):
    """The FP32 basis (nekoamd_proj_set_fp32) on the same sequence as the
    oracle's FP64 projection: the projected right-hand side is exactly
    b - A xbar for the xbar the restore hands back (to 1e-10 relative: the
    FP64 rounding of recovering xbar as x - dx, seen through the operator;
    an FP32 image of the basis would miss by ~1e-7), xbar and the restored solution agree with the
    FP64 projection's to 1e-5 (the FP32 rounding of the basis), and the basis
    sizes are the same, through the restarts."""
    nx, ny, nz = dims
    g = sem.BoxCase(nx, ny, nz, lx, jitter=0.05, seed=1234)
    o = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234)
    if kind == "pressure":
        faces, h1v, h2v, ifh2 = sem.PRS_DIRICHLET, 1.0, None, False
    else:
        faces, h1v, h2v, ifh2 = sem.VEL_DIRICHLET, 0.01, 1500.0, True
    mask = oracle.face_mask(nx, ny, nz, lx, faces)
    h1 = np.full(o.n, h1v)
    h2 = np.full(o.n, h2v) if ifh2 else None

    def A(x):
        w = o.gs(o.ax(x, h1, h2, ifh2))
        w[mask] = 0.0
        return w

    L = 4
    P = OracleProjection(o.n, L, A, o.mult, keep=keep)
    Q = sem.Projection(g, L, keep=keep, fp32=True)
    dmask = g.mask(faces)
    h1d = g.field(h1)
    h2d = g.field(h2) if ifh2 else None
    rng = np.random.default_rng(9)
    base = o.gs(rng.standard_normal(o.n)) * o.mult
    for s in range(7):
        x = base * (1.0 + 0.1 * s) + 0.05 * o.gs(rng.standard_normal(o.n)) * o.mult
        x[mask] = 0.0
        b = A(x)
        b0 = b.copy()
        bd = g.field(b)
        Q.pre(bd)
        P.pre(b)                             # (in place)
        dx = x - (P.xbar if P.m > 0 else 0.0)
        xd = g.field(dx)
        Q.post(xd, dmask, h1d, h2d, ifh2)
        xo = P.post(dx.copy())
        xr = xd.get()
        xbar = xr - dx                       # the xbar the restore added
        bexp = b0 - A(xbar)
        eb = np.abs(bd.get() - bexp).max() / np.abs(b0).max()
        ex = rel(xr, xo)
        ebar = rel(xbar, P.xbar) if s > 0 and np.abs(P.xbar).max() > 0 else 0.0
        print(f"{kind} keep={keep} solve {s}: m {Q.m}/{P.m}, b' - (b - A xbar) {eb:.1e}, "
              f"xbar vs FP64 {ebar:.1e}, x vs FP64 {ex:.1e}")
        assert eb < 1e-10, s
        assert ex < 1e-5, s
        assert ebar < 1e-5, s
        assert Q.m == P.m, (s, Q.m, P.m)
    assert P.m < L


def test_augment_is_galerkin_scaled(sem):
    """gamma = <q,b>/<q,Aq>: b' is orthogonal to q, and a q with a
    non-positive <q,Aq> is ignored."""
    g = sem.BoxCase(4, 3, 2, 6, jitter=0.05, seed=1234)
    o = oracle.build_case(4, 3, 2, 6, jitter=0.05, seed=1234)
    mask = oracle.face_mask(4, 3, 2, 6, sem.PRS_DIRICHLET)

    def A(x):
        w = o.gs(o.ax(x))
        w[mask] = 0.0
        return w

    rng = np.random.default_rng(4)
    x = o.gs(rng.standard_normal(o.n)) * o.mult
    x[mask] = 0
    q = x + 0.3 * o.gs(rng.standard_normal(o.n)) * o.mult
    q[mask] = 0
    b = A(x)
    Q = sem.Projection(g, 2)
    bd = g.field(b)
    gam = Q.augment(bd, g.field(q), g.field(A(q)))
    ref = np.sum(q * b * o.mult) / np.sum(q * A(q) * o.mult)
    assert abs(gam - ref) < 1e-12 * abs(ref)
    bp = bd.get()
    assert abs(np.sum(bp * q * o.mult)) < 1e-10 * np.sqrt(np.sum(b * b * o.mult)) * \
        np.sqrt(np.sum(q * q * o.mult))
    bd2 = g.field(b)
    assert Q.augment(bd2, g.field(q), g.field(np.zeros(o.n))) == 0.0
    assert np.array_equal(bd2.get(), b)


@pytest.mark.parametrize("Lp,Lv,kp,kv,fp32", [(20, 20, 0, 0, False), (6, 4, 3, 2, False),
                                               (20, 0, 10, 0, True), (6, 4, 3, 2, True)])
def test_channel_with_projection_matches_without(sem, Lp, Lv, kp, kv, fp32):
    """(6, 4, 3, 2): small bases that restart inside the run, keeping the
    span of the newest solutions; fp32: the bases in FP32."""
    from nekoamd.fortran import Channel
    c = oracle.build_case(8, 4, 4, 6, jitter=0.05, seed=1234)
    _, _, chi = oracle.seeded_design(c, 777, npass=3)
    a = Channel(8, 4, 4, 6, abstol_p=1e-8, abstol_v=1e-8, max_iter=3000)
    b = Channel(8, 4, 4, 6, abstol_p=1e-8, abstol_v=1e-8, max_iter=3000)
    try:
        a.set("chi", chi)
        b.set("chi", chi)
        b.projection(Lp, Lv, activ_step=2, keep_prs=kp, keep_vel=kv, fp32=fp32)
        ita, _ = a.step(12)
        itb, resb = b.step(12)
        print("pressure iters without", ita[:, 0].tolist(), "with", itb[:, 0].tolist())
        print("velocity iters without", ita[:, 1:].sum(1).tolist(), "with",
              itb[:, 1:].sum(1).tolist())
        assert np.all(resb < 1e-8)
        for k in ("u", "v", "w", "p"):
            e = rel(b.get(k), a.get(k))
            print(k, f"{e:.1e}")
            assert e < 1e-6, k
        assert itb[6:, 0].sum() < ita[6:, 0].sum()
    finally:
        a.free()
        b.free()
