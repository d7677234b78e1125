"""Pins of the hybrid Schwarz multigrid oracle (oracle/hsmg.py).

  - the ghost exchange puts exactly the neighbour's first interior layer
    into each ghost layer (zero on the domain boundary);
  - the overlap counts (the inverse Schwarz weights) are 1 inside an
    element, 2 on shared faces and on the layer next to a shared face, 4 on
    shared edges, 8 at shared vertices;
  - the fast-diagonalisation solve is the inverse of the element operator:
    on an undistorted single element, ax_helm (pinned elsewhere against
    dense assembly) applied to the FDM result returns the input, with
    Neumann faces (mean-free input) and with a Dirichlet face;
  - the unweighted Schwarz operator is symmetric in <.,.>_mult (a wrong
    layer in the exchange or in the return sum breaks this);
  - as a GMRES preconditioner it cuts the iterations of Jacobi-GMRES at
    least fourfold.
"""
import numpy as np
import pytest

import oracle
from oracle import hsmg, krylov


@pytest.fixture(scope="module")
def sch():
    c = oracle.build_case(3, 2, 2, 6, jitter=0.05, seed=1234)
    return c, hsmg.Schwarz(c, oracle.PRS_DIRICHLET)


def test_exchange_fills_ghosts_with_neighbour_layer(sch):
    c, s = sch
    lx, ne = c.lx, c.nelv
    f = np.random.default_rng(1).standard_normal(c.n)
    W = s._exchange(s._to_ext(f))
    F = f.reshape(ne, lx, lx, lx)                 # [e][k][j][i]
    # (a + b) - a rounds: equal to 1e-15
    eq = lambda a, b: np.allclose(a, np.broadcast_to(b, (lx, lx)), rtol=0, atol=1e-15 * 8)
    for e in range(ne):
        ex, ey, ez = e % c.nx, (e // c.nx) % c.ny, e // (c.nx * c.ny)
        sl = slice(1, lx + 1)
        # r direction (last axis)
        lo = F[e - 1, :, :, lx - 2] if ex > 0 else 0.0
        hi = F[e + 1, :, :, 1] if ex < c.nx - 1 else 0.0
        assert eq(W[e, sl, sl, 0], lo)
        assert eq(W[e, sl, sl, lx + 1], hi)
        # s direction
        lo = F[e - c.nx, :, lx - 2, :] if ey > 0 else 0.0
        hi = F[e + c.nx, :, 1, :] if ey < c.ny - 1 else 0.0
        assert eq(W[e, sl, 0, sl], lo)
        assert eq(W[e, sl, lx + 1, sl], hi)
        # t direction
        st = c.nx * c.ny
        lo = F[e - st, lx - 2, :, :] if ez > 0 else 0.0
        hi = F[e + st, 1, :, :] if ez < c.nz - 1 else 0.0
        assert eq(W[e, 0, sl, sl], lo)
        assert eq(W[e, lx + 1, sl, sl], hi)


def test_overlap_counts(sch):
    c, s = sch
    lx = c.lx
    cnt = (1.0 / s.wt).reshape(c.nelv, lx, lx, lx)
    # element 1 of a 3x2x2 box: neighbours at x-, x+, y+, z+
    e = 1
    assert cnt[e, 2:lx - 2, 2:lx - 2, 2:lx - 2].min() == 1.0 == cnt[e, 2:lx - 2, 2:lx - 2, 2:lx - 2].max()
    assert np.all(cnt[e, 2:lx - 2, 2:lx - 2, 0] == 2)       # shared face x-
    assert np.all(cnt[e, 2:lx - 2, 2:lx - 2, 1] == 2)       # layer next to it
    assert np.all(cnt[e, 2:lx - 2, 0, 2:lx - 2] == 1)       # domain boundary y-
    assert np.all(cnt[e, 2:lx - 2, lx - 1, 0] == 4)         # shared edge (x-, y+)
    assert cnt[e, lx - 1, lx - 1, 0] == 8                    # shared vertex


@pytest.mark.parametrize("faces", [0, oracle.F_XMAX])
def test_fdm_inverts_the_element_operator(faces):
    c = oracle.build_case(1, 1, 1, 6, jitter=0.0, seed=1)
    s = hsmg.Schwarz(c, faces)
    n = c.n
    A = np.array([c.ax(np.eye(n)[i]) for i in range(n)]).T
    act = np.ones(n, dtype=bool)
    act[oracle.face_mask(1, 1, 1, 6, faces)] = False
    rng = np.random.default_rng(4)
    f = rng.standard_normal(n)
    f[~act] = 0.0
    if faces == 0:
        f -= f.mean()
    u = s.fdm(s._to_ext(f)).reshape(1, 8, 8, 8)[:, 1:7, 1:7, 1:7].reshape(-1)
    assert np.all(u[~act] == 0.0)
    res = (A @ u)[act] - f[act]
    print(f"faces={faces}: |A FDM f - f| / |f| = {np.abs(res).max() / np.abs(f).max():.1e}")
    assert np.abs(res).max() < 1e-10 * np.abs(f).max()


def test_schwarz_is_symmetric(sch):
    c, s = sch
    rng = np.random.default_rng(6)

    def cont(v):
        v = c.mult * c.gs(v)
        v[s.mask] = 0.0
        return v

    a, b = cont(rng.standard_normal(c.n)), cont(rng.standard_normal(c.n))

    def S0(r):
        return s._sum_overlap(s.fdm(s._exchange(s._to_ext(r))))

    lhs = np.dot(S0(a) * c.mult, b)
    rhs = np.dot(a * c.mult, S0(b))
    print(f"<S a, b> = {lhs:.15e}, <a, S b> = {rhs:.15e}")
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_hsmg_cuts_gmres_iterations():
    nx, ny, nz, lx = 4, 3, 3, 6
    M = hsmg.HSMG(nx, ny, nz, lx)
    c = M.c3
    mask = oracle.face_mask(nx, ny, nz, lx, oracle.PRS_DIRICHLET)
    rng = np.random.default_rng(9)
    f = c.gs(c.B * rng.standard_normal(c.n))
    f[mask] = 0.0
    d = c.jacobi()

    def A(u):
        w = c.gs(c.ax(u))
        w[mask] = 0.0
        return w

    xj, itj, _, rj, _ = krylov.gmres(A, lambda v: d * v, f, c.mult, c.volume, m=30,
                                     abstol=1e-8, max_iter=3000)
    xh, ith, _, rh, _ = krylov.gmres(A, M.apply, f, c.mult, c.volume, m=30,
                                     abstol=1e-8, max_iter=3000)
    print(f"GMRES iterations: Jacobi {itj}, HSMG {ith}")
    assert rj < 1e-8 and rh < 1e-8
    assert ith * 4 <= itj
    xs = c.cg(f, mask, abstol=1e-13, max_iter=5000, dinv=d)[0]
    for x in (xh, xj):
        assert np.abs(x - xs).max() < 1e-4 * np.abs(xs).max()
