"""Pins the oracle's Pn/Pn step operators and the step itself.

The residual operators (Neko pnpn_prs_res / pnpn_vel_res) are not in the
reference tree; they are pinned here by exactness on linear fields, the
opgrad/cdtp adjoint identity, curl(grad) = 0, the divergence theorem for the
boundary flux, the BDF/EXT order conditions, and conservation properties of
the step (a quiescent state stays at rest; net volume flux through the
channel vanishes).
"""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def jit():
    return oracle.build_case(3, 2, 2, 6, jitter=0.05, seed=13)


@pytest.fixture(scope="module")
def flat():
    return oracle.build_case(3, 2, 2, 6, jitter=0.0, seed=13)


A = np.array([[0.3, -1.2, 0.7], [2.0, 0.5, -0.4], [-0.6, 1.1, 0.9]])


def lin(c):
    X = np.stack([c.x, c.y, c.zc])
    U = A @ X + np.array([[0.2], [-0.1], [0.4]])
    return U[0], U[1], U[2]


def test_opgrad_exact_on_linear_fields(jit):
    c = jit
    p = 1.5 * c.x - 2.0 * c.y + 0.5 * c.zc + 3.0
    gx, gy, gz = oracle.opgrad(c, p)
    np.testing.assert_allclose(gx, 1.5 * c.B, atol=1e-12)
    np.testing.assert_allclose(gy, -2.0 * c.B, atol=1e-12)
    np.testing.assert_allclose(gz, 0.5 * c.B, atol=1e-12)


def test_cdtp_is_the_transpose_of_opgrad(jit):
    c = jit
    rng = np.random.default_rng(0)
    p = rng.standard_normal(c.n)
    t = [rng.standard_normal(c.n) for _ in range(3)]
    g = oracle.opgrad(c, p)
    lhs = sum(np.dot(t[a], g[a]) for a in range(3))
    rhs = np.dot(p, oracle.cdtp(c, *t))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # the weak divergence of a constant field integrates to zero per element
    one = np.ones(c.n)
    d = oracle.cdtp(c, one, 2 * one, -one)
    assert abs(d.sum()) < 1e-12


def test_curl_exact_on_linear_fields(jit):
    c = jit
    u, v, w = lin(c)
    cx, cy, cz = oracle.curl_local(c, u, v, w)
    np.testing.assert_allclose(cx, A[2, 1] - A[1, 2], atol=1e-11)
    np.testing.assert_allclose(cy, A[0, 2] - A[2, 0], atol=1e-11)
    np.testing.assert_allclose(cz, A[1, 0] - A[0, 1], atol=1e-11)


def test_curl_of_gradient_vanishes(flat):
    c = flat
    phi = c.x ** 2 * c.y - c.zc ** 3 + c.x * c.y * c.zc
    gx = 2 * c.x * c.y + c.y * c.zc
    gy = c.x ** 2 + c.x * c.zc
    gz = -3 * c.zc ** 2 + c.x * c.y
    cx, cy, cz = oracle.curl_local(c, gx, gy, gz)
    assert max(np.abs(cx).max(), np.abs(cy).max(), np.abs(cz).max()) < 1e-11
    del phi


def test_advection_exact_on_linear_fields(jit):
    c = jit
    u, v, w = lin(c)
    U = np.stack([u, v, w])
    conv = A @ U  # (u.grad) u = A u for u = A x + b
    fx, fy, fz = oracle.advection(c, u, v, w)
    for a, f in enumerate((fx, fy, fz)):
        np.testing.assert_allclose(f, -c.B * conv[a], atol=1e-11)


def test_surface_flux_divergence_theorem(jit):
    c = jit
    Lx, Ly, Lz = 1.5, 1.0, 1.0
    u, v, w = c.x, 2 * c.y, -c.zc
    tot = oracle.surface_flux(c, u, v, w, 63).sum()
    assert abs(tot - 2.0 * Lx * Ly * Lz) < 1e-12
    # one face: int_{x=L} u.n = L * Ly * Lz ; int_{x=0} = 0
    assert abs(oracle.surface_flux(c, u, v, w, oracle.F_XMAX).sum() - Lx * Ly * Lz) < 1e-12
    assert abs(oracle.surface_flux(c, u, v, w, oracle.F_XMIN).sum()) < 1e-12
    # uniform inflow through x = 0: int u.n = -area
    one = np.ones(c.n)
    assert abs(oracle.surface_flux(c, one, 0 * one, 0 * one, oracle.F_XMIN).sum() + 1.0) < 1e-12


def test_bdf_ext_order_conditions():
    for k, order in ((1, 1), (2, 2), (5, 2)):
        bd, ext = oracle.pnpn_coeffs(k)
        # BDF: bd0 u^{n+1} - bd1 u^n - bd2 u^{n-1} = dt u' exactly for degree <= order
        for deg in range(order + 1):
            t = np.array([1.0, 0.0, -1.0])  # t_{n+1}, t_n, t_{n-1} with dt = 1
            val = bd[0] * t[0] ** deg - bd[1] * t[1] ** deg - bd[2] * t[2] ** deg
            assert abs(val - deg * t[0] ** max(deg - 1, 0)) < 1e-14
        # EXT: ext1 u^n + ext2 u^{n-1} = u^{n+1} exactly for degree < order
        for deg in range(order):
            assert abs(ext[0] * 0.0 ** deg + ext[1] * (-1.0) ** deg - 1.0) < 1e-14


def test_quiescent_state_stays_at_rest():
    c = oracle.build_case(3, 2, 2, 5, jitter=0.05, seed=3)
    ch = oracle.Channel(c, np.full(c.n, 50.0))
    for a in (ch.u, ch.v, ch.w, *ch.lags):
        a[:] = 0.0
    for _ in range(3):
        it, res = ch.step()
    assert not ch.u.any() and not ch.v.any() and not ch.w.any() and not ch.p.any()


def test_channel_step_conserves_volume_flux():
    c = oracle.build_case(6, 3, 3, 5, jitter=0.05, seed=5)
    chi = np.zeros(c.n)
    ch = oracle.Channel(c, chi, Re=100.0, dt=1e-3)
    for _ in range(6):
        it, res = ch.step()
        assert it[0] > 0 and res[0] < 1e-8
    fin = oracle.surface_flux(c, ch.u, ch.v, ch.w, oracle.F_XMIN).sum()
    fout = oracle.surface_flux(c, ch.u, ch.v, ch.w, oracle.F_XMAX).sum()
    print("inflow flux", fin, "outflow flux", fout)
    assert fin < 0 and abs(fin + fout) < 2e-2 * abs(fin)
    # the walls stay no-slip and the inflow stays prescribed
    walls = oracle.face_mask(c.nx, c.ny, c.nz, c.lx, oracle.WALLS)
    assert not ch.u[walls].any()
    inflow = np.setdiff1d(oracle.face_mask(c.nx, c.ny, c.nz, c.lx, oracle.F_XMIN), walls)
    assert np.all(ch.u[inflow] == 1.0)


def test_adjoint_advection_is_the_discrete_adjoint(jit):
    """<v, AdjAdv_ub(ua)> = <ua, LinAdv_ub(v)> (local Euclidean products),
    LinAdv_ub(v) = -B[(ub.grad)v + (v.grad)ub] built from opgrad alone; it
    holds exactly because cdtp = opgrad^T.  The reference's CPU branch
    (f - w1 + w2 + w3) fails it; its device branch, restated here, holds."""
    c = jit
    rng = np.random.default_rng(41)
    ua = [rng.standard_normal(c.n) for _ in range(3)]
    v = [rng.standard_normal(c.n) for _ in range(3)]
    ub = [rng.standard_normal(c.n) for _ in range(3)]
    f = oracle.adjoint_advection(c, *ua, *ub)
    gv = [oracle.opgrad(c, v[k]) for k in range(3)]     # gv[c][d] = B d v_c / dx_d
    gb = [oracle.opgrad(c, ub[k]) for k in range(3)]
    lin = [-sum(ub[d] * gv[k][d] for d in range(3)) - sum(v[d] * gb[k][d] for d in range(3))
           for k in range(3)]
    lhs = sum(np.dot(v[k], f[k]) for k in range(3))
    rhs = sum(np.dot(ua[k], lin[k]) for k in range(3))
    scale = sum(np.abs(v[k]).dot(np.abs(f[k])) for k in range(3))
    assert abs(lhs - rhs) < 1e-12 * scale, (lhs, rhs)


def test_adjoint_step_linear_in_the_forcing():
    """From a zero adjoint state the adjoint step is linear in its forcing:
    doubling obj_scale and K doubles the adjoint fields (to the solver
    tolerance)."""
    c = oracle.build_case(3, 2, 2, 5, jitter=0.05, seed=3)
    rng = np.random.default_rng(5)
    chi = rng.uniform(0, 50.0, c.n)
    ub = oracle.channel_initial_velocity(c)
    ub = (ub[0] * (1 + 0.3 * np.sin(3 * c.y)), 0.1 * np.cos(2 * c.x), 0.05 * c.zc)
    out = []
    for s in (1.0, 2.0):
        ch = oracle.Channel(c, chi, abstol_p=1e-13, abstol_v=1e-13, max_iter=3000)
        ch.make_adjoint(*ub, K=s, obj_scale=s)
        for _ in range(3):
            ch.step()
        out.append([ch.u.copy(), ch.v.copy(), ch.w.copy(), ch.p.copy()])
    for a, b in zip(out[0], out[1]):
        assert np.abs(b - 2 * a).max() < 1e-9 * np.abs(b).max()
    assert np.abs(out[0][0]).max() > 0
