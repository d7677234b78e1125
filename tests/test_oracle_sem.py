"""Pins the FP64 CPU oracle (GLL, mesh, coef, ax_helm, gs, Jacobi, CG).

Neko, where these algorithms live, is not part of the reference tree, so no
golden vector from the reference exists for them.  Each check below is one a
plausible mistake would fail: published test vectors (SplitMix64), closed
forms (quadrature exactness, volumes, A*1 = 0, energy = int |grad u|^2), and
dense brute-force assembly from independent numpy code (tests/brute.py).
"""
import numpy as np
import pytest

import oracle
from brute import element_matrices, gll_numpy, global_matrix, gs_numpy


def test_splitmix64_published_vectors():
    # SplitMix64 with state 0: first outputs of the reference generator
    # (Steele, Lea & Flood 2014; Vigna's splitmix64.c).
    L = oracle.lib()
    assert L.or_splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert L.or_splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert L.or_splitmix64(0, 2) == 0x06C45D188009454F
    u = np.array([oracle.u01(7, i) for i in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01


@pytest.mark.parametrize("lx", [2, 3, 4, 6, 8, 10])
def test_gll_against_numpy_legendre(lx):
    z, w, D = oracle.gll(lx)
    zn, wn, Dn = gll_numpy(lx)
    np.testing.assert_allclose(z, zn, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, wn, rtol=1e-13)
    np.testing.assert_allclose(D, Dn, rtol=0, atol=1e-12 * lx * lx)
    assert abs(w.sum() - 2.0) < 1e-14
    # D differentiates polynomials of degree <= N exactly
    for k in range(lx):
        np.testing.assert_allclose(D @ z ** k, k * z ** max(k - 1, 0) * (k > 0), atol=1e-11)
    # quadrature exact to degree 2N-1
    for k in range(2 * lx - 2):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-13


def test_box_vertices_jitter_rules():
    nx, ny, nz = 5, 4, 3
    v = oracle.box_vertices(nx, ny, nz, jitter=0.05, seed=99).reshape(-1, 3)
    Lx, Ly, Lz = nx / ny, 1.0, nz / ny
    h = np.array([Lx / nx, Ly / ny, Lz / nz])
    idx = np.array([(vx, vy, vz) for vz in range(nz + 1) for vy in range(ny + 1)
                    for vx in range(nx + 1)])
    base = idx * h
    d = v - base
    assert np.all(np.abs(d) <= 0.05 * h + 1e-15)
    for a, nn in enumerate((nx, ny, nz)):
        on_bnd = (idx[:, a] == 0) | (idx[:, a] == nn)
        assert np.all(d[on_bnd, a] == 0.0)
        assert np.all(np.abs(d[~on_bnd, a]) > 0)
    # same seed -> same mesh, other seed -> different
    v2 = oracle.box_vertices(nx, ny, nz, jitter=0.05, seed=99)
    v3 = oracle.box_vertices(nx, ny, nz, jitter=0.05, seed=100)
    assert np.array_equal(v.ravel(), v2)
    assert not np.array_equal(v.ravel(), v3)


@pytest.mark.parametrize("lx,jit", [(3, 0.0), (4, 0.05), (6, 0.05)])
def test_coef_against_brute(lx, jit):
    c = oracle.build_case(3, 2, 2, lx, jitter=jit, seed=5)
    z, w, D = gll_numpy(lx)
    npt = lx ** 3
    for e in range(c.nelv):
        Xc = c.corners[e * 24:(e + 1) * 24].reshape(8, 3)
        K, Bm, xyz, Gab = element_matrices(Xc, z, w, D)
        sl = slice(e * npt, (e + 1) * npt)
        np.testing.assert_allclose(c.B[sl], Bm, rtol=1e-12)
        np.testing.assert_allclose(np.stack([c.x[sl], c.y[sl], c.zc[sl]], 1), xyz, atol=1e-14)
        pairs = {"G11": (0, 0), "G12": (0, 1), "G13": (0, 2), "G22": (1, 1),
                 "G23": (1, 2), "G33": (2, 2)}
        scale = np.abs(Gab).max()
        for k, (a, b) in pairs.items():
            np.testing.assert_allclose(c.G[k][sl], Gab[:, a, b], rtol=0, atol=1e-12 * scale)
    # the jittered elements still tile the box: sum B = volume
    assert abs(c.B.sum() - (3 / 2) * 1.0 * 1.0) < 1e-13
    assert abs(c.volume - 1.5) < 1e-13


def test_coef_closed_form_unjittered():
    lx = 5
    c = oracle.build_case(2, 2, 2, lx, jitter=0.0)
    h = 0.5
    w3 = np.einsum("k,j,i->kji", c.w, c.w, c.w).ravel()
    for e in range(c.nelv):
        sl = slice(e * lx ** 3, (e + 1) * lx ** 3)
        np.testing.assert_allclose(c.G["G11"][sl], w3 * h / 2, rtol=1e-13)
        np.testing.assert_allclose(c.G["G33"][sl], w3 * h / 2, rtol=1e-13)
        assert np.abs(c.G["G12"][sl]).max() < 1e-15
        np.testing.assert_allclose(c.B[sl], w3 * h ** 3 / 8, rtol=1e-13)


@pytest.mark.parametrize("lx,jit", [(4, 0.05), (6, 0.05), (8, 0.05)])
def test_ax_helm_against_dense_element_matrices(lx, jit):
    c = oracle.build_case(2, 2, 1, lx, jitter=jit, seed=11)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, c.n)
    h1 = rng.uniform(0.5, 2.0, c.n)
    h2 = rng.uniform(0.5, 2.0, c.n)
    w = c.ax(u, h1, h2, ifh2=True)
    z, wq, D = gll_numpy(lx)
    npt = lx ** 3
    for e in range(c.nelv):
        Xc = c.corners[e * 24:(e + 1) * 24].reshape(8, 3)
        sl = slice(e * npt, (e + 1) * npt)
        K, Bm, _, _ = element_matrices(Xc, z, wq, D, h1[sl])
        ref = K @ u[sl] + h2[sl] * Bm * u[sl]
        np.testing.assert_allclose(w[sl], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_ax_helm_closed_forms():
    lx = 6
    c = oracle.build_case(3, 2, 2, lx, jitter=0.05, seed=21)
    n = c.n
    # constants are in the kernel of the stiffness (h2 = 0)
    w = c.ax(np.ones(n))
    assert np.abs(w).max() < 1e-12
    # symmetry of the element-local operator
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    a, b = v @ c.ax(u), u @ c.ax(v)
    assert abs(a - b) < 1e-12 * abs(a)
    # h1 = 0, h2 = 1  ->  A u = B u
    w = c.ax(u, h1=np.zeros(n), h2=np.ones(n), ifh2=True)
    np.testing.assert_allclose(w, c.B * u, rtol=1e-15)


def test_ax_helm_energy_equals_gradient_integral():
    # u of total degree <= N on rectilinear elements: GLL integrates
    # |grad u|^2 exactly, so u^T A u = int |grad u|^2 over the box.
    lx = 5
    c = oracle.build_case(2, 2, 2, lx, jitter=0.0)
    x, y, z = c.x, c.y, c.zc
    u = x ** 2 * y + 3 * y * z - z ** 3 + x
    energy = u @ c.ax(u)
    gq, gw = np.polynomial.legendre.leggauss(8)
    X = 0.5 * (gq + 1)
    Wt = 0.5 * gw
    xx, yy, zz = np.meshgrid(X, X, X, indexing="ij")
    W = np.einsum("i,j,k->ijk", Wt, Wt, Wt)
    ux = 2 * xx * yy + 1
    uy = xx ** 2 + 3 * zz
    uz = 3 * yy - 3 * zz ** 2
    exact = np.sum(W * (ux ** 2 + uy ** 2 + uz ** 2))
    assert abs(energy - exact) < 1e-12 * exact


def test_gs_maps_and_op():
    nx, ny, nz, lx = 3, 2, 2, 4
    c = oracle.build_case(nx, ny, nz, lx, jitter=0.05)
    p = lx - 1
    mult = c.gs(np.ones(c.n))
    assert set(np.unique(mult)) <= {1.0, 2.0, 4.0, 8.0}
    nunique = (nx * p + 1) * (ny * p + 1) * (nz * p + 1)
    assert abs(np.sum(1.0 / mult) - nunique) < 1e-9
    # canonical map order: buckets by m, tuples by gid, members ascending
    off = 0
    for m, cnt in zip(c.gs_m, c.gs_cnt):
        t = c.gs_idx[off:off + m * cnt].reshape(cnt, m)
        assert np.all(np.diff(t, axis=1) > 0)
        g = c.gid[t]
        assert np.all(g == g[:, :1])
        assert np.all(np.diff(g[:, 0]) > 0)
        off += m * cnt
    assert off == c.gs_idx.size
    # integer-valued data: exact against numpy's scatter-add
    rng = np.random.default_rng(1)
    u = rng.integers(-1000, 1000, c.n).astype(np.float64)
    assert np.array_equal(c.gs(u), gs_numpy(u, c.gid))
    # MIN / MAX / MUL
    g = c.gid
    ng = g.max() + 1
    mn = np.full(ng, np.inf); np.minimum.at(mn, g, u)
    mx = np.full(ng, -np.inf); np.maximum.at(mx, g, u)
    assert np.array_equal(c.gs(u, oracle.GS_MIN), mn[g])
    assert np.array_equal(c.gs(u, oracle.GS_MAX), mx[g])
    v = rng.integers(1, 4, c.n).astype(np.float64)
    pr = np.ones(ng); np.multiply.at(pr, g, v)
    assert np.array_equal(c.gs(v, oracle.GS_MUL), pr[g])
    # gs(u / mult) is continuous and idempotent
    q = c.gs(u * c.mult)
    assert np.array_equal(c.gs(q * c.mult), q)


@pytest.mark.parametrize("jit", [0.0, 0.05])
def test_jacobi_against_dense_diagonal(jit):
    lx = 4
    c = oracle.build_case(2, 2, 2, lx, jitter=jit, seed=8)
    rng = np.random.default_rng(2)
    h1 = np.full(c.n, 0.7)
    h2 = rng.uniform(1.0, 2.0, c.n)
    dinv = c.jacobi(h1, h2, ifh2=True)
    z, w, D = gll_numpy(lx)
    npt = lx ** 3
    dloc = np.zeros(c.n)
    for e in range(c.nelv):
        Xc = c.corners[e * 24:(e + 1) * 24].reshape(8, 3)
        sl = slice(e * npt, (e + 1) * npt)
        K, Bm, _, Gab = element_matrices(Xc, z, w, D, h1[sl])
        dk = np.diag(K).copy()
        # the Jacobi diagonal omits the G12/G13/G23 cross terms
        i = np.arange(npt) % lx
        j = (np.arange(npt) // lx) % lx
        k = np.arange(npt) // (lx * lx)
        cross = 2 * h1[sl] * (D[i, i] * D[j, j] * Gab[:, 0, 1] + D[i, i] * D[k, k] * Gab[:, 0, 2]
                              + D[j, j] * D[k, k] * Gab[:, 1, 2])
        dloc[sl] = dk - cross + h2[sl] * Bm
    ref = 1.0 / gs_numpy(dloc, c.gid)
    np.testing.assert_allclose(dinv, ref, rtol=1e-12)
    if jit == 0.0:
        A = global_matrix(c, h1, h2)
        np.testing.assert_allclose(1.0 / dinv, np.diag(A)[c.gid], rtol=1e-12)


@pytest.mark.parametrize("lx,ifh2", [(4, False), (5, True)])
def test_cg_against_dense_solve(lx, ifh2):
    c = oracle.build_case(3, 2, 2, lx, jitter=0.05, seed=4)
    n = c.n
    mask = oracle.face_mask(c.nx, c.ny, c.nz, lx, oracle.PRS_DIRICHLET if not ifh2
                            else oracle.VEL_DIRICHLET)
    rng = np.random.default_rng(5)
    h1 = np.full(n, 1.0 if not ifh2 else 0.01)
    h2 = np.full(n, 1500.0) if ifh2 else None
    f = c.gs(c.B * rng.uniform(-1, 1, n))
    f[mask] = 0.0
    x, it, r0, r1 = c.cg(f, mask, h1, h2, ifh2, abstol=1e-13, max_iter=2000)
    assert r1 < 1e-13 and it > 1
    A = global_matrix(c, h1, h2)
    ng = A.shape[0]
    fixed = np.zeros(ng, bool)
    fixed[c.gid[mask]] = True
    b = np.zeros(ng)
    b[c.gid] = f
    free = ~fixed
    xg = np.zeros(ng)
    xg[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    np.testing.assert_allclose(x, xg[c.gid], rtol=0, atol=1e-9 * np.abs(xg).max())
    # the solution is continuous (all copies equal) and zero on the mask
    assert np.array_equal(x[mask], np.zeros(mask.size))
    np.testing.assert_allclose(c.gs(x * c.mult), x, rtol=1e-12, atol=1e-14)


def test_cg_reports_nonconvergence_and_zero_rhs():
    c = oracle.build_case(2, 2, 2, 4, jitter=0.05)
    mask = oracle.face_mask(2, 2, 2, 4, oracle.PRS_DIRICHLET)
    f = c.gs(c.B)
    f[mask] = 0
    x, it, r0, r1 = c.cg(f, mask, abstol=0.0, max_iter=3)
    assert it == 3 and r1 > 0
    x, it, r0, r1 = c.cg(np.zeros(c.n), mask)
    assert it == 0 and r0 == 0.0 and not x.any()
