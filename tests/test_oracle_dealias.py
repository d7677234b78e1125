"""Pins of the dealiased-advection oracle (oracle/dealias.c).

Each operator or rule is checked against something the mathematics fixes,
not against a restatement of its own formula:
  - the Gauss-Legendre rule against numpy's leggauss;
  - the GLL->GL interpolation and the GL derivative matrix on polynomials
    they must reproduce exactly;
  - the GL metric (interpolated GLL cofactors, adv_adjoint_dealias.f90:
    155-163) against cofactors computed directly at the GL points from the
    trilinear element map;
  - forward and adjoint dealiased operators against the no-dealias oracle on
    quadratic fields, where both quadratures are exact (the integrands have
    degree <= 12 per direction, GLL with 8 points is exact to 13) -- on the
    jittered meshes, where aliasing would otherwise separate them;
  - the linearised and adjoint operators against each other by the discrete
    duality <a, L(u'; U)> = <u', A(a; U)>, which holds only with the
    transpose map back to GLL and consistent indices;
  - the forward operator against the reference's linearised one by the exact
    quadratic identity F(U + e u') = F(U) + e L(u'; U) + e^2 F(u').
"""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def case():
    return oracle.build_case(3, 2, 2, 8, jitter=0.05, seed=1234)


@pytest.fixture(scope="module")
def dl(case):
    return oracle.dealias_setup(case)


@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_gauss_legendre_matches_numpy(n):
    z, w = oracle.gauss_legendre(n)
    zr, wr = np.polynomial.legendre.leggauss(n)
    assert np.abs(z - zr).max() < 1e-15
    assert np.abs(w - wr).max() < 1e-15


def test_interpolation_and_derivative_exact_on_polynomials():
    zl, _, _ = oracle.gll(8)
    zg, _ = oracle.gauss_legendre(12)
    J = oracle.lagrange_matrix(zl, zg)
    D = oracle.deriv_matrix(zg)
    assert np.abs(J.sum(1) - 1).max() < 1e-14
    for k in range(8):
        assert np.abs(J @ zl ** k - zg ** k).max() < 1e-13, k
    for k in range(12):
        d = k * zg ** (k - 1) if k else 0 * zg
        assert np.abs(D @ zg ** k - d).max() < 1e-11 * max(1, k * k), k


def _trilinear_gl_cofactors(case, zg, Dg):
    """Cofactors at the GL points straight from the element corners."""
    lxd = zg.size
    corners = case.corners.reshape(case.nelv, 8, 3)
    r = zg[None, None, :]
    s = zg[None, :, None]
    t = zg[:, None, None]
    out = np.zeros((9, case.nelv, lxd ** 3))
    for e in range(case.nelv):
        X = np.zeros((3, lxd, lxd, lxd))
        for cidx in range(8):
            cx, cy, cz = cidx & 1, (cidx >> 1) & 1, (cidx >> 2) & 1
            wgt = ((1 + (2 * cx - 1) * r) * (1 + (2 * cy - 1) * s) * (1 + (2 * cz - 1) * t)) / 8
            for d in range(3):
                X[d] += wgt * corners[e, cidx, d]
        # derivatives along r (last axis), s, t
        dr = np.einsum("il,dtsl->dtsi", Dg, X)
        ds = np.einsum("il,dtlr->dtir", Dg, X)
        dt = np.einsum("il,dlsr->disr", Dg, X)
        ar, as_, at = dr.reshape(3, -1), ds.reshape(3, -1), dt.reshape(3, -1)
        cof = np.zeros((9, lxd ** 3))
        cof[0:3] = np.cross(as_.T, at.T).T
        cof[3:6] = np.cross(at.T, ar.T).T
        cof[6:9] = np.cross(ar.T, as_.T).T
        out[:, e] = cof
    return out.reshape(9, -1)


def test_gl_metric_equals_direct_cofactors(case, dl):
    ref = _trilinear_gl_cofactors(case, dl.zg, dl.Dg)
    got = dl.cof.reshape(9, -1)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"GL metric vs direct trilinear cofactors: {err:.1e}")
    # the direct route differentiates with the 12-point GL matrix (entries
    # up to ~70), which costs a few hundred ulps
    assert err < 1e-12


def _quadratic_fields(case, seed):
    rng = np.random.default_rng(seed)
    X = np.stack([case.x, case.y, case.zc])
    out = []
    for _ in range(3):
        a = rng.standard_normal(10)
        x, y, z = X
        out.append(a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * y + a[5] * y * z
                   + a[6] * x * z + a[7] * x * x + a[8] * y * y + a[9] * z * z)
    return out


def _rel(a, b):
    a = np.concatenate(a)
    b = np.concatenate(b)
    return np.abs(a - b).max() / np.abs(b).max()


def test_forward_equals_no_dealias_on_quadratics(case, dl):
    u, v, w = _quadratic_fields(case, 1)
    fd = oracle.dealias_advection(dl, u, v, w)
    fn = oracle.advection(case, u, v, w)
    err = _rel(fd, fn)
    print(f"dealiased vs no-dealias forward advection, quadratic fields: {err:.1e}")
    assert err < 1e-12


def test_forward_differs_from_no_dealias_on_rough_fields(case, dl):
    rng = np.random.default_rng(2)
    u, v, w = (rng.standard_normal(case.n) for _ in range(3))
    err = _rel(oracle.dealias_advection(dl, u, v, w), oracle.advection(case, u, v, w))
    assert err > 1e-3  # aliasing: the rules really differ


def test_adjoint_equals_no_dealias_on_quadratics(case, dl):
    ua, va, wa = _quadratic_fields(case, 3)
    ub, vb, wb = _quadratic_fields(case, 4)
    fd = oracle.dealias_adjoint(dl, ua, va, wa, ub, vb, wb)
    fn = oracle.adjoint_advection(case, ua, va, wa, ub, vb, wb)
    err = _rel(fd, fn)
    print(f"dealiased vs no-dealias adjoint advection, quadratic fields: {err:.1e}")
    assert err < 1e-12


def test_linear_and_adjoint_are_dual(case, dl):
    rng = np.random.default_rng(5)
    U = [rng.standard_normal(case.n) for _ in range(3)]
    up = [rng.standard_normal(case.n) for _ in range(3)]
    a = [rng.standard_normal(case.n) for _ in range(3)]
    L = oracle.dealias_linear(dl, *up, *U)
    A = oracle.dealias_adjoint(dl, *a, *U)
    lhs = sum(np.dot(a[c], L[c]) for c in range(3))
    rhs = sum(np.dot(up[c], A[c]) for c in range(3))
    scale = sum(np.abs(a[c]) @ np.abs(L[c]) for c in range(3))
    print(f"<a, L u'> = {lhs:.12e}, <u', A a> = {rhs:.12e}")
    assert abs(lhs - rhs) < 1e-13 * scale


def test_duality_fails_for_a_planted_index_mistake(case, dl):
    """The duality check is sharp: transposing the GL derivative matrix in
    the adjoint (a plausible indexing mistake) breaks it."""
    import dataclasses
    rng = np.random.default_rng(6)
    U = [rng.standard_normal(case.n) for _ in range(3)]
    up = [rng.standard_normal(case.n) for _ in range(3)]
    a = [rng.standard_normal(case.n) for _ in range(3)]
    bad = dataclasses.replace(dl, Dg=np.ascontiguousarray(dl.Dg.T))
    L = oracle.dealias_linear(dl, *up, *U)
    A = oracle.dealias_adjoint(bad, *a, *U)
    lhs = sum(np.dot(a[c], L[c]) for c in range(3))
    rhs = sum(np.dot(up[c], A[c]) for c in range(3))
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


def test_forward_is_the_quadratic_form_of_the_linear_operator(case, dl):
    rng = np.random.default_rng(7)
    U = [rng.standard_normal(case.n) for _ in range(3)]
    up = [rng.standard_normal(case.n) for _ in range(3)]
    eps = 0.37
    F_sum = oracle.dealias_advection(dl, *[U[c] + eps * up[c] for c in range(3)])
    F_U = oracle.dealias_advection(dl, *U)
    F_p = oracle.dealias_advection(dl, *up)
    L = oracle.dealias_linear(dl, *up, *U)
    pred = [F_U[c] + eps * L[c] + eps * eps * F_p[c] for c in range(3)]
    err = _rel(F_sum, pred)
    print(f"F(U + e u') against F(U) + e L + e^2 F(u'): {err:.1e}")
    assert err < 1e-13


def test_subtracts_from_existing_forcing(case, dl):
    rng = np.random.default_rng(8)
    u, v, w = (rng.standard_normal(case.n) for _ in range(3))
    f0 = [rng.standard_normal(case.n) for _ in range(3)]
    a = oracle.dealias_advection(dl, u, v, w, f=f0)
    b = oracle.dealias_advection(dl, u, v, w)
    for c in range(3):
        assert np.allclose(a[c], f0[c] + b[c], rtol=0, atol=1e-13 * np.abs(b[c]).max())


def test_p5_rule_size():
    c = oracle.build_case(2, 2, 2, 6, jitter=0.05, seed=3)
    d = oracle.dealias_setup(c)
    assert d.lxd == 9 and d.cof.size == 9 * c.nelv * 9 ** 3
    # GLL with 6 points is exact to degree 9: linear fields (integrand degree
    # 5 + 1 + 2 = 8 per direction) keep both rules exact, quadratic ones not
    rng = np.random.default_rng(9)
    u, v, w = (a[0] + a[1] * c.x + a[2] * c.y + a[3] * c.zc
               for a in rng.standard_normal((3, 4)))
    assert _rel(oracle.dealias_advection(d, u, v, w), oracle.advection(c, u, v, w)) < 1e-12
    q = _quadratic_fields(c, 9)
    assert _rel(oracle.dealias_advection(d, *q), oracle.advection(c, *q)) > 1e-9
