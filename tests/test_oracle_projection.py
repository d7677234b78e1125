"""The oracle's solution projection (oracle/projection.py) pinned by the
properties Fischer's method fixes: Galerkin orthogonality of the projected
right-hand side, an A-orthonormal basis with b_i = A x_i, exact recovery of
solutions inside the span, A-norm optimality of the initial guess against a
dense least-squares solve, the restart when the basis is full and the
rejection of a dependent vector; the restart that keeps the newest
solutions: the recorded coefficients reproduce every solution, and after
the restart the basis is A-orthonormal with b_i = A x_i and spans exactly
the kept solutions (each is recovered by the projection, an older one is
not)."""
import numpy as np
import pytest

import oracle
from oracle.projection import Projection


@pytest.fixture(scope="module")
def prob():
    c = oracle.build_case(4, 3, 2, 5, jitter=0.05, seed=7)
    mask = oracle.face_mask(4, 3, 2, 5, oracle.PRS_DIRICHLET)

    def A(x):
        w = c.gs(c.ax(x))
        w[mask] = 0.0
        return w

    def field(rng):
        x = c.gs(rng.standard_normal(c.n)) * c.mult   # continuous
        x[mask] = 0.0
        return x

    return c, mask, A, field


def _feed(P, A, x):
    """Use the projection as a solver would: pre on b = A x, the correction
    dx with A dx = b' is x - xbar, post(dx) restores x."""
    b = A(x)
    P.pre(b)
    dx = x - (P.xbar if P.m > 0 else 0.0)
    return P.post(dx)


def test_basis_is_A_orthonormal_and_b_is_Ax(prob):
    c, mask, A, field = prob
    rng = np.random.default_rng(1)
    P = Projection(c.n, 6, A, c.mult)
    xs = [field(rng) for _ in range(5)]
    for x in xs:
        out = _feed(P, A, x)
        assert np.allclose(out, x, rtol=0, atol=1e-12 * np.abs(x).max())
    assert P.m == 5
    G = np.array([[P.dot(P.xx[i], P.bb[j]) for j in range(5)] for i in range(5)])
    assert np.abs(G - np.eye(5)).max() < 1e-10
    for i in range(5):
        Ax = A(P.xx[i])
        assert np.abs(Ax - P.bb[i]).max() < 1e-10 * np.abs(Ax).max()


def test_projected_rhs_is_orthogonal_to_the_basis(prob):
    c, mask, A, field = prob
    rng = np.random.default_rng(2)
    P = Projection(c.n, 8, A, c.mult)
    for _ in range(4):
        _feed(P, A, field(rng))
    b = A(field(rng))
    nb = np.sqrt(P.dot(b, b))
    P.pre(b)
    for i in range(P.m):
        nx = np.sqrt(P.dot(P.xx[i], P.xx[i]))
        assert abs(P.dot(b, P.xx[i])) < 1e-10 * nb * nx


def test_solution_in_span_is_recovered(prob):
    c, mask, A, field = prob
    rng = np.random.default_rng(3)
    P = Projection(c.n, 8, A, c.mult)
    x1, x2 = field(rng), field(rng)
    _feed(P, A, x1)
    _feed(P, A, x2)
    xt = 2.0 * x1 - 3.0 * x2
    b = A(xt)
    nb = np.abs(b).max()
    P.pre(b)
    assert np.abs(P.xbar - xt).max() < 1e-9 * np.abs(xt).max()
    assert np.abs(b).max() < 1e-9 * nb


def test_guess_minimises_the_A_norm_error(prob):
    """xbar = argmin over span{x_i} of ||x* - y||_A, from the dense Galerkin
    system on the raw (not orthonormalised) vectors."""
    c, mask, A, field = prob
    rng = np.random.default_rng(4)
    P = Projection(c.n, 8, A, c.mult)
    X = [field(rng) for _ in range(4)]
    for x in X:
        _feed(P, A, x)
    xs = field(rng)
    b = A(xs)
    M = np.array([[P.dot(xi, A(xj)) for xj in X] for xi in X])
    rhs = np.array([P.dot(xi, b) for xi in X])
    y = np.linalg.solve(M, rhs) @ np.array(X)
    P.pre(b.copy())
    assert np.abs(P.xbar - y).max() < 1e-8 * np.abs(y).max()


def test_restart_when_full_and_dependent_vector_dropped(prob):
    c, mask, A, field = prob
    rng = np.random.default_rng(5)
    P = Projection(c.n, 3, A, c.mult)
    xs = [field(rng) for _ in range(4)]
    for x in xs[:3]:
        _feed(P, A, x)
    assert P.m == 3
    _feed(P, A, xs[3])          # full: restart with the newest solution only
    assert P.m == 1
    nrm = np.sqrt(P.dot(xs[3], A(xs[3])))
    assert np.abs(P.xx[0] - xs[3] / nrm).max() < 1e-12 * np.abs(xs[3] / nrm).max()
    _feed(P, A, 5.0 * xs[3])    # inside the span: rejected, m unchanged
    assert P.m == 1


def test_recorded_coefficients_reproduce_the_solutions(prob):
    """x_j = sum_i R_ij x_i for every accepted solution: a wrong sign or a
    dropped Gram-Schmidt pass in the record fails this."""
    c, mask, A, field = prob
    rng = np.random.default_rng(6)
    P = Projection(c.n, 6, A, c.mult, keep=2)
    base = field(rng)
    xs = [base * (1 + 0.1 * j) + 0.3 * field(rng) for j in range(5)]
    for x in xs:
        _feed(P, A, x)
    assert P.m == 5
    for j, x in enumerate(xs):
        y = P.R[:5, j] @ P.xx[:5]
        assert np.abs(y - x).max() < 1e-10 * np.abs(x).max(), j


@pytest.mark.parametrize("keep", [1, 2, 3])
def test_restart_keeps_the_span_of_the_newest_solutions(prob, keep):
    c, mask, A, field = prob
    rng = np.random.default_rng(7)
    L = 4
    P = Projection(c.n, L, A, c.mult, keep=keep)
    base = field(rng)
    xs = [base * (1 + 0.1 * j) + 0.3 * field(rng) for j in range(L + 1)]
    for x in xs[:L]:
        _feed(P, A, x)
    assert P.m == L
    _feed(P, A, xs[L])           # full: rebase on xs[L-keep:L], then add xs[L]
    assert P.m == keep + 1
    G = np.array([[P.dot(P.xx[i], P.bb[j]) for j in range(P.m)] for i in range(P.m)])
    assert np.abs(G - np.eye(P.m)).max() < 1e-10
    for i in range(P.m):
        Ax = A(P.xx[i])
        assert np.abs(Ax - P.bb[i]).max() < 1e-10 * np.abs(Ax).max()
    # the kept solutions and the newest one are in the span ...
    for x in xs[L - keep:]:
        b = A(x)
        nb = np.abs(b).max()
        P.pre(b)
        assert np.abs(P.xbar - x).max() < 1e-9 * np.abs(x).max()
        assert np.abs(b).max() < 1e-9 * nb
    # ... an older one is not
    b = A(xs[L - keep - 1])
    nb = np.abs(b).max()
    P.pre(b)
    assert np.abs(b).max() > 1e-3 * nb
