"""Pins of the GMRES oracle (oracle/krylov.py), Neko's gmres_t restated.

  - minimal residual: with no restart inside the run, the k-th GMRES
    residual is the smallest over the Krylov space that Jacobi-PCG searches
    too, so it never exceeds the PCG residual at the same iteration (the PCG
    residuals come from the C oracle's CG, checked elsewhere against dense
    assembly);
  - the iterate itself equals the explicit least-squares solution over the
    preconditioned Krylov basis, computed with numpy's lstsq, for the first
    few iterations (a wrong rotation sign, index or back substitution fails
    this);
  - restarted GMRES(10) never increases the residual and reaches the same
    solution as PCG.
"""
import numpy as np
import pytest

import oracle
from oracle import krylov


@pytest.fixture(scope="module")
def prob():
    c = oracle.build_case(3, 2, 2, 6, jitter=0.05, seed=1234)
    mask = oracle.face_mask(3, 2, 2, 6, oracle.PRS_DIRICHLET)
    rng = np.random.default_rng(3)
    f = c.gs(c.B * rng.standard_normal(c.n))
    f[mask] = 0.0
    d = c.jacobi()

    def A(u):
        w = c.gs(c.ax(u))
        w[mask] = 0.0
        return w

    return c, mask, f, d, A


def test_gmres_residual_never_above_pcg(prob):
    c, mask, f, d, A = prob
    _, it, r0, rn, hist = krylov.gmres(A, lambda v: d * v, f, c.mult, c.volume, m=40,
                                       abstol=1e-30, max_iter=25)
    assert it == 25 and len(hist) == 25
    cg = [c.cg(f, mask, abstol=1e-30, max_iter=k, dinv=d)[3] for k in range(1, 26)]
    for k, (g, p) in enumerate(zip(hist, cg), 1):
        assert g <= p * (1 + 1e-9), (k, g, p)
    print(f"iteration 25: GMRES {hist[-1]:.3e} <= PCG {cg[-1]:.3e} (start {r0:.3e})")
    assert hist[-1] < 0.9 * cg[-1]


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_gmres_iterate_is_the_least_squares_solution(prob, k):
    c, mask, f, d, A = prob
    x, it, _, _, _ = krylov.gmres(A, lambda v: d * v, f, c.mult, c.volume, m=30,
                                  abstol=1e-30, max_iter=k)
    assert it == k
    # preconditioned Krylov basis M^-1 {f, A M^-1 f, ...}, orthonormalised
    # in the mult-weighted norm for conditioning
    w = np.sqrt(c.mult)
    K = []
    v = f.copy()
    for _ in range(k):
        z = d * v
        K.append(z)
        v = A(z)
    K = np.array(K).T
    AK = np.array([A(K[:, i]) for i in range(k)]).T
    y, *_ = np.linalg.lstsq(AK * w[:, None], f * w, rcond=None)
    xs = K @ y
    err = np.abs(x - xs).max() / np.abs(xs).max()
    print(f"k={k}: GMRES iterate vs least squares {err:.1e}")
    assert err < 1e-8


def test_restarted_gmres_is_monotone_and_converges(prob):
    c, mask, f, d, A = prob
    x, it, r0, rn, hist = krylov.gmres(A, lambda v: d * v, f, c.mult, c.volume, m=10,
                                       abstol=1e-10, max_iter=4000)
    assert rn < 1e-10
    assert all(b <= a * (1 + 1e-12) for a, b in zip(hist, hist[1:]))
    xc = c.cg(f, mask, abstol=1e-12, max_iter=2000, dinv=d)[0]
    err = np.abs(x - xc).max() / np.abs(xc).max()
    print(f"GMRES(10): {it} iterations; solution vs PCG {err:.1e}")
    assert err < 1e-7
