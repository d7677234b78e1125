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

Pins of the flexible CG oracle (krylov.fcg, Notay 2000):
  - with a fixed symmetric preconditioner (Jacobi) its iterates are PCG's:
    the same iteration count as the C oracle's CG and the same x;
  - with a preconditioner that changes every iteration, consecutive search
    directions stay A-conjugate, <p_k, A p_{k-1}> = 0 (to rounding), and
    the solve converges; the same run with the Fletcher-Reeves beta (the
    plain CG formula, a plausible slip) breaks that conjugacy.
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


def test_fcg_with_jacobi_is_pcg(prob):
    c, mask, f, d, A = prob
    x_ref, it_ref, r0_ref, _ = c.cg(f, mask, abstol=1e-10, max_iter=2000)
    x, it, r0, r1, _ = krylov.fcg(A, lambda r, k: d * r, f, c.mult, c.volume, abstol=1e-10,
                                  max_iter=2000)
    assert abs(it - it_ref) <= 1
    assert abs(r0 - r0_ref) < 1e-12 * r0_ref
    assert np.abs(x - x_ref).max() < 1e-8 * np.abs(x_ref).max()


def _variable_pc(d, seed):
    rng = np.random.default_rng(seed)
    scales = {}

    def pc(r, k):
        if k not in scales:
            scales[k] = 1.0 + 0.5 * rng.uniform(-1, 1, r.size)
        return d * scales[k] * r
    return pc


def test_fcg_keeps_consecutive_directions_conjugate(prob):
    c, mask, f, d, A = prob
    dot = lambda a, b: float(np.dot(a * b, c.mult))
    x, it, r0, r1, hist, P, W = krylov.fcg(A, _variable_pc(d, 9), f, c.mult, c.volume,
                                           abstol=1e-9, max_iter=3000, trace=True)
    assert r1 < 1e-9
    worst = max(abs(dot(P[k], W[k - 1])) / np.sqrt(dot(P[k], W[k]) * dot(P[k - 1], W[k - 1]))
                for k in range(1, min(len(P), 60)))
    assert worst < 1e-10, worst
    # the Fletcher-Reeves slip: beta = <z_k, r_k> / <z_{k-1}, r_{k-1}>
    pc = _variable_pc(d, 9)
    r = f.copy()
    z = pc(r, 0)
    p = z.copy()
    rz = dot(z, r)
    cos = []
    for k in range(1, 30):
        w = A(p)
        al = rz / dot(p, w)
        r = r - al * w
        z = pc(r, k)
        rz_new = dot(z, r)
        p_new = z + rz_new / rz * p
        cos.append(abs(dot(p_new, w)) / np.sqrt(dot(p_new, A(p_new)) * dot(p, w)))
        p, rz = p_new, rz_new
    assert max(cos) > 1e-3
