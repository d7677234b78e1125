"""Test infrastructure only: the CPU restatement of the solution projection
for successive right-hand sides (Fischer, "Projection techniques for
iterative solution of Ax = b with successive right-hand sides", CMAME 163,
1998), the algorithm of Neko's projection_t.

The reference declares and drives it but does not vendor it:
  proj_prs, proj_u/v/w           sources/adjoint/adjoint_pnpn.f90:102-105
  init(n, L, activ_step)         sources/adjoint/adjoint_pnpn.f90:365-373
  pre_solving(b, tstep, ...)     sources/adjoint/adjoint_pnpn.f90:729-730, :777-779
  post_solving(x, Ax, ...)       sources/adjoint/adjoint_pnpn.f90:739-740, :788-793
  size default 20, activation    sources/adjoint/adjoint_scheme.f90:385-392
Restated here from the published method:
  pre:  alpha_i = <b, x_i>_mult (i < m);  xbar = sum alpha_i x_i;
        b <- b - sum alpha_i (A x_i)
  post: x <- x + xbar;  if m == L: m = 0 (restart from the newest solution);
        the new vector v, b_v = A v (gs-assembled, masked);  classical
        Gram-Schmidt in the A-inner product against x_0..x_{m-1};  keep the
        vector (scaled to A-norm 1) unless its A-norm fell below
        tol * (its A-norm before orthogonalisation), tol = 1e-7.
The new vector: Neko orthogonalises the full solution x, twice.  Here, when
the basis was not restarted in this post, v is the solver's correction
dx = x - xbar (x and dx span the same space with the basis, as xbar lies in
it): dx is A-orthogonal to the basis up to the solve's residual, so one pass
leaves it A-orthogonal to rounding, where x (|x| >> |dx| once the basis is
good) cancels to |dx| and needs a second pass.  A second pass runs whenever
the first removed more than half of v's A-norm (so always for the full
solution after a restart).  The dependence test compares the A-norm after
Gram-Schmidt with the full solution's: for v = dx that is
||x||_A^2 = sum_i a_i^2 + 2 sum_i a_i d_i + ||dx||_A^2 (a: the pre
coefficients, d: the first pass's, in the A-orthonormal basis).
All inner products are mult-weighted (the gs-assembled <.,.> of Neko's
glsc3(a, b, mult)).

Restart that keeps the newest solutions (keep > 0; not in Neko, whose
restart is keep = 0): each accepted solution x_j is recorded by its
coefficients in the A-orthonormal basis, column j of an upper-triangular
R (x_j = sum_i R_ij x_i: the two Gram-Schmidt passes' coefficients summed,
and its A-norm after them on the diagonal).  When the basis is full, the
last `keep` columns R_q (the newest solutions) are factored R_q = Q T
(thin QR), the basis becomes x'_j = sum_i Q_ij x_i (with b'_j likewise:
A-orthonormal because Q has orthonormal columns, spanning the newest
solutions because T is nonsingular) and R restarts as T.  The new solution
then joins as usual.

Parity of this file is pinned by the properties the method fixes, checked in
tests/test_oracle_projection.py: the projected b is orthogonal to the basis,
the basis is A-orthonormal with b_i = A x_i, a solution inside the span is
recovered exactly, and the projected initial guess minimises the A-norm of
the error over the span (checked against a dense least-squares solve).
"""
import numpy as np

PROJ_TOL = 1e-7


class Projection:
    def __init__(self, n, L, apply_A, mult, keep=0):
        """apply_A(x) -> mask(QQ^T A x); mult: inverse multiplicity;
        keep: solutions kept when the full basis restarts (0 <= keep < L)."""
        assert 0 <= keep < max(L, 1)
        self.n, self.L, self.keep = n, L, keep
        self.A = apply_A
        self.mult = mult
        self.xx = np.zeros((L, n))
        self.bb = np.zeros((L, n))
        self.R = np.zeros((L, L))
        self.xbar = np.zeros(n)
        self.m = 0

    def dot(self, a, b):
        return float(np.sum(a * b * self.mult))

    def clear(self):
        self.m = 0

    def pre(self, b):
        """b -> b - A xbar in place; xbar kept for post."""
        m = self.m
        if m == 0:
            return b
        alpha = np.array([self.dot(b, self.xx[i]) for i in range(m)])
        self.alpha = alpha
        self.xbar = alpha @ self.xx[:m]
        b -= alpha @ self.bb[:m]
        return b

    def post(self, x):
        """x (the solver's correction) -> x + xbar in place; the new solution
        joins the basis."""
        dx = x.copy()
        had_pre = self.m > 0
        coef = np.zeros(self.L)
        if had_pre:
            x += self.xbar
            coef[:self.m] = self.alpha
        restarted = self.m == self.L
        if restarted:
            self.rebase()
            coef[:] = 0.0
        k = self.m
        v = dx if (had_pre and not restarted) else x
        self.xx[k] = v
        self.bb[k] = self.A(v)
        nv2 = self.dot(self.xx[k], self.bb[k])
        nrm0 = np.sqrt(max(nv2, 0.0))
        nfull = nrm0
        nrm = nrm0
        for p in range(2):
            if k == 0 or (p == 1 and nrm >= 0.5 * nrm0):
                break
            alpha = np.array([self.dot(self.xx[k], self.bb[i]) for i in range(k)])
            if p == 0 and v is dx:
                a = coef[:k]
                nfull = np.sqrt(max(np.sum(a * a) + 2.0 * np.sum(a * alpha) + nv2, 0.0))
            self.xx[k] -= alpha @ self.xx[:k]
            self.bb[k] -= alpha @ self.bb[:k]
            coef[:k] += alpha
            nrm = np.sqrt(max(self.dot(self.xx[k], self.bb[k]), 0.0))
        if nrm > PROJ_TOL * nfull and nrm > 0.0:
            self.xx[k] /= nrm
            self.bb[k] /= nrm
            self.R[:, k] = 0.0
            self.R[:k, k] = coef[:k]
            self.R[k, k] = nrm
            self.m = k + 1
        return x

    def rebase(self):
        """The full basis restarts from the span of the `keep` newest
        solutions (keep = 0: empty, Neko's restart)."""
        m, q = self.m, self.keep
        if q == 0:
            self.m = 0
            return
        Q, T = np.linalg.qr(self.R[:m, m - q:m])
        self.xx[:q] = Q.T @ self.xx[:m]
        self.bb[:q] = Q.T @ self.bb[:m]
        self.R[:] = 0.0
        self.R[:q, :q] = T
        self.m = q
