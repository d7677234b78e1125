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
        x_m = x, b_m = A x (gs-assembled, masked);  two passes of classical
        Gram-Schmidt in the A-inner product against x_0..x_{m-1};  keep the
        vector (scaled to A-norm 1) unless its A-norm fell below
        tol * (its A-norm before orthogonalisation), tol = 1e-7.
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
        self.xbar = alpha @ self.xx[:m]
        b -= alpha @ self.bb[:m]
        return b

    def post(self, x):
        """x -> x + xbar in place; the new solution joins the basis."""
        if self.m > 0:
            x += self.xbar
        if self.m == self.L:
            self.rebase()
        k = self.m
        self.xx[k] = x
        self.bb[k] = self.A(x)
        nrm0 = np.sqrt(max(self.dot(self.xx[k], self.bb[k]), 0.0))
        coef = np.zeros(k)
        for _ in range(2):
            if k == 0:
                break
            alpha = np.array([self.dot(self.xx[k], self.bb[i]) for i in range(k)])
            self.xx[k] -= alpha @ self.xx[:k]
            self.bb[k] -= alpha @ self.bb[:k]
            coef += alpha
        nrm = np.sqrt(max(self.dot(self.xx[k], self.bb[k]), 0.0))
        if nrm > PROJ_TOL * nrm0 and nrm > 0.0:
            self.xx[k] /= nrm
            self.bb[k] /= nrm
            self.R[:, k] = 0.0
            self.R[:k, k] = coef
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
