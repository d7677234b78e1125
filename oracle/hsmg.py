"""Hybrid Schwarz multigrid preconditioner (Neko's pc_hsmg, the 'hsmg' of
the shipped cases' pressure solver, rugby_ball.case:46-52), restated from
the method it implements: Lottes & Fischer, "Hybrid multigrid/Schwarz
algorithms for the spectral element method", J. Sci. Comput. 24 (2005).
Neko is not vendored, so the restatement follows the published algorithm
with Neko's level structure; it is pinned by the properties in
tests/test_oracle_hsmg.py, not against Neko's code.

Levels: fine lx, mid lx_mid = 4 (lx >= 5; else max(lx - 1, 3)), coarse 2.
One application z = M^-1 r (additive across levels):
  z  = S_fine(r)
  r2 = gs_mid(J_fm^T (mult_fine r));           e2 = S_mid(r2)
  r1 = mask(gs_crs(J_mc^T (mult_mid r2)));     e1 = mask(CG_crs(r1), 10 its)
  e2 += J_mc e1;  z += J_fm e2;  z = mult_fine gs_fine(z)
J_fm, J_mc: Lagrange interpolation from the coarser GLL points to the finer.

Overlapping Schwarz S on a level (one-point overlap): each element is
extended by one layer of points on every face (the neighbour's first
interior layer); r is masked, copied into the extended element, the ghost
layers filled with the neighbours' values by a gs on the (lx + 2)-point
lattice (ghost := own layer 2; gs; ghost -= own layer 2), the local problem
solved by fast diagonalisation, the ghost results sent back to the
neighbours the same way and added to their layer 2, restricted to the
element, gs-summed, masked and weighted by 1 / (the number of subdomains
covering each point).

Fast diagonalisation: per element and direction d the 1-D extended
operators A_d = (2 / L_d) A_hat and B_d = (L_d / 2) diag(w) on the element's
points, plus the neighbour's 1-D matrices (scaled with its length)
restricted to its two points next to the shared face; the ghost point is
dropped where there is no neighbour, and the boundary point too where the
face is Dirichlet.  Generalised eigenproblem A_d s = lambda B_d s with
S^T B S = I; the local solve is
  u = (S_t x S_s x S_r) diag(1 / (lam_r + lam_s + lam_t)) (S_t x S_s x S_r)^T f
with the zero mode (a pure-Neumann element) dropped.  Element lengths: the
mean length of the element's four edges in each direction.

TEST INFRASTRUCTURE ONLY (the oracle).
"""
from __future__ import annotations

import numpy as np

import oracle


def lx_mid_of(lx):
    return 4 if lx >= 5 else max(lx - 1, 3)


def _one_d(lx):
    z, w, D = oracle.gll(lx)
    Ah = D.T @ np.diag(w) @ D
    return z, w, Ah


def element_lengths(c):
    cor = c.corners.reshape(c.nelv, 8, 3)
    L = np.zeros((c.nelv, 3))
    for d in range(3):
        bit = 1 << d
        acc = 0.0
        for k in range(8):
            if k & bit:
                continue
            acc = acc + np.linalg.norm(cor[:, k | bit] - cor[:, k], axis=1)
        L[:, d] = acc / 4.0
    return L


def _tensor(M0, M1, M2, u, n_in, n_out):
    """v[k,j,i] = sum M2[k,c] M1[j,b] M0[i,a] u[c,b,a] for one element
    (arrays indexed [t][s][r], r fastest)."""
    u = u.reshape(n_in, n_in, n_in)
    v = np.einsum("ia,cba->cbi", M0, u)
    v = np.einsum("jb,cbi->cji", M1, v)
    v = np.einsum("kc,cji->kji", M2, v)
    return v.reshape(-1)


class Schwarz:
    """Overlapping Schwarz with FDM local solves on one level (case c)."""

    def __init__(self, c, dirichlet_faces):
        self.c = c
        lx = c.lx
        self.lx, self.lxe = lx, lx + 2
        self.nelv = c.nelv
        self.gid_ext = np.zeros(self.nelv * self.lxe ** 3, dtype=np.int64)
        oracle.lib().or_dofmap_gids(c.nx, c.ny, c.nz, self.lxe, self.gid_ext)
        self.gs_ext = oracle.gs_build(self.gid_ext)
        self.mask = oracle.face_mask(c.nx, c.ny, c.nz, lx, dirichlet_faces)
        self.dirichlet_faces = dirichlet_faces
        self._setup_fdm()
        ones = np.ones(c.n)
        cnt = self._sum_overlap(self._exchange(self._to_ext(ones)))
        self.wt = 1.0 / cnt

    # -- data movement -----------------------------------------------------
    def _gs_ext(self, a):
        idx, m, cnt = self.gs_ext
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        oracle.lib().or_gs_op(a, len(m), m, cnt, idx, oracle.GS_ADD)
        return a

    def _to_ext(self, r):
        lx, lxe, ne = self.lx, self.lxe, self.nelv
        W = np.zeros((ne, lxe, lxe, lxe))
        W[:, 1:lx + 1, 1:lx + 1, 1:lx + 1] = r.reshape(ne, lx, lx, lx)
        return W

    def _faces(self, W, l):
        """views of the six face layers at depth l (index l and lxe-1-l),
        over the original points' range [1, lx] in the other directions"""
        lx = self.lx
        s = slice(1, lx + 1)
        e = self.lxe - 1 - l
        return [(np.s_[:, s, s, l], np.s_[:, s, s, e]),
                (np.s_[:, s, l, s], np.s_[:, s, e, s]),
                (np.s_[:, l, s, s], np.s_[:, e, s, s])]

    def _extrude(self, A, l1, f1, B, l2, f2):
        for (a0, a1), (b0, b1) in zip(self._faces(A, l1), self._faces(B, l2)):
            A[a0] = f1 * A[a0] + f2 * B[b0]
            A[a1] = f1 * A[a1] + f2 * B[b1]

    def _exchange(self, W):
        """ghost layers := the neighbours' first interior layer"""
        self._extrude(W, 0, 0.0, W, 2, 1.0)
        W = self._gs_ext(W.reshape(-1)).reshape(W.shape)
        self._extrude(W, 0, 1.0, W, 2, -1.0)
        return W

    def _sum_overlap(self, W2):
        """send the ghost results to the neighbours, restrict, gs"""
        lx = self.lx
        W1 = np.zeros_like(W2)
        self._extrude(W1, 0, 0.0, W2, 0, 1.0)
        W2 = self._gs_ext(W2.reshape(-1)).reshape(W2.shape)
        self._extrude(W2, 0, 1.0, W1, 0, -1.0)
        self._extrude(W2, 2, 1.0, W2, 0, 1.0)
        e = W2[:, 1:lx + 1, 1:lx + 1, 1:lx + 1].reshape(-1).copy()
        return self.c.gs(e)

    # -- fast diagonalisation ------------------------------------------------
    def _setup_fdm(self):
        c, lx, lxe, ne = self.c, self.lx, self.lxe, self.nelv
        _, w, Ah = _one_d(lx)
        L = element_lengths(c)
        dims = (c.nx, c.ny, c.nz)
        self.S = np.zeros((ne, 3, lxe, lxe))
        self.lam = np.zeros((ne, 3, lxe))
        for e in range(ne):
            ec = (e % c.nx, (e // c.nx) % c.ny, e // (c.nx * c.ny))
            for d in range(3):
                stride = (1, c.nx, c.nx * c.ny)[d]
                lo = ec[d] > 0
                hi = ec[d] < dims[d] - 1
                A = np.zeros((lxe, lxe))
                Bm = np.zeros(lxe)
                Le = L[e, d]
                A[1:lx + 1, 1:lx + 1] += (2.0 / Le) * Ah
                Bm[1:lx + 1] += (Le / 2.0) * w
                if lo:
                    Ll = L[e - stride, d]
                    A[0:2, 0:2] += (2.0 / Ll) * Ah[lx - 2:lx, lx - 2:lx]
                    Bm[0:2] += (Ll / 2.0) * w[lx - 2:lx]
                if hi:
                    Lr = L[e + stride, d]
                    A[lxe - 2:, lxe - 2:] += (2.0 / Lr) * Ah[0:2, 0:2]
                    Bm[lxe - 2:] += (Lr / 2.0) * w[0:2]
                act = np.ones(lxe, dtype=bool)
                if not lo:
                    act[0] = False
                    if self.dirichlet_faces & (1 << (2 * d)):
                        act[1] = False
                if not hi:
                    act[lxe - 1] = False
                    if self.dirichlet_faces & (1 << (2 * d + 1)):
                        act[lxe - 2] = False
                ia = np.flatnonzero(act)
                Aa = A[np.ix_(ia, ia)]
                bh = 1.0 / np.sqrt(Bm[ia])
                lam, Q = np.linalg.eigh(bh[:, None] * Aa * bh[None, :])
                S = np.zeros((lxe, lxe))
                S[np.ix_(ia, np.arange(ia.size))] = bh[:, None] * Q
                self.S[e, d] = S
                self.lam[e, d, :ia.size] = lam
                self.lam[e, d, ia.size:] = np.inf  # inactive modes: no contribution
        self.active = None

    def fdm(self, W):
        lxe, ne = self.lxe, self.nelv
        out = np.zeros_like(W).reshape(ne, -1)
        for e in range(ne):
            Sr, Ss, St = self.S[e]
            f = W[e].reshape(-1)
            g = _tensor(Sr.T, Ss.T, St.T, f, lxe, lxe).reshape(lxe, lxe, lxe)
            lr, ls, lt = self.lam[e]
            den = lt[:, None, None] + ls[None, :, None] + lr[None, None, :]
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = np.where(np.isfinite(den) & (np.abs(den) > 1e-12 * np.abs(
                    np.where(np.isfinite(den), den, 0)).max()), 1.0 / den, 0.0)
            g = (g * inv).reshape(-1)
            out[e] = _tensor(Sr, Ss, St, g, lxe, lxe)
        return out.reshape(W.shape)

    # -- the smoother --------------------------------------------------------
    def apply(self, r):
        rm = np.array(r, dtype=np.float64)
        rm[self.mask] = 0.0
        W1 = self._exchange(self._to_ext(rm))
        W2 = self.fdm(W1)
        e = self._sum_overlap(W2)
        e[self.mask] = 0.0
        return e * self.wt


class HSMG:
    def __init__(self, nx, ny, nz, lx, jitter=0.05, seed=1234, dirichlet_faces=None,
                 crs_iter=10):
        faces = oracle.PRS_DIRICHLET if dirichlet_faces is None else dirichlet_faces
        lm = lx_mid_of(lx)
        self.lx, self.lx_mid = lx, lm
        self.c3 = oracle.build_case(nx, ny, nz, lx, jitter=jitter, seed=seed)
        self.c2 = oracle.build_case(nx, ny, nz, lm, jitter=jitter, seed=seed)
        self.c1 = oracle.build_case(nx, ny, nz, 2, jitter=jitter, seed=seed)
        self.J_fm = oracle.lagrange_matrix(self.c2.z, self.c3.z)   # lx x lm
        self.J_mc = oracle.lagrange_matrix(self.c1.z, self.c2.z)   # lm x 2
        self.s3 = Schwarz(self.c3, faces)
        self.s2 = Schwarz(self.c2, faces)
        self.mask1 = oracle.face_mask(nx, ny, nz, 2, faces)
        self.d1 = self.c1.jacobi()
        self.crs_iter = crs_iter

    def _restrict(self, J, r, cf, cc):
        """per element J^T (mult_f r), then gs on the coarser level"""
        ne = cf.nelv
        nf, nc = cf.lx, cc.lx
        out = np.zeros(cc.n).reshape(ne, -1)
        rm = (cf.mult * r).reshape(ne, -1)
        for e in range(ne):
            out[e] = _tensor(J.T, J.T, J.T, rm[e], nf, nc)
        return cc.gs(out.reshape(-1))

    def _prolong(self, J, e_c, cc, cf):
        ne = cc.nelv
        out = np.zeros(cf.n).reshape(ne, -1)
        ec = e_c.reshape(ne, -1)
        for e in range(ne):
            out[e] = _tensor(J, J, J, ec[e], cc.lx, cf.lx)
        return out.reshape(-1)

    def coarse_solve(self, r1):
        x, _, _, _ = self.c1.cg(r1, self.mask1, abstol=0.0, max_iter=self.crs_iter,
                                dinv=self.d1)
        return x

    def apply(self, r):
        z = self.s3.apply(r)
        r2 = self._restrict(self.J_fm, r, self.c3, self.c2)
        e2 = self.s2.apply(r2)
        r1 = self._restrict(self.J_mc, r2, self.c2, self.c1)
        r1[self.mask1] = 0.0
        e1 = self.coarse_solve(r1)
        e1[self.mask1] = 0.0
        e2 = e2 + self._prolong(self.J_mc, e1, self.c1, self.c2)
        z = z + self._prolong(self.J_fm, e2, self.c2, self.c3)
        return self.c3.mult * self.c3.gs(z)
