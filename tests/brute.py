"""Independent numpy brute force used to pin the oracle.

Nothing here reuses the oracle's formulas: GLL nodes come from numpy's
Legendre roots, the derivative matrix from barycentric Lagrange weights, the
element geometry from the analytic derivatives of the trilinear map and a
numpy 3x3 inverse, and the element stiffness is assembled as a dense matrix
K_e = sum_a P_a^T diag(w3 |J| h1) P_a from the physical gradient operators.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as leg


def gll_numpy(lx):
    N = lx - 1
    cN = np.zeros(N + 1)
    cN[N] = 1.0
    if N >= 2:
        r = np.sort(np.real(leg.legroots(leg.legder(cN))))
    else:
        r = np.array([])
    z = np.concatenate([[-1.0], r, [1.0]])
    PN = leg.legval(z, cN)
    w = 2.0 / (N * (N + 1) * PN ** 2)
    lam = np.array([1.0 / np.prod([z[j] - z[k] for k in range(lx) if k != j]) for j in range(lx)])
    D = np.zeros((lx, lx))
    for i in range(lx):
        for j in range(lx):
            if i != j:
                D[i, j] = (lam[j] / lam[i]) / (z[i] - z[j])
        D[i, i] = -np.sum(D[i, :])
    return z, w, D


def trilinear(Xc, r, s, t):
    """x(r,s,t) and dx/d(r,s,t) of the trilinear map of 8 corners Xc[c,3],
    corner c = cx + 2 cy + 4 cz.  Returns (x[3], J[3,3] with J[a,b] = dx_a/dxi_b)."""
    x = np.zeros(3)
    J = np.zeros((3, 3))
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        fr = 0.5 * (1 + r) if cx else 0.5 * (1 - r)
        fs = 0.5 * (1 + s) if cy else 0.5 * (1 - s)
        ft = 0.5 * (1 + t) if cz else 0.5 * (1 - t)
        dr = 0.5 if cx else -0.5
        ds = 0.5 if cy else -0.5
        dt = 0.5 if cz else -0.5
        x += fr * fs * ft * Xc[c]
        J[:, 0] += dr * fs * ft * Xc[c]
        J[:, 1] += fr * ds * ft * Xc[c]
        J[:, 2] += fr * fs * dt * Xc[c]
    return x, J


def element_geometry(Xc, z, w):
    lx = len(z)
    npt = lx ** 3
    xyz = np.zeros((npt, 3))
    inv = np.zeros((npt, 3, 3))
    det = np.zeros(npt)
    w3 = np.zeros(npt)
    for k in range(lx):
        for j in range(lx):
            for i in range(lx):
                p = i + lx * (j + lx * k)
                x, J = trilinear(Xc, z[i], z[j], z[k])
                xyz[p] = x
                inv[p] = np.linalg.inv(J)  # inv[b, a] = dxi_b / dx_a
                det[p] = np.linalg.det(J)
                w3[p] = w[i] * w[j] * w[k]
    return xyz, inv, det, w3


def element_matrices(Xc, z, w, D, h1=None):
    """Dense element stiffness K (h1-weighted) and diagonal mass of one hex."""
    lx = len(z)
    npt = lx ** 3
    I = np.eye(lx)
    Dr = np.kron(I, np.kron(I, D))
    Ds = np.kron(I, np.kron(D, I))
    Dt = np.kron(D, np.kron(I, I))
    xyz, inv, det, w3 = element_geometry(Xc, z, w)
    h1 = np.ones(npt) if h1 is None else h1
    K = np.zeros((npt, npt))
    for a in range(3):
        Pa = (inv[:, 0, a][:, None] * Dr + inv[:, 1, a][:, None] * Ds
              + inv[:, 2, a][:, None] * Dt)
        K += Pa.T @ ((w3 * det * h1)[:, None] * Pa)
    Gab = np.einsum("pax,pbx->pab", inv, inv) * (w3 * det)[:, None, None]
    return K, w3 * det, xyz, Gab


def gs_numpy(u, gid):
    ng = int(gid.max()) + 1
    s = np.zeros(ng)
    np.add.at(s, gid, u)
    return s[gid]


def global_matrix(case, h1=None, h2=None):
    """Assembled stiffness (+ h2 mass) on unique dofs, dense."""
    lx, nelv = case.lx, case.nelv
    npt = lx ** 3
    ng = int(case.gid.max()) + 1
    A = np.zeros((ng, ng))
    z, w, D = gll_numpy(lx)
    for e in range(nelv):
        Xc = case.corners[e * 24:(e + 1) * 24].reshape(8, 3)
        sl = slice(e * npt, (e + 1) * npt)
        K, Bm, _, _ = element_matrices(Xc, z, w, D,
                                       None if h1 is None else h1[sl])
        if h2 is not None:
            K = K + np.diag(h2[sl] * Bm)
        g = case.gid[sl]
        A[np.ix_(g, g)] += K
    return A
