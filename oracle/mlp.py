"""Test infrastructure only: CPU restatement of the element-local MLP of the
learned solver aid, forward and backward, against which the HIP training and
inference kernels are checked.

The reference has no learned component (its initial-guess accelerator is
the projection of sources/adjoint/adjoint_pnpn.f90:102-105, :365-373,
:729-741); the network is this build's own and is specified by its layer
sizes: 518 inputs (the 512 GLL values of an element's residual and the
means of the residual over its 6 face neighbours, all divided by the RMS of
the residual field) -> 256 -> 256 -> 512, SiLU between layers, BF16
operands with FP32 accumulation on the device.

Parameter layout (FP32, 330,240 values): W1[256,518], b1[256], W2[256,256],
b2[256], W3[512,256], b3[512], each row-major (out x in).  On the device the
input dimension is zero-padded to KIN = 544 (a multiple of the 32-wide
K tile); the padding never carries data or gradient.

Pinned in tests/test_oracle_mlp.py: the backward against central finite
differences of the forward (float64, no BF16 rounding), SiLU and its
derivative against closed forms, the features against a direct loop over
the box mesh, and BF16 rounding against the bit pattern rule
(round-to-nearest-even on the upper 16 bits of the float32).
"""
import numpy as np

NP = 512
NIN = 518
KIN = 544
H = 256
SIZES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
         ("W3", (NP, H)), ("b3", (NP,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for k, s in SIZES:
        m = int(np.prod(s))
        out[k] = np.asarray(theta[o:o + m], dtype=np.float64).reshape(s)
        o += m
    return out


def pack(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k, _ in SIZES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as
    float64.  NaN/Inf do not occur here."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def neighbours(nx, ny, nz):
    """[nelv, 6] face neighbours (-x, +x, -y, +y, -z, +z) of the x-fastest
    box numbering e = ex + nx (ey + ny ez); -1 on the boundary."""
    nelv = nx * ny * nz
    nb = np.full((nelv, 6), -1, dtype=np.int32)
    for e in range(nelv):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        for k, (dx, dy, dz) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
                                          (0, 0, -1), (0, 0, 1)]):
            jx, jy, jz = ex + dx, ey + dy, ez + dz
            if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz:
                nb[e, k] = jx + nx * (jy + ny * jz)
    return nb


def features(r, nb, scale=None):
    """[nelv, 518] inputs from an element-contiguous field r (float64):
    r_e / s and the 6 neighbour means / s (0 for a missing neighbour),
    s = RMS of r unless given."""
    nelv = nb.shape[0]
    if scale is None:
        scale = np.sqrt(np.mean(r * r))
    re = r.reshape(nelv, NP) / scale
    mean = re.mean(axis=1)
    nm = np.where(nb >= 0, mean[np.maximum(nb, 0)], 0.0)
    return np.concatenate([re, nm], axis=1), scale


def forward(p, X, round_bf16=True):
    """X [M, 518] -> dict with Z1, H1, Z2, H2, Y (float64).  With
    round_bf16 the GEMM operands are rounded as on the device (X, W, H)."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Xq = q(X)
    Z1 = Xq @ q(p["W1"]).T + p["b1"]
    H1 = q(silu(Z1))
    Z2 = H1 @ q(p["W2"]).T + p["b2"]
    H2 = q(silu(Z2))
    Y = H2 @ q(p["W3"]).T + p["b3"]
    return {"X": Xq, "Z1": Z1, "H1": H1, "Z2": Z2, "H2": H2, "Y": Y}


def loss_grad(p, X, T, round_bf16=True):
    """MSE loss = mean over M x 512 of (Y - T)^2 and its gradient (a dict
    shaped like p).  With round_bf16 the backward GEMM operands (dZ, W) are
    rounded as on the device."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    f = forward(p, X, round_bf16)
    M = X.shape[0]
    d = f["Y"] - T
    loss = float(np.mean(d * d))
    dZ3 = 2.0 * d / (M * NP)
    g = {"b3": dZ3.sum(0)}
    dZ3q = q(dZ3)
    g["W3"] = dZ3q.T @ f["H2"]
    dH2 = dZ3q @ q(p["W3"])
    dZ2 = dH2 * dsilu(f["Z2"])
    g["b2"] = dZ2.sum(0)
    dZ2q = q(dZ2)
    g["W2"] = dZ2q.T @ f["H1"]
    dH1 = dZ2q @ q(p["W2"])
    dZ1 = dH1 * dsilu(f["Z1"])
    g["b1"] = dZ1.sum(0)
    g["W1"] = q(dZ1).T @ f["X"]
    return loss, g


def adam(theta, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba 2015) with bias correction, in place."""
    m *= b1
    m += (1.0 - b1) * grad
    v *= b2
    v += (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    theta -= lr * mh / (np.sqrt(vh) + eps)
    return theta
