"""The oracle's MLP (oracle/mlp.py) pinned independently of its own
formulas: BF16 rounding against hand-worked ties and torch's bfloat16 cast,
SiLU' against finite differences, the backward against central finite
differences of the forward, the features against a direct loop, and Adam's
first step against its closed form."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_rule():
    one = 1.0
    assert mlp.bf16(np.float32(one)) == 1.0
    assert mlp.bf16(np.float32(1 + 2 ** -8)) == 1.0                 # tie -> even
    assert mlp.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6     # tie -> even
    assert mlp.bf16(np.float32(1 + 2 ** -8 + 2 ** -12)) == 1 + 2 ** -7
    assert mlp.bf16(np.float32(-3.0)) == -3.0
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h)
    assert np.abs(fd - mlp.dsilu(z)).max() < 1e-8
    assert mlp.silu(0.0) == 0.0 and abs(mlp.dsilu(0.0) - 0.5) < 1e-15


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = rng.uniform(-1, 1, mlp.NPARAM) * 0.05
    p = mlp.unpack(theta)
    X = rng.standard_normal((4, mlp.NIN))
    T = rng.standard_normal((4, mlp.NP))
    _, g = mlp.loss_grad(p, X, T, round_bf16=False)
    gv = mlp.pack(g)

    def loss(th):
        return mlp.loss_grad(mlp.unpack(th), X, T, round_bf16=False)[0]

    o = 0
    for k, s in mlp.SIZES:
        m = int(np.prod(s))
        for idx in rng.choice(m, 12, replace=False):
            i = o + idx
            e = np.zeros_like(theta)
            e[i] = 1e-5
            fd = (loss(theta + e) - loss(theta - e)) / 2e-5
            assert abs(fd - gv[i]) < 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, idx, fd, gv[i])
        o += m


def test_features_direct_loop():
    nx, ny, nz = 3, 2, 2
    nb = mlp.neighbours(nx, ny, nz)
    r = np.random.default_rng(2).standard_normal(nx * ny * nz * 512)
    X, s = mlp.features(r, nb)
    assert abs(s - np.sqrt(np.mean(r * r))) < 1e-15
    for e in range(nx * ny * nz):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        assert np.allclose(X[e, :512], r[e * 512:(e + 1) * 512] / s, rtol=0, atol=1e-15)
        for k, (dx, dy, dz) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
                                          (0, 0, -1), (0, 0, 1)]):
            jx, jy, jz = ex + dx, ey + dy, ez + dz
            if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz:
                j = jx + nx * (jy + ny * jz)
                ref = r[j * 512:(j + 1) * 512].mean() / s
            else:
                ref = 0.0
            assert abs(X[e, 512 + k] - ref) < 1e-14


def test_adam_first_step():
    th = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 0.0])
    m, v = np.zeros(3), np.zeros(3)
    mlp.adam(th, g, m, v, 1, lr=0.01)
    # step 1: m^ = g, v^ = g^2 -> theta -= lr g / (|g| + eps)
    ref = np.array([1.0 - 0.01 * 0.3 / (0.3 + 1e-8), -2.0 + 0.01 * 4.0 / (4.0 + 1e-8), 0.5])
    assert np.allclose(th, ref, rtol=0, atol=1e-15)
