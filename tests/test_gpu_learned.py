"""The learned solver aid on the GPU against the CPU oracle (oracle/mlp.py).

Bars (BF16 operands, FP32 accumulation on the device; FP64 in the oracle on
the same BF16-rounded operands):
  - each layer, fed the device's own BF16 inputs: hidden activations equal
    the oracle's BF16 values to one BF16 unit in the last place (2^-7
    relative) on >= 99% of entries; the output layer to 2e-6 of max|Y|
    (FP32 accumulation over K = 256);
  - end to end: max-norm error of Y below 1e-2 of max|Y| (BF16 rounding
    of the hidden activations differs where the FP32 and FP64 sums straddle
    a rounding boundary);
  - loss within 1e-4 relative, every gradient tensor within 2e-3 relative
    (Frobenius) and cosine >= 0.99999 of the oracle's;
  - training on harvested (residual, solution) pairs lowers the loss, and
    the Galerkin-scaled guess never costs accuracy: the guessed solve
    reaches the same FP64 solution as the plain one;
  - on the channel, what the network gains is measured and asserted: the
    learned guess gains nothing (within 5 %), the learned correction inside
    flexible CG costs iterations (at least 2x).
"""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from nekoamd import device, learned
    device.init(0)
    return learned


def _net(L, rng, batch=256, nx=4, ny=4, nz=2, capacity=0):
    net = L.MLP(nx, ny, nz, capacity=capacity, batch=batch, seed=3)
    theta = (rng.uniform(-1, 1, L.NPARAM) * 0.06).astype(np.float32)
    net.set_params(theta)
    assert np.array_equal(net.get_params(), theta)
    return net, omlp.unpack(theta.astype(np.float64))


def _features(L, rng, M):
    from nekoamd import device as dev
    X = rng.standard_normal((M, omlp.NIN)).astype(np.float32)
    Xp = np.zeros((M, L.KIN), dtype=np.float32)
    Xp[:, :omlp.NIN] = X
    bits = L.to_bf16_bits(Xp)
    return dev.DeviceArray.from_host(bits.ravel()), L.from_bf16_bits(bits)[:, :omlp.NIN]


def _ulp_close(gpu_bits, ref, L):
    a = L.from_bf16_bits(gpu_bits)
    b = omlp.bf16(ref)
    bad = np.abs(a - b) > 2.0 ** -7 * np.abs(b) + 1e-30
    return bad.mean()


@pytest.mark.parametrize("M", [256, 200])
def test_forward_layers_match_oracle(L, M):
    rng = np.random.default_rng(M)
    net, p = _net(L, rng)
    Xd, Xq = _features(L, rng, M)
    Y, H1, H2 = net.forward(Xd, M, hidden=True)
    q = omlp.bf16
    Z1 = Xq @ q(p["W1"]).T + p["b1"]
    f1 = _ulp_close(H1, omlp.silu(Z1), L)
    Z2 = L.from_bf16_bits(H1) @ q(p["W2"]).T + p["b2"]
    f2 = _ulp_close(H2, omlp.silu(Z2), L)
    Yr = L.from_bf16_bits(H2) @ q(p["W3"]).T + p["b3"]
    e3 = np.abs(Y - Yr).max() / np.abs(Yr).max()
    full = omlp.forward(p, Xq)["Y"]
    ee = np.abs(Y - full).max() / np.abs(full).max()
    print(f"M={M}: H1 off-by-ulp {f1:.2%}, H2 {f2:.2%}, layer-3 err {e3:.1e}, "
          f"end-to-end {ee:.1e}")
    assert f1 < 0.01 and f2 < 0.01
    assert e3 < 2e-6
    assert ee < 1e-2


def test_gradients_match_oracle(L):
    from nekoamd import device as dev
    rng = np.random.default_rng(5)
    net, p = _net(L, rng)
    M = 128
    Xd, Xq = _features(L, rng, M)
    T = (rng.standard_normal((M, omlp.NP)) * 0.5).astype(np.float32)
    loss, g = net.loss_grad(Xd, dev.DeviceArray.from_host(T.ravel()), M)
    lo, go = omlp.loss_grad(p, Xq, T.astype(np.float64))
    gu = omlp.unpack(g.astype(np.float64))
    print(f"loss gpu {loss:.6e} oracle {lo:.6e}")
    assert abs(loss - lo) < 1e-4 * lo
    for k, _ in omlp.SIZES:
        a, b = gu[k].ravel(), go[k].ravel()
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"{k}: rel {rel:.2e} cos {cos:.6f}")
        assert rel < 2e-3 and cos > 0.99999, k


def test_training_lowers_loss_and_guess_is_safe(L):
    """Pairs (b, x) with A x = b from the device's FP64 CG on a 4^3 box of
    order 7; after training, a guessed solve reaches the same solution as the
    plain one."""
    from nekoamd import sem
    rng = np.random.default_rng(8)
    g = sem.BoxCase(4, 4, 4, 8, jitter=0.05, seed=1234)
    mask = g.mask(sem.PRS_DIRICHLET)
    m_host = oracle.face_mask(4, 4, 4, 8, sem.PRS_DIRICHLET)
    h1 = g.field(np.ones(g.n))
    d = g.field()
    g.jacobi(d, h1)
    cg = sem.CG(g)
    net = L.MLP(4, 4, 4, capacity=8, batch=64, seed=11)
    base = rng.standard_normal(g.n)

    Bh = g.B.get()

    def rhs():
        fd = g.field(Bh * (base + 0.3 * rng.standard_normal(g.n)))
        g.gs.op(fd)
        f = fd.get()
        f[m_host] = 0.0
        return f

    for _ in range(6):
        f = rhs()
        x = g.field()
        mon = cg.solve(x, g.field(f), d, mask, h1, abstol=1e-10, max_iter=2000)
        assert mon.res_final < 1e-10
        net.harvest(g.field(f), x)
    assert net.pairs == 6
    l0, l1 = net.train(epochs=40, batch=64, lr=1e-3)
    print(f"training loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.5 * l0
    f = rhs()
    x_plain = g.field()
    mp = cg.solve(x_plain, g.field(f), d, mask, h1, abstol=1e-10, max_iter=2000)
    b = g.field(f)
    gamma = net.pre(g, b, mask, h1, apply=True)
    x = g.field()
    ml = cg.solve(x, b, d, mask, h1, abstol=1e-10, max_iter=2000)
    net.post(x)
    xp, xl = x_plain.get(), x.get()
    err = np.abs(xl - xp).max() / np.abs(xp).max()
    print(f"gamma {gamma:.3e}; iterations plain {mp.iter}, with guess {ml.iter}; "
          f"start residual {mp.res_start:.3e} -> {ml.res_start:.3e}; x diff {err:.1e}")
    assert gamma > 0.0
    assert ml.res_final < 1e-10
    assert err < 1e-7


def test_bad_network_falls_back_to_plain_pcg(L):
    """A deliberately bad network (huge random weights) must not cost the
    solve: the Galerkin scaling shrinks its direction, the solve still
    reaches its tolerance and the same solution as plain Jacobi-PCG."""
    from nekoamd import sem
    rng = np.random.default_rng(21)
    g = sem.BoxCase(4, 4, 4, 8, jitter=0.05, seed=1234)
    mask = g.mask(sem.PRS_DIRICHLET)
    m_host = oracle.face_mask(4, 4, 4, 8, sem.PRS_DIRICHLET)
    h1 = g.field(np.ones(g.n))
    d = g.field()
    g.jacobi(d, h1)
    cg = sem.CG(g)
    net = L.MLP(4, 4, 4, capacity=0, batch=64, seed=2)
    net.set_params((rng.standard_normal(L.NPARAM) * 50.0).astype(np.float32))
    fd = g.field(g.B.get() * rng.standard_normal(g.n))
    g.gs.op(fd)
    f = fd.get()
    f[m_host] = 0.0
    x_plain = g.field()
    mp = cg.solve(x_plain, g.field(f), d, mask, h1, abstol=1e-10, max_iter=3000)
    b = g.field(f)
    gamma = net.pre(g, b, mask, h1, apply=True)
    x = g.field()
    ml = cg.solve(x, b, d, mask, h1, abstol=1e-10, max_iter=3000)
    net.post(x)
    err = np.abs(x.get() - x_plain.get()).max() / np.abs(x_plain.get()).max()
    print(f"bad network: gamma {gamma:.2e}, iterations plain {mp.iter} guessed {ml.iter}, "
          f"x diff {err:.1e}")
    assert np.isfinite(gamma)
    assert ml.res_final < 1e-10 and err < 1e-7
    assert ml.iter <= mp.iter + 10


def test_channel_learned_guess_against_projection_alone():
    """Two channels (8x4x4, p = 7) with the projection of the pressure and
    velocity solves; the second harvests its pressure (residual, correction)
    pairs for 10 steps, trains on them and applies the learned direction for
    8 more.  Both reach 1e-8 on every solve and agree to 1e-6.  What the
    network gains over the method without it, measured and asserted: nothing
    -- the pressure iterations of the 8 steps are the projection-only ones
    to within 5 % (measured: 1.6 % fewer; after the projection the residual
    left for the solve is not predictable element by element)."""
    from nekoamd.fortran import Channel
    c = oracle.build_case(8, 4, 4, 8, jitter=0.05, seed=1234)
    _, _, chi = oracle.seeded_design(c, 777, npass=3)
    a = Channel(8, 4, 4, 8, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000, mixed=True)
    b = Channel(8, 4, 4, 8, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000, mixed=True)
    try:
        for ch in (a, b):
            ch.set("chi", chi)
            ch.projection(20, 20, activ_step=2)
        a.step(2)
        b.step(2)
        b.learned(Channel.LEARNED_HARVEST, capacity=16, batch=256, seed=5)
        ia, _ = a.step(10)
        ib, _ = b.step(10)
        l0, l1, npairs = b.train(epochs=30, batch=256, lr=1e-3)
        print(f"trained on {npairs} pairs: loss {l0:.3e} -> {l1:.3e}")
        b.learned(Channel.LEARNED_APPLY)
        ia2, ra = a.step(8)
        ib2, rb = b.step(8)
        print("pressure iterations, projection only:", ia2[:, 0].tolist(),
              "with the learned direction:", ib2[:, 0].tolist(), "gamma", f"{b.gamma:.3e}")
        assert npairs == 10 and l1 < l0
        assert np.all(ra < 1e-8) and np.all(rb < 1e-8)
        for k in ("u", "v", "w", "p"):
            e = np.abs(b.get(k) - a.get(k)).max() / np.abs(a.get(k)).max()
            assert e < 1e-6, (k, e)
        gain = ia2[:, 0].sum() / ib2[:, 0].sum()
        print(f"learned direction: {gain:.3f}x fewer pressure iterations than projection alone")
        assert abs(gain - 1.0) < 0.05
    finally:
        a.free()
        b.free()


def test_channel_learned_correction_in_flexible_cg():
    """The learned coarse correction inside flexible CG against plain
    Jacobi-PCG on the 8x4x4 p=7 channel: the FCG channel harvests 12 steps
    of pressure pairs, trains the network on them, then runs 2 steps with the
    network's Galerkin-weighted correction every 4th FCG iteration (and 2
    with the learned guess alone).  Every solve reaches 1e-8 and the fields
    agree with the PCG channel's to 1e-6.  Measured and asserted: the
    correction costs iterations -- at least twice as many as PCG (measured
    1224/1269 against 323/311, about 4x: the truncated flexible CG loses the
    global conjugacy of CG whenever the preconditioner changes, and the
    element-local prediction of A^-1 r carries too little of the smooth
    global error to make up for it) -- and the guess alone changes them by
    less than 5 % (measured 329/331 against 329/330)."""
    from nekoamd.fortran import Channel
    c = oracle.build_case(8, 4, 4, 8, jitter=0.05, seed=1234)
    _, _, chi = oracle.seeded_design(c, 777, npass=3)
    a = Channel(8, 4, 4, 8, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
    b = Channel(8, 4, 4, 8, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
    try:
        for ch in (a, b):
            ch.set("chi", chi)
        b.pressure_solver("fcg", "jacobi")
        ia, _ = a.step(2)
        ib, _ = b.step(2)
        assert np.array_equal(ia[:, 0], ib[:, 0])      # FCG with Jacobi is PCG
        b.learned(Channel.LEARNED_HARVEST, capacity=16, batch=256, seed=5)
        a.step(12)
        b.step(12)
        l0, l1, npairs = b.train(epochs=30, batch=256, lr=1e-3)
        b.learned(Channel.LEARNED_OFF)
        b.learned_correction(4)
        ia, ra = a.step(2)
        ib, rb = b.step(2)
        print(f"trained on {npairs} pairs (loss {l0:.2e} -> {l1:.2e}); pressure iterations "
              f"PCG {ia[:, 0].tolist()}, FCG + learned correction every 4 {ib[:, 0].tolist()}")
        assert np.all(ra < 1e-8) and np.all(rb < 1e-8)
        assert ib[:, 0].sum() >= 2 * ia[:, 0].sum()
        b.learned_correction(0)
        b.learned(Channel.LEARNED_APPLY)
        ia, ra = a.step(2)
        ib, rb = b.step(2)
        print(f"learned guess alone: PCG {ia[:, 0].tolist()}, FCG + guess {ib[:, 0].tolist()}")
        assert np.all(rb < 1e-8)
        assert abs(ib[:, 0].sum() / ia[:, 0].sum() - 1.0) < 0.05
        for k in ("u", "v", "w", "p"):
            e = np.abs(b.get(k) - a.get(k)).max() / np.abs(a.get(k)).max()
            assert e < 1e-6, (k, e)
    finally:
        a.free()
        b.free()
