"""GPU parity of the topology-optimisation kernels and the device MMA with
the FP64 oracle (through the C-ABI)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from nekoamd import device, sem, topopt
    device.init(0)
    g = sem.BoxCase(4, 3, 2, 8, jitter=0.05, seed=77)
    o = oracle.build_case(4, 3, 2, 8, jitter=0.05, seed=77)
    return device, sem, topopt, g, o


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_elementwise_terms(env):
    device, sem, T, g, o = env
    n = g.n
    rng = np.random.default_rng(1)
    u, v, w, ua, va, wa = (rng.standard_normal(n) for _ in range(6))
    rho = rng.uniform(0, 1, n)
    chi = oracle.ramp_fwd(rho, 1e4, 0.0, 0.1, True)
    f = [rng.standard_normal(n) for _ in range(3)]
    D = {k: g.field(val) for k, val in dict(u=u, v=v, w=w, ua=ua, va=va, wa=wa, rho=rho,
                                            chi=chi).items()}
    F = [g.field(x) for x in f]
    T.brinkman(*F, D["u"], D["v"], D["w"], D["chi"])
    ref = oracle.brinkman(*f, u, v, w, chi)
    for a, b in zip(F, ref):
        assert rel(a.get(), b) < 1e-15
    F = [g.field(x) for x in f]
    T.adjoint_lube(*F, D["u"], D["v"], D["w"], D["chi"], K=2.0)
    ref = oracle.adjoint_lube(*f, u, v, w, chi, K=2.0)
    for a, b in zip(F, ref):
        assert rel(a.get(), b) < 1e-15
    s = g.field()
    T.dissipation_sensitivity(s, D["u"], D["v"], D["w"], D["ua"], D["va"], D["wa"], 1.0, 1.0)
    assert rel(s.get(), oracle.dissipation_sensitivity(u, v, w, ua, va, wa)) < 1e-15
    r = g.field()
    T.update_permeability(r, D["rho"], 1000.0, 0.0, 1.0)
    assert rel(r.get(), oracle.update_permeability(rho, 1000.0, 0.0, 1.0)) < 1e-15
    for up in (True, False):
        out = g.field()
        T.ramp_fwd(out, D["rho"], 1e4, 0.0, 0.1, up)
        assert rel(out.get(), oracle.ramp_fwd(rho, 1e4, 0.0, 0.1, up)) < 1e-15
        T.ramp_bwd(out, D["u"], D["rho"], 1e4, 0.0, 0.1, up)
        assert rel(out.get(), oracle.ramp_bwd(u, rho, 1e4, 0.0, 0.1, up)) < 1e-15
    out = g.field()
    T.linear_fwd(out, D["rho"], 2.0, 5.0)
    assert rel(out.get(), oracle.linear_fwd(rho, 2.0, 5.0)) < 1e-15
    T.linear_bwd(out, D["u"], 2.0, 5.0)
    assert rel(out.get(), oracle.linear_bwd(u, 2.0, 5.0)) < 1e-15


def test_masked_col3_is_the_legacy_brinkman_force(env):
    # design.f90:266-294: f = resistance * u on the design mask (1-based)
    device, sem, T, g, o = env
    import ctypes as C
    from nekoamd._lib import check, lib, ptr, ri
    n = g.n
    rng = np.random.default_rng(2)
    res, u, f = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    mask = (np.nonzero(rng.uniform(size=n) < 0.3)[0] + 1).astype(np.intc)
    fd, ud, rd_ = g.field(f), g.field(u), g.field(res)
    md = device.DeviceArray.from_host(mask)
    check(lib().hip_col3_mask(ptr(fd), ptr(ud), ptr(rd_), ri(n), ptr(md), ri(mask.size)), "c3m")
    ref = f.copy()
    oracle.lib().or_col3_mask(ref, u, res, mask.ctypes.data_as(C.c_void_p), mask.size)
    assert np.array_equal(fd.get(), ref)


def test_gradient_and_objective(env):
    device, sem, T, g, o = env
    n = g.n
    rng = np.random.default_rng(3)
    x, y, z = o.x, o.y, o.zc
    u = np.sin(2 * x) * y + 0.1 * rng.standard_normal(n)
    v = x * z * z
    w = np.cos(y + z)
    chi = rng.uniform(0, 10, n)
    gx, gy, gz = T.grad(g, g.field(u))
    rx, ry, rz = oracle.grad(o, u)
    for a, b in ((gx, rx), (gy, ry), (gz, rz)):
        assert rel(a.get(), b) < 1e-12
    out = T.dissipation_objective(g, g.field(u), g.field(v), g.field(w), g.field(chi), K=1.0)
    ref = oracle.dissipation_objective(o, u, v, w, chi, K=1.0)
    print("objective gpu", out, "oracle", ref)
    np.testing.assert_allclose(out, ref, rtol=1e-12)
    vc = T.volume_constraint(g.field(np.clip(u, 0, 1)), g.B, g.volume, 0.4)
    vr = oracle.volume_constraint(np.clip(u, 0, 1), o.B, o.volume, 0.4)
    np.testing.assert_allclose(vc, vr, rtol=1e-12)
    s = g.field()
    T.volume_sensitivity(s, o.volume)
    assert np.all(s.get() == oracle.volume_sensitivity(n, o.volume))


def _func1(x):
    c = float(np.float32(0.0624))
    k = np.array([61.0, 37.0, 19.0, 7.0, 1.0])
    return np.full(5, c), np.array([np.sum(k / x ** 3) - 1]), (-3 * k / x ** 4).reshape(1, 5)


def _func2(x):
    p1, p2 = np.array([5.0, 2.0, 1.0]), np.array([3.0, 4.0, 3.0])
    fval = np.array([np.sum((x - p1) ** 2) - 9.0, np.sum((x - p2) ** 2) - 9.0])
    return 2 * x, fval, np.vstack([2.0 * (x - p1), 2.0 * (x - p2)])


@pytest.mark.parametrize("prob", [1, 2])
def test_mma_known_answers_on_device(env, prob):
    device, sem, T, g, o = env
    if prob == 1:
        x0, m, a0, a, c, d, lo, hi, func = [5.0] * 5, 1, 1.0, [0.0], [1000.0], [0.0], 0.0, 10.0, _func1
        want = [6.0160159, 5.30917391, 4.49432957, 3.50147495, 2.1526653]
    else:
        x0, m, a0, a, c, d, lo, hi, func = ([4.0, 3.0, 2.0], 2, 1.0, [0.0, 0.0], [1000.0, 1000.0],
                                           [1.0, 1.0], 0.0, 5.0, _func2)
        want = [2.01751859, 1.78001144, 1.23750715]
    x = np.array(x0)
    n = x.size
    xd = device.DeviceArray.from_host(x)
    opt = T.MMA(xd, m, a0, a, c, d, device.DeviceArray.from_host(np.full(n, lo)),
                device.DeviceArray.from_host(np.full(n, hi)))
    ref = oracle.MMA(x, m, a0, a, c, d, np.full(n, lo), np.full(n, hi))
    xr = x.copy()
    df0, fval, dfdx = func(x)
    for it in range(1, 41):
        opt.update(it, xd, device.DeviceArray.from_host(df0), fval,
                   device.DeviceArray.from_host(np.asfortranarray(dfdx).ravel(order="F")))
        xr = ref.update(it, xr, df0, fval, dfdx)
        x = xd.get()
        assert np.abs(x - xr).max() < 1e-9 * np.abs(xr).max(), (it, x, xr)
        df0, fval, dfdx = func(x)
        kd = opt.kkt(xd, device.DeviceArray.from_host(df0), fval,
                     device.DeviceArray.from_host(np.asfortranarray(dfdx).ravel(order="F")))
        kr = ref.kkt(xr, df0, fval, dfdx)
        assert abs(kd[1] - kr[1]) <= 1e-6 * max(kr[1], 1e-12) + 1e-12
    np.testing.assert_allclose(x, want, atol=2e-7)


@pytest.mark.parametrize("n", [20000, 4194304])
def test_mma_large_n_parity(env, n):
    """n design variables (20,000, and 4,194,304 = the 32x16x16 p=7 loop's
    per-GLL design), m = 1 volume constraint: three updates of a synthetic
    smooth objective against the oracle."""
    device, sem, T, g, o = env
    t = np.linspace(0, 1, n)
    x = 0.4 + 0.1 * np.sin(7 * t)
    target = 0.5 + 0.4 * np.cos(5 * t)
    xd = device.DeviceArray.from_host(x)
    opt = T.MMA(xd, 1, 0.0, [0.0], [100.0], [0.0], device.DeviceArray.from_host(np.zeros(n)),
                device.DeviceArray.from_host(np.ones(n)))
    ref = oracle.MMA(x, 1, 0.0, [0.0], [100.0], [0.0], np.zeros(n), np.ones(n))
    xr = x.copy()
    for it in range(1, 4):
        df0 = 2 * (x - target) / n
        fval = np.array([100 * (np.mean(x) - 0.4)])
        dfdx = np.full(n, 100.0 / n)
        opt.update(it, xd, device.DeviceArray.from_host(df0), fval,
                   device.DeviceArray.from_host(dfdx))
        xr = ref.update(it, xr, df0, fval, dfdx.reshape(1, n))
        x = xd.get()
        err = np.abs(x - xr).max()
        print(f"mma n={n} it={it}: max |x_gpu - x_oracle| = {err:.2e}, newton iters "
              f"{opt.residuals()[2]:.0f}")
        assert err < 1e-9
