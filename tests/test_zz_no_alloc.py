"""The allocation-free objective (nekoamd_dissipation_objective_noalloc)
against the allocating entry point (nekoamd_dissipation_objective) and the
oracle, through the C-ABI: two calls in a row, the second on a smaller
mesh, then the size check.  The two entry points run the same kernels with
the same reduction order, so their results are bit-identical."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _fields(g, o, seed):
    rng = np.random.default_rng(seed)
    x, y, z = o.x, o.y, o.zc
    u = np.sin(2 * x) * y + 0.1 * rng.standard_normal(g.n)
    v = x * z * z + 0.05 * rng.standard_normal(g.n)
    w = np.cos(y + z)
    chi = rng.uniform(0, 1e4, g.n)
    return u, v, w, chi


def _check(T, g, o, seed, K, obj_scale, if_lube):
    u, v, w, chi = _fields(g, o, seed)
    D = [g.field(a) for a in (u, v, w, chi)]
    ref = T.dissipation_objective(g, *D, K=K, obj_scale=obj_scale, if_lube=if_lube)
    got = T.dissipation_objective_noalloc(g, *D, K=K, obj_scale=obj_scale, if_lube=if_lube)
    want = oracle.dissipation_objective(o, u, v, w, chi, K=K, obj_scale=obj_scale,
                                        if_lube=if_lube)
    print(f"nelv {g.nelv}: noalloc {got}, allocating {ref}, oracle {want}", flush=True)
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_noalloc_objective_matches_allocating_entry_point():
    from nekoamd import device, sem, topopt as T
    device.init(0)
    big = (8, 4, 4, 8)
    small = (4, 3, 2, 6)
    for (nx, ny, nz, lx), seed, K, s, lube in ((big, 5, 1.0, 1.0, True),
                                               (small, 6, 2.5, 0.5, True),
                                               (small, 7, 1.0, 1.0, False)):
        g = sem.BoxCase(nx, ny, nz, lx, jitter=0.05, seed=31)
        o = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=31)
        _check(T, g, o, seed, K, s, lube)


def test_noalloc_objective_rejects_sizes_beyond_its_scratch():
    """nelv = 262145 is refused by the host-side check before any device
    work (the arrays passed are those of a 24-element mesh)."""
    from nekoamd import device, sem, topopt as T
    from nekoamd._lib import NekoAmdError
    device.init(0)
    g = sem.BoxCase(4, 3, 2, 6, jitter=0.05, seed=31)
    f = [g.field(np.ones(g.n)) for _ in range(4)]
    with pytest.raises(NekoAmdError, match="262144"):
        T.dissipation_objective_noalloc(g, *f, nelv=262145)
