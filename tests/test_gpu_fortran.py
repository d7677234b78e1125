"""The Fortran host (ax_helm_factory / precon_factory / krylov_solver_factory /
set_pc / solve over the C-ABI) against the FP64 oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,lx", [((8, 4, 4), 6), ((4, 3, 2), 8)])
@pytest.mark.parametrize("kind", [0, 1])
def test_fortran_pcg_matches_oracle(dims, lx, kind):
    from nekoamd import fortran
    nx, ny, nz = dims
    o = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234)
    rng = np.random.default_rng(21 + kind)
    f = o.B * rng.uniform(-1, 1, o.n)
    faces = oracle.PRS_DIRICHLET if kind == 0 else oracle.VEL_DIRICHLET
    mask = oracle.face_mask(nx, ny, nz, lx, faces)
    fa = o.gs(f)
    fa[mask] = 0.0
    if kind == 0:
        xr, itr, r0, r1 = o.cg(fa, mask, abstol=1e-8)
    else:
        h1, h2 = np.full(o.n, 0.01), np.full(o.n, 1500.0)
        xr, itr, r0, r1 = o.cg(fa, mask, h1, h2, True, abstol=1e-8)
    x, it, s0, s1 = fortran.pcg_solve(nx, ny, nz, lx, f, kind=kind)
    print(f"fortran pcg kind={kind} lx={lx}: {it} iters (oracle {itr}), res {s1:.2e}")
    assert abs(s0 - r0) < 1e-12 * r0
    assert abs(it - itr) <= 1 and s1 < 1e-8
    assert np.abs(x - xr).max() < 1e-6 * np.abs(xr).max()
