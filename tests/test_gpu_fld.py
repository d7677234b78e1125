"""fld output of the Fortran host on the GPU (Neko's fld_file_output_t):
the channel's p, u, v, w series and the optimisation loop's 13-field series
(steady_state_problem.f90:216-244), read back with nekoamd.output.fld_read.

Bars: the mesh block equals the case coordinates and the field blocks equal
the device fields, both to float32 rounding (wdsize 4) or exactly
(wdsize 8); the mesh is in the first file of a series only; the index
counts the samples."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _blk(a, nelv, lx):
    return a.reshape(nelv, lx ** 3)


def test_channel_series(tmp_path):
    from nekoamd.fortran import Channel
    from nekoamd.output import fld_read
    nx, ny, nz, lx = 4, 2, 2, 6
    c = oracle.build_case(nx, ny, nz, lx, jitter=0.05, seed=1234)
    ch = Channel(nx, ny, nz, lx)
    name = str(tmp_path / "chan")
    try:
        ch.step(2)
        ch.sample(name, 0.002, wdsize=8)
        r0 = fld_read(name + "0.f00000")
        assert r0["rdcode"] == "XPU" and r0["wdsize"] == 8 and r0["nelv"] == c.nelv
        for d, a in enumerate((c.x, c.y, c.zc)):
            assert np.abs(r0["groups"]["X"][d] - _blk(a, c.nelv, lx)).max() < 1e-13
        assert np.array_equal(r0["groups"]["P"][0], _blk(ch.get("p"), c.nelv, lx))
        for d, k in enumerate("uvw"):
            assert np.array_equal(r0["groups"]["U"][d], _blk(ch.get(k), c.nelv, lx))
        ch.step(1)
        ch.sample(name, 0.003, wdsize=8)
        r1 = fld_read(name + "0.f00001")
        assert r1["rdcode"] == "PU" and r1["step"] == 1
        assert r1["time"] == pytest.approx(0.003, rel=1e-12)
        assert np.array_equal(r1["groups"]["U"][0], _blk(ch.get("u"), c.nelv, lx))
        idx = open(name + ".nek5000").read()
        assert idx == f"filetemplate: {name}%01d.f%05d\nfirsttimestep: 0\nnumtimesteps: 2\n"
    finally:
        ch.free()


def test_optimisation_series(tmp_path):
    from nekoamd.fortran import TopOpt
    from nekoamd.output import fld_read
    nx, ny, nz, lx = 4, 2, 2, 6
    t = TopOpt(nx, ny, nz, lx, abstol_p=1e-8, abstol_v=1e-8, max_iter=2000)
    name = str(tmp_path / "optimization")
    try:
        t.output(name)
        x0 = t.get("x")
        t.iterate(2, 2)
        r = fld_read(name + "0.f00000")
        ne = nx * ny * nz
        assert r["rdcode"] == "XPUTS08" and r["wdsize"] == 4
        f32 = lambda a: a.astype(np.float32).astype(np.float64).reshape(ne, lx ** 3)
        g = r["groups"]
        # sampled after the sensitivities, before the MMA update
        assert np.array_equal(g["T"][0], f32(x0))
        assert np.array_equal(g["U"][0], f32(t.get("u")))
        assert np.array_equal(g["S"][0], f32(t.get("ua")))
        assert np.array_equal(g["S"][5], f32(t.get("dF_dchi")))
        assert np.array_equal(g["S"][7], f32(t.get("dF_drho")))
        assert r["time"] == 1.0
    finally:
        t.free()
