"""Field output in Neko's fld format (csrc/fld.cpp), host writes (no GPU).

Pins:
  - the series index equals, byte for byte, the reference's own index file
    examples/power_iterations/baseflow0.nek5000 (copied as the fixture
    tests/golden/fld/baseflow0.nek5000);
  - the 132-byte header follows the fixed-width Fortran record
    '#std',1x,i1,1x,i2,1x,i2,1x,i2,1x,i10,1x,i10,1x,e20.13,1x,i9,1x,i6,1x,i6,1x,a
    column by column (checked against Python's own formatting of each field
    at its columns), with the Nek5000 byte-order tag 6.54321 after it;
  - the list-size -> group mapping of Neko's writer (1 P, 2 P T, 3 U, 4 P U,
    5 P U T, n > 5 P U T S(n-5)) matches the reference's own description of
    its 13-field output (steady_state_problem.f90:216-226: design ->
    temperature, the rest -> s1..s8);
  - data blocks and the per-element min/max metadata against numpy on the
    written arrays (float32 rounding for wdsize 4, exact for wdsize 8).
"""
import os

import numpy as np
import pytest

from nekoamd import output

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fld", "baseflow0.nek5000")


def test_series_index_matches_reference_file(tmp_path):
    p = tmp_path / "baseflow0.nek5000"
    output.fld_series(str(p), "baseflow", 0, 1)
    assert p.read_bytes() == open(GOLD, "rb").read()


def _fortran_e(v):
    """Ew.d, w = 20, d = 13, via the decimal exponent of the value."""
    if v == 0:
        return "0.0000000000000E+00".rjust(20)
    s = f"{abs(v):.12e}"                 # d.dddddddddddde+xx
    mant, ex = s.split("e")
    digits = mant.replace(".", "")      # 13 significant digits
    e = int(ex) + 1
    return (("-" if v < 0 else "") + "0." + digits + f"E{'-' if e < 0 else '+'}{abs(e):02d}").rjust(20)


@pytest.mark.parametrize("t", [0.0, 0.1, 1.0, 12.5, 3e-4, 123456.789])
def test_header_columns(tmp_path, t):
    lx, nelv = 4, 3
    rng = np.random.default_rng(1)
    f = [rng.standard_normal(nelv * lx ** 3) for _ in range(4)]
    p = tmp_path / "a0.f00000"
    output.fld_write(str(p), f, lx, nelv, time=t, step=7)
    h = open(p, "rb").read(132).decode("ascii")
    assert len(h) == 132
    cols = {"tag": h[0:4], "ws": h[5:6], "nx": h[7:9], "ny": h[10:12], "nz": h[13:15],
            "nel": h[16:26], "nelg": h[27:37], "time": h[38:58], "step": h[59:68],
            "fid0": h[69:75], "nfiles": h[76:82], "code": h[83:]}
    assert cols["tag"] == "#std" and cols["ws"] == "4"
    assert cols["nx"] == cols["ny"] == cols["nz"] == f"{lx:2d}"
    assert cols["nel"] == cols["nelg"] == f"{nelv:10d}"
    assert cols["time"] == _fortran_e(t), (cols["time"], _fortran_e(t))
    assert float(cols["time"]) == pytest.approx(t, rel=1e-12, abs=0)
    assert cols["step"] == f"{7:9d}" and cols["fid0"] == f"{0:6d}" and cols["nfiles"] == f"{1:6d}"
    assert cols["code"].rstrip() == "PU"
    assert all(h[i] == " " for i in (4, 6, 9, 12, 15, 26, 37, 58, 68, 75, 82))


@pytest.mark.parametrize("nf,code", [(1, "P"), (2, "PT"), (3, "U"), (4, "PU"), (5, "PUT"),
                                     (13, "PUTS08")])
def test_group_mapping(tmp_path, nf, code):
    lx, nelv = 3, 2
    rng = np.random.default_rng(nf)
    f = [rng.standard_normal(nelv * lx ** 3) for _ in range(nf)]
    mesh = [rng.standard_normal(nelv * lx ** 3) for _ in range(3)]
    p = tmp_path / "b0.f00000"
    output.fld_write(str(p), f, lx, nelv, mesh=mesh, wdsize=8)
    r = output.fld_read(str(p))
    assert r["rdcode"] == "X" + code
    assert r["end"] == r["size"]
    g = r["groups"]
    blk = lambda a: a.reshape(nelv, lx ** 3)
    assert all(np.array_equal(g["X"][d], blk(mesh[d])) for d in range(3))
    want = {1: {"P": [0]}, 2: {"P": [0], "T": [1]}, 3: {"U": [0, 1, 2]},
            4: {"P": [0], "U": [1, 2, 3]}, 5: {"P": [0], "U": [1, 2, 3], "T": [4]},
            13: {"P": [0], "U": [1, 2, 3], "T": [4], "S": list(range(5, 13))}}[nf]
    for grp, idx in want.items():
        for k, i in enumerate(idx):
            assert np.array_equal(g[grp][k], blk(f[i])), (grp, k)


def test_single_precision_data_ids_tag_and_metadata(tmp_path):
    lx, nelv = 5, 4
    rng = np.random.default_rng(3)
    f = [rng.standard_normal(nelv * lx ** 3) * 10 for _ in range(7)]
    mesh = [rng.uniform(0, 2, nelv * lx ** 3) for _ in range(3)]
    p = tmp_path / "c0.f00000"
    output.fld_write(str(p), f, lx, nelv, time=0.25, step=2, mesh=mesh)
    r = output.fld_read(str(p))
    assert r["wdsize"] == 4 and r["tag"] == np.float32(6.54321)
    assert np.array_equal(r["ids"], np.arange(1, nelv + 1))
    expect_size = 132 + 4 + 4 * nelv + 4 * (3 + 1 + 3 + 1 + 2) * nelv * lx ** 3 \
        + 4 * 2 * nelv * (3 + 1 + 3 + 1 + 2)
    assert r["size"] == expect_size == r["end"]
    g, md = r["groups"], r["meta"]
    f32 = lambda a: a.astype(np.float32).astype(np.float64).reshape(nelv, lx ** 3)
    assert np.array_equal(g["P"][0], f32(f[0]))
    assert np.array_equal(g["S"][1], f32(f[6]))
    for grp in g:
        for k in range(g[grp].shape[0]):
            src = {"X": mesh, "U": f[1:4], "P": f[:1], "T": f[4:5], "S": f[5:]}[grp][k]
            b = src.reshape(nelv, lx ** 3)
            assert np.array_equal(md[grp][k][:, 0], b.min(1).astype(np.float32))
            assert np.array_equal(md[grp][k][:, 1], b.max(1).astype(np.float32))


def test_rejects_bad_arguments(tmp_path):
    from nekoamd._lib import NekoAmdError
    f = [np.zeros(8)]
    with pytest.raises(NekoAmdError):
        output.fld_write(str(tmp_path / "d"), f, 2, 1, wdsize=2)
    with pytest.raises(NekoAmdError):
        output.fld_write(str(tmp_path / "nodir" / "e"), f, 2, 1)
