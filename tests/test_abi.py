"""CPU-side checks of the product boundary: the HIP library and the Fortran
host library load, and export every entry point include/*.h declares.  No
compute call is made (there is no GPU here)."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "neko-top_amd")


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*|uint64_t)\s+(\w+)\s*\(", txt, re.M)


LIB_OF_HEADER = {"nekoamd.h": "libnekoamd.so", "nekoamd_fortran.h": "libnekoamd_f.so"}


def test_every_declared_symbol_is_exported():
    headers = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
    assert {os.path.basename(h) for h in headers} <= set(LIB_OF_HEADER)
    for h in headers:
        lib = C.CDLL(os.path.join(PKG, LIB_OF_HEADER[os.path.basename(h)]))
        names = declared(h)
        assert names, h
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, f"{os.path.basename(h)}: not exported: {missing}"


def test_host_setup_entry_points_match_oracle():
    """The host-side set-up functions need no GPU: GLL, vertices, gids,
    face masks and the gs map must equal the oracle's exactly."""
    import numpy as np

    import oracle
    from nekoamd import sem

    for lx in (2, 5, 8):
        z, w, D = sem.gll(lx)
        zo, wo, Do = oracle.gll(lx)
        np.testing.assert_allclose(z, zo, rtol=0, atol=1e-15)
        np.testing.assert_allclose(w, wo, rtol=2e-15)
        np.testing.assert_allclose(D, Do, rtol=1e-14, atol=1e-13)
    v = sem.box_vertices(8, 4, 4, 0.05, 77)
    vo = oracle.box_vertices(8, 4, 4, jitter=0.05, seed=77)
    np.testing.assert_allclose(v, vo, rtol=1e-15, atol=1e-16)
    g = sem.box_gids(8, 4, 4, 6)
    c = oracle.build_case(8, 4, 4, 6, jitter=0.05, seed=77)
    assert np.array_equal(g, c.gid)
    for faces in (sem.VEL_DIRICHLET, sem.PRS_DIRICHLET, 0, 63):
        assert np.array_equal(sem.face_mask(8, 4, 4, 6, faces),
                              oracle.face_mask(8, 4, 4, 6, faces))


def test_fortran_host_links_the_hip_library():
    """libnekoamd_f.so resolves libnekoamd.so next to it (rpath $ORIGIN)."""
    import subprocess
    out = subprocess.run(["ldd", os.path.join(PKG, "libnekoamd_f.so")], capture_output=True,
                         text=True, check=True).stdout
    line = [l for l in out.splitlines() if "libnekoamd.so" in l]
    assert line and os.path.join("neko-top_amd", "libnekoamd.so") in line[0].replace("/./", "/")
