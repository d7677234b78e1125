"""Checkpoint and restart of the Pn/Pn channel (the reference's chkp_t and
adjoint_pnpn_restart, sources/adjoint/adjoint_pnpn.f90:418-484) on the
config-1 mesh (8x4x4, p = 5).

* The file round trip is bit-exact: ten fields and the header written by
  nekoamd_chkp_write come back unchanged through nekoamd_chkp_read.
* A run checkpointed after 3 steps and restarted in a new case object
  continues as the uninterrupted run does.  The restart makes u, v, w, p
  and the velocity lags continuous again by col2(mult) and gs ADD, as the
  reference's restart does; for continuous fields that sum reproduces each
  value exactly at face points (u/2 + u/2) but can round by one unit in the
  last place at edges and vertices (u/4 + u/4 + u/4 + u/4 summed in member
  order), so the restarted state equals the checkpointed one to 2 ulp and
  the continued fields equal the uninterrupted run's to 1e-11 (the step
  parity bar; solves to 1e-13)."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

NX, NY, NZ, LX = 8, 4, 4, 6
STATE = ("u", "v", "w", "p", "ulag", "vlag", "wlag", "fxlag", "fylag", "fzlag")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_checkpoint_file_round_trip_is_bit_exact(tmp_path):
    from nekoamd import device
    from nekoamd._lib import check, lib
    device.init(0)
    n = 12345
    rng = np.random.default_rng(5)
    src = [device.DeviceArray.from_host(rng.standard_normal(n)) for _ in range(10)]
    dst = [device.zeros(n) for _ in range(10)]
    path = str(tmp_path / "state.chkp").encode()
    ptrs = (C.c_void_p * 10)(*[a.ptr for a in src])
    check(lib().nekoamd_chkp_write(path, C.byref(C.c_int(17)), C.byref(C.c_double(0.017)),
                                   C.byref(C.c_double(1e-3)), C.byref(C.c_int(10)), ptrs,
                                   C.byref(C.c_int64(n))), "chkp_write")
    ts, t, dt = C.c_int(0), C.c_double(0), C.c_double(0)
    ptrs = (C.c_void_p * 10)(*[a.ptr for a in dst])
    check(lib().nekoamd_chkp_read(path, C.byref(ts), C.byref(t), C.byref(dt),
                                  C.byref(C.c_int(10)), ptrs, C.byref(C.c_int64(n))),
          "chkp_read")
    assert (ts.value, t.value, dt.value) == (17, 0.017, 1e-3)
    for a, b in zip(src, dst):
        assert np.array_equal(a.get(), b.get())
    raw = open(tmp_path / "state.chkp", "rb").read()
    assert raw[:8] == b"NKAMDCHK" and len(raw) == 48 + 10 * 8 * n
    # a file of another size is refused
    rc = lib().nekoamd_chkp_read(path, C.byref(ts), C.byref(t), C.byref(dt),
                                 C.byref(C.c_int(10)), ptrs, C.byref(C.c_int64(n - 1)))
    assert rc != 0


def test_restart_continues_the_run(tmp_path):
    from nekoamd.fortran import Channel
    c = oracle.build_case(NX, NY, NZ, LX, jitter=0.05, seed=1234)
    rho, rf, chi = oracle.seeded_design(c, 777, npass=3)
    kw = dict(abstol_p=1e-13, abstol_v=1e-13, max_iter=3000)
    ref = Channel(NX, NY, NZ, LX, **kw)
    first = Channel(NX, NY, NZ, LX, **kw)
    try:
        for ch in (ref, first):
            ch.set("chi", chi)
        ref.step(3)
        first.step(3)
        path = tmp_path / "channel.chkp"
        first.checkpoint(path)
        saved = {k: first.get(k) for k in STATE}
    finally:
        first.free()
    second = Channel(NX, NY, NZ, LX, **kw)
    try:
        second.set("chi", chi)
        second.restart(path)
        assert second.tstep() == 3
        for k in STATE:
            got = second.get(k)
            if k in ("fxlag", "fylag", "fzlag"):
                assert np.array_equal(got, saved[k]), k   # read as written
            else:
                ulp = np.spacing(np.abs(saved[k]))
                assert np.all(np.abs(got - saved[k]) <= 2 * ulp), k
        ref.step(3)
        second.step(3)
        assert second.tstep() == 6
        errs = {k: rel(second.get(k), ref.get(k)) for k in ("u", "v", "w", "p")}
        print("restarted run against the uninterrupted one after 6 steps: "
              + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
        assert max(errs.values()) < 1e-11
    finally:
        second.free()
        ref.free()


def test_restart_with_time_order_3(tmp_path):
    """With time_order 3 the checkpoint also holds the six second lags
    (sixteen fields); a run restarted after 4 steps continues as the
    uninterrupted one to 1e-11 (steps 5 and 6 are BDF3/EXT3 steps that read
    the second lags)."""
    from nekoamd.fortran import Channel
    c = oracle.build_case(NX, NY, NZ, LX, jitter=0.05, seed=1234)
    rho, rf, chi = oracle.seeded_design(c, 777, npass=3)
    kw = dict(abstol_p=1e-13, abstol_v=1e-13, max_iter=3000)
    ref = Channel(NX, NY, NZ, LX, **kw)
    first = Channel(NX, NY, NZ, LX, **kw)
    path = tmp_path / "bdf3.chkp"
    try:
        for ch in (ref, first):
            ch.set("chi", chi)
            ch.time_order(3)
        ref.step(4)
        first.step(4)
        first.checkpoint(path)
        lag2 = first.get("ulag2")
    finally:
        first.free()
    raw = open(path, "rb").read()
    assert len(raw) == 48 + 16 * 8 * ref.n
    second = Channel(NX, NY, NZ, LX, **kw)
    try:
        second.set("chi", chi)
        second.time_order(3)
        second.restart(path)
        assert np.all(np.abs(second.get("ulag2") - lag2) <= 2 * np.spacing(np.abs(lag2)))
        ref.step(2)
        second.step(2)
        errs = {k: rel(second.get(k), ref.get(k)) for k in ("u", "v", "w", "p")}
        print("BDF3 restarted run against the uninterrupted one after 6 steps: "
              + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
        assert max(errs.values()) < 1e-11
    finally:
        second.free()
        ref.free()
