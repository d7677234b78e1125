"""Field output in Neko's fld format (csrc/fld.cpp): one file per sample,
the series index <name>.nek5000, and a reader for those files.

Writers take device arrays (DeviceArray) or host numpy arrays; host writes
need no GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from ._lib import check, lib, ptr


def fld_write(path, fields, lx, nelv, time=0.0, step=0, mesh=None, wdsize=4):
    """Write one fld file.  fields: list of n arrays (all host numpy or all
    device arrays); mesh: (x, y, z) or None."""
    dev = not isinstance((fields or mesh)[0], np.ndarray)
    keep = [np.ascontiguousarray(a, dtype=np.float64) if not dev else a for a in fields]
    arr = (C.c_void_p * max(len(keep), 1))(*[_addr(a, dev) for a in keep])
    xyz = [None, None, None]
    if mesh is not None:
        mk = [np.ascontiguousarray(a, dtype=np.float64) if not dev else a for a in mesh]
        keep += mk
        xyz = [_addr(a, dev) for a in mk]
    check(lib().nekoamd_fld_write(
        os.fsencode(path), C.byref(C.c_double(time)), C.byref(C.c_int(step)),
        C.byref(C.c_int(lx)), C.byref(C.c_int(nelv)), C.byref(C.c_int(len(fields))), arr,
        C.c_void_p(xyz[0]), C.c_void_p(xyz[1]), C.c_void_p(xyz[2]),
        C.byref(C.c_int(wdsize)), C.byref(C.c_int(1 if dev else 0))), "fld_write")


def fld_series(path, prefix, first, count):
    check(lib().nekoamd_fld_series(os.fsencode(path), os.fsencode(prefix),
                                   C.byref(C.c_int(first)), C.byref(C.c_int(count))),
          "fld_series")


def _addr(a, dev):
    if dev:
        return ptr(a).value
    return a.ctypes.data


def fld_read(path):
    """Parse one fld file: dict with the header tokens, element numbers,
    the groups {code: array [ncomp, nelv, lx^3]} and the metadata."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = raw[:132].decode("ascii")
    tok = hdr.split()
    assert tok[0] == "#std", hdr
    ws, lx, ly, lz, nel, nelg = (int(t) for t in tok[1:7])
    time, step, fid0, nfiles = float(tok[7]), int(tok[8]), int(tok[9]), int(tok[10])
    code = tok[11]
    tag = np.frombuffer(raw, np.float32, 1, 132)[0]
    off = 136
    ids = np.frombuffer(raw, np.int32, nel, off)
    off += 4 * nel
    np_ = lx * ly * lz
    dt = np.float32 if ws == 4 else np.float64
    groups = {}
    ncomp = {"X": 3, "U": 3, "P": 1, "T": 1}
    m = re.match(r"^([XUPT]*)(?:S(\d\d))?$", code)
    assert m, code
    order = list(m.group(1)) + (["S"] if m.group(2) else [])
    ns = int(m.group(2)) if m.group(2) else 0
    for g in order:
        if g == "S":
            a = np.frombuffer(raw, dt, ns * nel * np_, off).reshape(ns, nel, np_)
        else:
            k = ncomp[g]
            a = np.frombuffer(raw, dt, nel * k * np_, off).reshape(nel, k, np_).transpose(1, 0, 2)
        off += a.size * ws
        groups[g] = a.astype(np.float64)
    nmd = sum(a.shape[0] for a in groups.values()) * nel * 2
    md_raw = np.frombuffer(raw, np.float32, nmd, off)
    off += 4 * nmd
    meta, o = {}, 0
    for g in order:
        k = groups[g].shape[0]
        if g == "S":
            meta[g] = md_raw[o:o + 2 * k * nel].reshape(k, nel, 2)
        else:
            meta[g] = md_raw[o:o + 2 * k * nel].reshape(nel, k, 2).transpose(1, 0, 2)
        o += 2 * k * nel
    return {"wdsize": ws, "lx": lx, "nelv": nel, "nelgt": nelg, "time": time, "step": step,
            "fid0": fid0, "nfiles": nfiles, "rdcode": code, "tag": tag, "ids": ids,
            "groups": groups, "meta": meta, "size": len(raw), "end": off, "header": hdr}
