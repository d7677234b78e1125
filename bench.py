#!/usr/bin/env python3
"""Benchmark of the spectral-element operator core on one MI355X.

Workload (32^3 elements, p = 7, 16,777,216 GLL points; Brinkman channel box
with +-5% seeded vertex jitter): one step is the solve phase of a Pn/Pn time
step -- Jacobi update and Jacobi-PCG pressure solve to 1e-8 (Dirichlet at the
outflow), then Jacobi update and three Jacobi-PCG velocity Helmholtz solves
to 1e-8 (h1 = 1/Re, h2 = 1.5/dt; Dirichlet at inflow and walls).  Right-hand
sides are fixed seeded smooth-plus-noise fields, assembled (gs) and masked,
resident in HBM before timing starts.

Prints one JSON line: value = seconds per step (lower is better), the
ax_helm roofline (HIP events on the library stream), gs_op throughput, PCG
iteration counts and the CPU baseline (the FP64 oracle on the host cores,
timed on a bounded sample and scaled to one step).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neko-top_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "s/timestep + ax_helm GB/s at 32^3 el p=7; CG iters, learned vs plain @1e-8"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def smooth_rhs(case, seed):
    """B * (sin(pi x / Lx) sin(pi y) sin(pi z) + 0.1 * U(-1,1)), assembled."""
    from nekoamd._lib import check, lib, ptr, ri64
    x, y, z = case.x.get(), case.y.get(), case.zc.get()
    Lx = case.nx / case.ny
    rng = np.random.default_rng(seed)
    xi = np.sin(np.pi * x / Lx) * np.sin(np.pi * y) * np.sin(np.pi * z)
    xi += 0.1 * rng.uniform(-1, 1, case.n)
    f = case.field(xi)
    check(lib().nekoamd_col2(ptr(f), ptr(case.B), ri64(case.n)), "col2")
    case.gs.op(f)
    return f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nel", type=int, default=32, help="elements per direction")
    ap.add_argument("--lx", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from nekoamd import device, sem
    from nekoamd._lib import check, lib, ptr, rd, ri, ri64

    device.init(0)
    dname = device.name()
    log("device:", dname)
    N, lx = args.nel, args.lx
    t0 = time.time()
    case = sem.BoxCase(N, N, N, lx, jitter=0.05, seed=1234)
    n = case.n
    log(f"case built: n={n} in {time.time() - t0:.1f}s")

    Re, dt = 100.0, 1e-3
    h1p = case.field(np.ones(n))
    h1v = case.field(np.full(n, 1.0 / Re))
    h2v = case.field(np.full(n, 1.5 / dt))
    mp = case.mask(sem.PRS_DIRICHLET)
    mv = case.mask(sem.VEL_DIRICHLET)
    fp = smooth_rhs(case, 11)
    check(lib().nekoamd_bc_zero(ptr(fp), ptr(mp), ri64(mp.count)), "bc")
    fv = []
    for s in (12, 13, 14):
        f = smooth_rhs(case, s)
        check(lib().nekoamd_bc_zero(ptr(f), ptr(mv), ri64(mv.count)), "bc")
        fv.append(f)
    dp, dv = case.field(), case.field()
    xp = case.field()
    xv = [case.field() for _ in range(3)]
    cg = sem.CG(case)
    iters = {}

    def step():
        case.jacobi(dp, h1p)
        m = cg.solve(xp, fp, dp, mp, h1p, abstol=1e-8, max_iter=2000, poll=16)
        iters["pressure"] = m.iter
        iters["pressure_res"] = m.res_final
        case.jacobi(dv, h1v, h2v, True)
        vi = []
        for c in range(3):
            m = cg.solve(xv[c], fv[c], dv, mv, h1v, h2v, True, abstol=1e-8, max_iter=2000,
                         poll=4)
            vi.append(m.iter)
        iters["velocity"] = vi

    for _ in range(args.warmup):
        step()
    device.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    device.sync()
    ms_step = (time.perf_counter() - t) * 1e3 / max(args.steps, 1)
    log(f"step: {ms_step:.2f} ms, iters {iters}")

    # ---- ax_helm roofline: HIP events on the library stream ---------------
    u = case.field(np.random.default_rng(5).uniform(-1, 1, n))
    w = case.field()
    e0, e1 = device.Event(), device.Event()

    def timed(fn, reps=20):
        fn()
        device.sync()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.sync()
        return e0.elapsed_ms(e1) / reps

    ax_ms = timed(lambda: case.ax(w, u, h1p))
    ax_bytes = 72 * n  # u, G11..G33, h1 read + w written, FP64
    ax_flops = (12 * lx ** 4 + 18 * lx ** 3) * case.nelv
    ax_gbs = ax_bytes / ax_ms / 1e6
    axv_ms = timed(lambda: case.ax(w, u, h1v, h2v, True))
    axv_gbs = 88 * n / axv_ms / 1e6
    u32 = case.field(np.random.default_rng(5).uniform(-1, 1, n), np.float32)
    w32 = case.field(None, np.float32)
    h132 = case.field(np.ones(n), np.float32)
    ax32_ms = timed(lambda: case.ax(w32, u32, h132, variant="f32"))
    ax32_gbs = 60 * n / ax32_ms / 1e6
    del u32, w32, h132
    idx, bm, bc = case.gs.maps()
    nocc = int((bm * bc).sum())
    gs_ms = timed(lambda: case.gs.op(w))
    gs_gbs = nocc * 20 / gs_ms / 1e6
    cp_ms = timed(lambda: check(lib().nekoamd_copy(ptr(w), ptr(u), ri64(n)), "copy"))
    copy_gbs = 16 * n / cp_ms / 1e6
    log(f"ax_helm f64: {ax_ms*1e3:.1f} us, {ax_gbs:.0f} GB/s; velocity form {axv_gbs:.0f} GB/s; "
        f"f32 {ax32_gbs:.0f} GB/s; gs {gs_ms*1e3:.1f} us {gs_gbs:.0f} GB/s; copy {copy_gbs:.0f} GB/s")

    # ---- CPU baseline: the FP64 oracle on the host cores ------------------
    cpu = None
    if not args.no_cpu:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        tb = time.time()
        oc = oracle.build_case(N, N, N, lx, jitter=0.05, seed=1234)
        log(f"oracle case built in {time.time() - tb:.1f}s")
        mask = oracle.face_mask(N, N, N, lx, oracle.PRS_DIRICHLET)
        f = fp.get()
        h1 = np.ones(n)
        tc = time.perf_counter()
        dinv = oc.jacobi(h1)
        t_jac = time.perf_counter() - tc
        k = args.cpu_iters
        tc = time.perf_counter()
        oc.cg(f, mask, h1, abstol=0.0, max_iter=k, dinv=dinv)
        t_it = (time.perf_counter() - tc) / k
        tot_iters = iters["pressure"] + sum(iters["velocity"])
        cpu_step = 2 * t_jac + tot_iters * t_it
        cpu = {"value": cpu_step, "unit": "s/timestep", "cores": cores, "kind": "port",
               "sample": f"FP64 oracle (oracle/sem.c, gcc -O3, OpenMP) at 32^3 p=7: one Jacobi "
                         f"update ({t_jac:.2f} s) and {k} pressure-PCG iterations "
                         f"({t_it:.3f} s/iter) timed, scaled to this step's {tot_iters} PCG "
                         f"iterations and 2 Jacobi updates"}
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC,
        "value": ms_step / 1e3,
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"solve phase of one Pn/Pn step (Jacobi update + PCG pressure "
                               f"to 1e-8, Jacobi update + 3 PCG velocity Helmholtz to 1e-8) on a "
                               f"{N}^3-element p={lx - 1} jittered box, {n} GLL points",
                   "elements": case.nelv, "lx": lx, "gll_points": n, "Re": Re, "dt": dt,
                   "device": dname},
        "roofline": {"bound": "hbm", "achieved": round(ax_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ax_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "ax_helm_k<8,f64> pressure form, 72 B/point algorithmic",
                     "avg_us": round(ax_ms * 1e3, 2)},
        "cpu_baseline": cpu,
        "cg_iters": {"pressure": iters["pressure"], "velocity": iters["velocity"],
                     "learned": None},
        "ax_helm": {"f64_GBps": round(ax_gbs, 1), "f64_TFLOPs": round(ax_flops / ax_ms / 1e9, 2),
                    "f64_velocity_GBps": round(axv_gbs, 1), "f32_GBps": round(ax32_gbs, 1)},
        "gs_op_GBps": round(gs_gbs, 1),
        "copy_GBps": round(copy_gbs, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
