#!/usr/bin/env python3
"""Benchmark of the spectral-element operator core on one MI355X.

Workload: the Brinkman-penalised channel on a 32^3-element box of order 7
(16,777,216 GLL points, +-5% seeded vertex jitter), Re = 100, dt = 1e-3,
BDF2/EXT2.  The design is seeded: rho ~ U[0,1) per global point, three
PDE-filter passes, chi = 1e4 * 0.1 (1 - rho) / (0.1 + rho).  One step is a
full Pn/Pn time step of the Fortran host (fluid_pnpn_t): explicit
advection, the pressure residual, the pressure solve to 1e-8, the velocity
residual and three Jacobi-PCG velocity solves to 1e-8.  All fields are
resident in HBM before the timed region.

Solver legs, timed on the same steps (after the same warm-up).  The timed
steps default to 20, two restart cycles of the 20-vector pressure
projection basis (it restarts when full, keeping the span of its 10 newest
solutions); with Neko's restart from the newest solution alone the steps
after a restart need 3-5 times the pressure iterations, which a window
shorter than a cycle can miss.  ms_per_step is the mean over the timed
steps; the pressure iterations fall as the flow settles, so every leg is
timed on the same steps.
  fp64                     FP64 Jacobi-PCG from zero (the reference's
                           configuration of the workload)
  mixed                    FP32 iterations with FP64 reliable updates
  mixed+proj               plus the previous-solution projection of the
                           pressure (basis 20, 10 kept on restart)
  mixed+proj+learned       plus the learned guess direction (trained on the
                           device on pairs harvested during the warm-up)
  gmres+hsmg               FP64 GMRES(30) + hybrid Schwarz multigrid for the
                           pressure (the shipped cases' pressure solver)
  mixed+proj+hsmg          GMRES + HSMG pressure, mixed velocity, projection
  mixed+proj+learned+hsmg  and the learned direction
Every solve stops on its FP64 residual < 1e-8.

Prints one JSON line: value = seconds per time step of the last leg (lower
is better), the per-leg times, speed-ups over fp64 and pressure iterations,
the ax_helm roofline (HIP events on the library stream; measured HBM
traffic), the whole-iteration PCG rate, gs_op, dealiased advection and
network throughput, and the CPU baseline (the FP64 oracle's step on the
host cores, timed on a bounded sample and scaled to one step).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neko-top_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of ax_helm_k<8,double,double> (pressure form) at
# 32^3 p=7: rocprofv3 FETCH_SIZE (x 1.9996, calibrated on an in-place stream
# of known size) + WRITE_SIZE (x 1.0), profiles/ax_helm_gs_traffic_32cubed_p7.md
AX_TRAFFIC_BYTES = {(32, 8): 1.2078e9}
METRIC = "s/timestep + ax_helm GB/s at 32^3 el p=7; CG iters, learned vs plain @1e-8"
DESIGN_SEED = 4321
MESH_SEED = 1234
PROJ_L = 20          # pressure projection size (Neko's default)
PROJ_KEEP = 10       # solutions kept when the full pressure basis restarts; mean
                     # over 20 steps at 32^3 with GMRES + HSMG (velocity basis 8):
                     # keep 0 (Neko's restart) 54.7 ms, (L, keep) = (20, 10) 40.9,
                     # (16, 8) 42.7, (12, 6) 47.5; over 40 steps (20, 0/10/14) 41.0/
                     # 34.8/35.7, (24, 16) 36.4, (30, 20) 37.9
PROJ_L_VEL = 0       # velocity projection off: the velocity solves take ~4
                     # iterations and their projection costs about what it saves
                     # (20 steps, pressure (20, 10): velocity basis 0: 39.8 ms,
                     # (4, 2) 41.9, (8, 0) 40.9, (8, 4) 42.0, (12, 6) 43.8)
PROJ_ACTIV = 2       # projection active from step 3 on
LEARN_PAIRS = 16     # pressure (residual, correction) pairs harvested
LEARN_EPOCHS = 20
LEARN_BATCH = 4096
# every leg is timed on the same steps: after PROJ_ACTIV + LEARN_PAIRS
# warmup steps
PROJ_WARMUP = PROJ_ACTIV + LEARN_PAIRS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def dealias_cost(lx, lxd, mode):
    """Algorithmic FP64 flops and HBM bytes per element of one dealiased
    advection call: one GLL->GL interpolation (three tensor passes) costs
    2 (lxd lx^3 + lxd^2 lx^2 + lxd^3 lx) flops, the map back the same; per
    GL point a reference gradient costs 6 lxd, opgrad 6 lxd + 21.  Bytes:
    the GLL inputs, the 9-field GL metric and f (read and written unless
    overwritten)."""
    iu = 2 * (lxd * lx ** 3 + lxd ** 2 * lx ** 2 + lxd ** 3 * lx)
    npg, np_ = lxd ** 3, lx ** 3
    opg = 6 * lxd + 21
    if mode == "fwd":
        flops = 6 * iu + npg * 3 * (opg + 6)
        nbytes = 8 * (3 * np_ + 9 * npg + 3 * np_)
    else:
        flops = 9 * iu + npg * (3 * opg + 36 + 3 * (5 + 6 * lxd))
        nbytes = 8 * (6 * np_ + 9 * npg + 6 * np_)
    return flops, nbytes


def kernel_rates(N, lx):
    """ax_helm / gs_op / copy rates on the library stream (HIP events)."""
    import ctypes as C
    from nekoamd import device, sem
    from nekoamd._lib import check, lib, ptr, ri, ri64
    case = sem.BoxCase(N, N, N, lx, jitter=0.05, seed=MESH_SEED, keep_xyz=False)
    n = case.n
    rng = np.random.default_rng(5)
    u = case.field(rng.uniform(-1, 1, n))
    w = case.field()
    one = case.field(np.ones(n))
    h2 = case.field(np.full(n, 1500.0))
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

    r = {}
    u32 = case.field(rng.uniform(-1, 1, n), np.float32)
    w32 = case.field(None, np.float32)
    o32 = case.field(np.ones(n), np.float32)
    # the three ax_helm forms are timed in interleaved bursts of 10 calls so
    # that no form sees only one phase of the clock governor; each figure is
    # the mean over its 4 bursts (40 calls)
    forms = {"ax": lambda: case.ax(w, u, one),
             "axv": lambda: case.ax(w, u, one, h2, True),
             "ax32": lambda: case.ax(w32, u32, o32, variant="f32")}
    # the GPU's clocks ramp up over the first few hundred ms of sustained load
    # (measured: 197 -> 253 -> 219 -> 209 us bursts, then a steady ~193 us):
    # run the operator for 1 s first, so every form is timed at the steady
    # state a solve runs in
    t_warm = time.perf_counter()
    while time.perf_counter() - t_warm < 1.0:
        for _ in range(50):
            forms["ax"]()
        device.sync()
    acc = {k: [] for k in forms}
    for _ in range(4):
        for k, fn in forms.items():
            acc[k].append(timed(fn, reps=10))
    r["ax_ms"] = statistics.mean(acc["ax"])
    r["ax_gbs"] = 72 * n / r["ax_ms"] / 1e6
    r["ax_tflops"] = (12 * lx ** 4 + 18 * lx ** 3) * case.nelv / r["ax_ms"] / 1e9
    r["ax_bursts_ms"] = acc["ax"]
    r["axv_gbs"] = 88 * n / statistics.mean(acc["axv"]) / 1e6
    r["ax32_gbs"] = 60 * n / statistics.mean(acc["ax32"]) / 1e6
    idx, bm, bc = case.gs.maps()
    nocc = int((bm * bc).sum())
    r["gs_ms"] = timed(lambda: case.gs.op(w))
    r["gs_gbs"] = nocc * 20 / r["gs_ms"] / 1e6
    r["gs_occ_per_point"] = nocc / n
    cp = timed(lambda: check(lib().nekoamd_copy(ptr(w), ptr(u), ri64(n)), "copy"))
    r["copy_gbs"] = 16 * n / cp / 1e6
    # one PCG iteration of the pressure solve (all its kernels), FP64 and
    # FP32-with-FP64-updates: a 40-iteration solve at abstol 0, one batch
    mask = case.mask(sem.PRS_DIRICHLET)
    dj = case.field()
    case.jacobi(dj, one)
    fh = case.field(rng.uniform(-1, 1, n))
    case.gs.op(fh)
    check(lib().nekoamd_bc_zero(ptr(fh), ptr(mask), ri64(mask.count)), "bc_zero")
    xs = case.field()
    for name, mixed in (("cg64", False), ("cg32", True)):
        cg = sem.CG(case, mixed=mixed)
        its = 40
        r[name + "_us"] = 1e3 * timed(lambda: cg.solve(xs, fh, dj, mask, one, abstol=0.0,
                                                       max_iter=its, poll=its), reps=3) / its
        del cg
    # dealiased advection (lxd = 3 lx / 2 Gauss-Legendre points): forward
    # term (overwrite) and adjoint term about a base flow, one fused kernel
    # per call
    dl_case = sem.BoxCase(N, N, N, lx, jitter=0.05, seed=MESH_SEED)
    dl = sem.Dealias(dl_case)
    uvw = [dl_case.field(rng.uniform(-1, 1, n)) for _ in range(3)]
    ub = [dl_case.field(rng.uniform(-1, 1, n)) for _ in range(3)]
    ff = [dl_case.field() for _ in range(3)]
    r["dl_fwd_ms"] = timed(lambda: dl.advection(ff, uvw, 0.0), reps=10)
    r["dl_adj_ms"] = timed(lambda: dl.adjoint(ff, uvw, ub), reps=10)
    r["dl_lxd"] = dl.lxd
    del dl, uvw, ub, ff, dl_case
    # the learned guess's network on every element: three BF16 MFMA GEMMs
    # (518 -> 256 -> 256 -> 512, input padded to 544), 2 x nelv x (544 x 256 +
    # 256 x 256 + 256 x 512) flops
    from nekoamd import learned
    net = learned.MLP(N, N, N, lx=lx, capacity=0, batch=4096, seed=3)
    ne = case.nelv
    X = device.DeviceArray.from_host(
        learned.to_bf16_bits(rng.standard_normal((ne, learned.KIN)).astype(np.float32)).ravel())
    Y = device.DeviceArray(ne * learned.NP, np.float32)
    fwd = lambda: check(lib().nekoamd_mlp_forward(C.c_void_p(net.h), ptr(X), ri(ne), ptr(Y),
                                                  None, None), "mlp_forward")
    r["mlp_ms"] = timed(fwd)
    r["mlp_tflops"] = 2.0 * ne * (544 * 256 + 256 * 256 + 256 * 512) / r["mlp_ms"] / 1e9
    del X, Y, net
    del u32, w32, o32
    return r


def cpu_baseline(N, lx, iters_per_step, cores):
    """The FP64 oracle on the host cores at the same size: one Pn/Pn step
    with every solve capped at one PCG iteration gives the residual work,
    a 10-iteration pressure PCG gives the cost of one iteration; one step is
    estimated as the residual work + (the GPU step's PCG iterations) x that."""
    import oracle
    oracle.set_threads(cores)
    tb = time.time()
    c = oracle.build_case(N, N, N, lx, jitter=0.05, seed=MESH_SEED)
    log(f"oracle case built in {time.time() - tb:.1f}s")
    ch = oracle.Channel(c, np.zeros(c.n), abstol_p=0.0, abstol_v=0.0, max_iter=1)
    t = time.perf_counter()
    ch.step()
    t_step1 = time.perf_counter() - t
    del ch
    mask = oracle.face_mask(N, N, N, lx, oracle.PRS_DIRICHLET)
    f = c.gs(c.B * np.sin(np.pi * c.x))
    f[mask] = 0.0
    h1 = np.ones(c.n)
    dinv = c.jacobi(h1)
    k = 10
    t = time.perf_counter()
    c.cg(f, mask, h1, abstol=0.0, max_iter=k, dinv=dinv)
    t_it = (time.perf_counter() - t) / k
    t_rest = max(t_step1 - 4 * t_it, 0.0)
    est = t_rest + iters_per_step * t_it
    return {"value": round(est, 3), "unit": "s/timestep", "cores": cores, "kind": "port",
            "sample": (f"FP64 oracle (oracle/pnpn.c + sem.c, gcc -O3, OpenMP {cores} threads) at "
                       f"{N}^3 p={lx - 1}: one Pn/Pn step with each solve capped at 1 PCG "
                       f"iteration ({t_step1:.2f} s) and 10 pressure-PCG iterations "
                       f"({t_it:.3f} s each); estimate = residual work {t_rest:.2f} s + "
                       f"{iters_per_step} PCG iterations of this step")}


SOLVERS = {
    # name: (FP32 iterations, projection basis size, learned guess, pressure
    # solver: Jacobi-PCG, or GMRES + hybrid Schwarz multigrid as the shipped
    # cases run the pressure, rugby_ball.case:46-52)
    "fp64": (False, 0, False, "cg"),        # the reference configuration
    "mixed": (True, 0, False, "cg"),
    "mixed+proj": (True, PROJ_L, False, "cg"),
    "mixed+proj+learned": (True, PROJ_L, True, "cg"),
    "gmres+hsmg": (False, 0, False, "gmres+hsmg"),
    "mixed+proj+hsmg": (True, PROJ_L, False, "gmres+hsmg"),
    "mixed+proj+learned+hsmg": (True, PROJ_L, True, "gmres+hsmg"),
    "mixed+proj+hsmg32": (True, PROJ_L, False, "gmres+hsmg32"),
    "mixed+proj+learned+hsmg32": (True, PROJ_L, True, "gmres+hsmg32"),
}
SOLVER_TEXT = {
    "cg": "Jacobi-PCG pressure",
    "gmres+hsmg": "GMRES(30) + hybrid Schwarz multigrid pressure",
    "gmres+hsmg32": "FP64 GMRES(30) + FP32 hybrid Schwarz multigrid pressure",
}


def run_channel(N, lx, solver, steps, warmup):
    """Time `steps` Pn/Pn steps (after `warmup`) of the channel with the given
    solver (see SOLVERS: 'fp64' is FP64 Jacobi-PCG from a zero initial guess;
    'mixed' runs FP32 iterations with FP64 reliable updates; '+proj' adds the
    previous-solution projection of the pressure and velocity solves;
    '+learned' harvests the pressure (residual, correction) pairs of
    LEARN_PAIRS steps, trains the element network on them and then adds its
    Galerkin-scaled direction to every pressure guess).  Every solve stops on
    its FP64 residual < 1e-8; projection legs are timed on the same steps."""
    from nekoamd import device
    from nekoamd.fortran import Channel
    mixed, proj, learned, prs = SOLVERS[solver]
    t0 = time.time()
    ch = Channel(N, N, N, lx, jitter=0.05, seed=MESH_SEED, re=100.0, dt=1e-3,
                 abstol_p=1e-8, abstol_v=1e-8, max_iter=5000, mixed=mixed)
    ch.design(DESIGN_SEED, npass=3, f_min=1e4, f_max=0.0, q=0.1)
    if prs == "gmres+hsmg":
        ch.pressure_solver("gmres", "hsmg")
    elif prs == "gmres+hsmg32":
        ch.pressure_solver("gmres", "hsmg_mixed")
    info = {}
    # every leg is timed on the same steps, after the projection legs have
    # filled their bases (and the learned leg has trained)
    warmup = max(warmup, PROJ_WARMUP)
    if proj:
        ch.projection(proj, PROJ_L_VEL, activ_step=PROJ_ACTIV, keep_prs=PROJ_KEEP)
    device.sync()
    log(f"[{solver}] channel built and design mapped: n={ch.n} in {time.time() - t0:.1f}s")
    done = 0
    if learned:
        ch.step(PROJ_ACTIV)
        ch.learned(Channel.LEARNED_HARVEST, capacity=LEARN_PAIRS, batch=LEARN_BATCH, seed=17)
        it, _ = ch.step(LEARN_PAIRS)
        done = PROJ_ACTIV + LEARN_PAIRS
        log(f"[{solver}] harvest steps: pressure iters {it[:, 0].tolist()}")
        device.sync()
        t = time.perf_counter()
        l0, l1, npairs = ch.train(epochs=LEARN_EPOCHS, batch=LEARN_BATCH, lr=1e-3, seed=23)
        device.sync()
        info = {"pairs": npairs, "epochs": LEARN_EPOCHS, "batch": LEARN_BATCH,
                "loss_first_epoch": l0, "loss_last_epoch": l1,
                "train_s": round(time.perf_counter() - t, 3)}
        log(f"[{solver}] trained on {npairs} pairs in {info['train_s']} s: loss {l0:.3e} -> "
            f"{l1:.3e}")
        ch.learned(Channel.LEARNED_APPLY)
    if warmup > done:
        it, res = ch.step(warmup - done)
        log(f"[{solver}] warmup iters {it.tolist()}")
    device.sync()
    step_ms, iters, finals, gammas = [], [], [], []
    for _ in range(steps):
        t = time.perf_counter()
        it, res = ch.step(1)
        device.sync()
        step_ms.append((time.perf_counter() - t) * 1e3)
        iters.append(it[0].tolist())
        finals.append(res[0].tolist())
        if learned:
            gammas.append(ch.gamma)
        log(f"[{solver}] step {len(step_ms)}: {step_ms[-1]:.1f} ms, iters {it[0].tolist()}, "
            f"res {['%.1e' % r for r in res[0]]}" + (f", gamma {gammas[-1]:.3e}" if learned else ""))
    u = ch.get("u")
    assert np.all(np.isfinite(u)), "non-finite velocity"
    assert all(r < 1e-8 for f in finals for r in f), "a solve missed 1e-8"
    n = ch.n
    ch.free()
    if learned:
        info["gamma_per_step"] = [float("%.4g" % g) for g in gammas]
    return {"ms": sum(step_ms) / len(step_ms), "step_ms": step_ms, "iters": iters, "n": n,
            "warmup": warmup,
            "max_final_res": max(max(f) for f in finals), "learned": info or None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=PROJ_L,
                    help="timed steps (default: two restart cycles of the pressure basis)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nel", type=int, default=32, help="elements per direction")
    ap.add_argument("--lx", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kernels", action="store_true", help="skip the kernel-rate leg")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the kernel-rate leg (ax_helm, gs_op, copy); prints its rates as "
                         "JSON (used for counter runs)")
    ap.add_argument("--solvers", default="fp64,mixed,mixed+proj,mixed+proj+learned,gmres+hsmg,"
                                         "mixed+proj+hsmg,mixed+proj+learned+hsmg",
                    help="comma list of channel solvers to time: " + ", ".join(SOLVERS))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from nekoamd import device

    device.init(0)
    dname = device.name()
    log("device:", dname)
    N, lx = args.nel, args.lx
    if args.kernels_only:
        kr = kernel_rates(N, lx)
        print(json.dumps({"kernel_rates": kr,
                          "elements": N ** 3, "lx": lx}), flush=True)
        return
    legs = {}
    for solver in args.solvers.split(","):
        legs[solver] = run_channel(N, lx, solver, args.steps, args.warmup)
    head = legs[args.solvers.split(",")[-1]]
    head_name = args.solvers.split(",")[-1]
    ms, iters, n = head["ms"], head["iters"], head["n"]

    kr = kernel_rates(N, lx) if not args.no_kernels else None
    if kr:
        log(f"ax_helm f64 {kr['ax_ms']*1e3:.1f} us {kr['ax_gbs']:.0f} GB/s; velocity form "
            f"{kr['axv_gbs']:.0f}; f32 {kr['ax32_gbs']:.0f}; gs {kr['gs_gbs']:.0f} GB/s; "
            f"copy {kr['copy_gbs']:.0f} GB/s; mlp {kr['mlp_ms']*1e3:.1f} us "
            f"{kr['mlp_tflops']:.1f} TFLOP/s")
    pit = [s[0] for s in iters]
    vit = [s[1] + s[2] + s[3] for s in iters]
    base = legs.get("fp64") or head
    tot = int(round(statistics.median([sum(s) for s in base["iters"]])))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(N, lx, tot, min(16, os.cpu_count() or 1))
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC,
        "value": round(ms / 1e3, 5),
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": head["warmup"],
        "ms_per_step": round(ms, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if SOLVERS[head_name][0] else "f64",
        "data": "synthetic",
        "config": {"workload": f"Brinkman channel Pn/Pn time step (BDF2/EXT2, Re=100, dt=1e-3, "
                               f"{SOLVER_TEXT[SOLVERS[head_name][3]]}, Jacobi-PCG velocity, "
                               f"every solve to an FP64 residual of 1e-8, seeded filtered "
                               f"design) on a {N}^3-element p={lx - 1} jittered box",
                   "solver": head_name,
                   "elements": N ** 3, "lx": lx, "gll_points": n, "device": dname,
                   "timing": f"mean wall clock over {args.steps} consecutive timed steps"
                             f" (the pressure projection basis restarts every "
                             f"{PROJ_L - PROJ_KEEP} steps)"},
        "roofline": None if not kr else {
            "bound": "hbm", "achieved": round(kr["ax_gbs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(kr["ax_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": AX_TRAFFIC_BYTES.get((N, lx)),
            "kernel": "ax_helm_k<8,double,double> pressure form, 72 B/point algorithmic",
            "avg_us": round(kr["ax_ms"] * 1e3, 2),
            "burst_avg_us": [round(t * 1e3, 1) for t in kr["ax_bursts_ms"]]},
        "cpu_baseline": cpu,
        "cg_iters": {"pressure_per_step": pit, "velocity_per_step": vit,
                     "learned": ({"pressure_per_step":
                                  [s[0] for s in legs["mixed+proj+learned"]["iters"]],
                                  "without_network": [s[0] for s in legs["mixed+proj"]["iters"]],
                                  "plain_fp64": [s[0] for s in legs["fp64"]["iters"]],
                                  "with_hsmg": ({"pressure_per_step": [
                                      s[0] for s in legs["mixed+proj+learned+hsmg"]["iters"]],
                                      "without_network": [
                                      s[0] for s in legs["mixed+proj+hsmg"]["iters"]]}
                                      if {"mixed+proj+learned+hsmg", "mixed+proj+hsmg"}
                                      <= set(legs) else None)}
                                 if {"mixed+proj+learned", "mixed+proj", "fp64"} <= set(legs)
                                 else None)},
        "solvers": {k: {"ms_per_step": round(v["ms"], 2),
                        "speedup_vs_fp64": (round(legs["fp64"]["ms"] / v["ms"], 3)
                                            if "fp64" in legs else None),
                        "pressure_solver": SOLVERS[k][3],
                        "step_ms": [round(t, 1) for t in v["step_ms"]],
                        "pressure_iters": [s[0] for s in v["iters"]],
                        "velocity_iters": [s[1] + s[2] + s[3] for s in v["iters"]],
                        "max_final_res": v["max_final_res"],
                        "warmup_steps": v["warmup"], "learned": v["learned"]}
                    for k, v in legs.items()},
        "speedup_vs_fp64": (round(legs["fp64"]["ms"] / ms, 3) if "fp64" in legs else None),
        "ax_helm": None if not kr else {
            "f64_GBps": round(kr["ax_gbs"], 1), "f64_TFLOPs": round(kr["ax_tflops"], 2),
            "f64_velocity_GBps": round(kr["axv_gbs"], 1), "f32_GBps": round(kr["ax32_gbs"], 1)},
        "gs_op_GBps": None if not kr else round(kr["gs_gbs"], 1),
        "pcg_iteration": None if not kr else {
            "fp64_us": round(kr["cg64_us"], 1),
            "fp64_GBps": round((152 + 20 * kr["gs_occ_per_point"]) * n / kr["cg64_us"] / 1e3, 1),
            "fp32_us": round(kr["cg32_us"], 1),
            "fp32_GBps": round((76 + 12 * kr["gs_occ_per_point"]) * n / kr["cg32_us"] / 1e3, 1),
            "bytes_per_point": "operator + update: FP64 152, FP32 76 (FP32 vectors, "
                               "geometry and x correction); gs: 20 (FP64) or "
                               "12 (FP32) bytes per shared-point occurrence"},
        "dealias_advection": None if not kr else {
            "lxd": kr["dl_lxd"],
            "forward_us": round(kr["dl_fwd_ms"] * 1e3, 1),
            "forward_TFLOPs": round(dealias_cost(lx, kr["dl_lxd"], "fwd")[0] * N ** 3
                                    / kr["dl_fwd_ms"] / 1e9, 2),
            "forward_GBps": round(dealias_cost(lx, kr["dl_lxd"], "fwd")[1] * N ** 3
                                  / kr["dl_fwd_ms"] / 1e6, 1),
            "adjoint_us": round(kr["dl_adj_ms"] * 1e3, 1),
            "adjoint_TFLOPs": round(dealias_cost(lx, kr["dl_lxd"], "adj")[0] * N ** 3
                                    / kr["dl_adj_ms"] / 1e9, 2),
            "adjoint_GBps": round(dealias_cost(lx, kr["dl_lxd"], "adj")[1] * N ** 3
                                  / kr["dl_adj_ms"] / 1e6, 1),
            "fp64_peak_TFLOPs": 78.6},
        "mlp_inference": None if not kr else {
            "elements": N ** 3, "us": round(kr["mlp_ms"] * 1e3, 1),
            "TFLOPs": round(kr["mlp_tflops"], 1), "mfma_bf16_peak_TFLOPs": 2500.0,
            "mfma_frac": round(kr["mlp_tflops"] / 2500.0, 4)},
        "copy_GBps": None if not kr else round(kr["copy_gbs"], 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
