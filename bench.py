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
                           pressure (basis 20, 10 kept on restart; the
                           mixed legs keep the basis in FP32, the
                           projected right-hand side exact in FP64)
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
# every leg runs the same prefill steps before its warm-up and timed steps:
# the projection legs fill their bases, the learned legs harvest and train
PREFILL = PROJ_ACTIV + LEARN_PAIRS
CONFIG2_STEPS = 100  # the plain FP64 leg runs the reference's 100-step protocol


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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may run on (the box's CPU share), not the host's
    total."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_kernel_rates(c, threads):
    """ax_helm (pressure form, 72 B/point) and gs_op (20 B per shared-point
    occurrence) of the FP64 oracle, GB/s of algorithmic bytes, per thread
    count (best of 3 calls)."""
    import oracle
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, c.n)
    nocc = int(c.gs_idx.size)
    out = {}
    for th in threads:
        oracle.set_threads(th)
        ta, tg = [], []
        for _ in range(3):
            t = time.perf_counter()
            c.ax(u)
            ta.append(time.perf_counter() - t)
            t = time.perf_counter()
            c.gs(u)
            tg.append(time.perf_counter() - t)
        out[str(th)] = {"ax_helm_GBps": round(72 * c.n / min(ta) / 1e9, 2),
                        "gs_op_GBps": round(20 * nocc / min(tg) / 1e9, 2)}
    return out


def config1_cpu(threads, steps=50):
    """Config 1 on the CPU: the 8x4x4 p=5 Brinkman channel with the seeded
    filtered design, 50 BDF2/EXT2 steps, Jacobi-PCG solves to 1e-8 -- all of
    it, timed per thread count."""
    import oracle
    c = oracle.build_case(8, 4, 4, 6, jitter=0.05, seed=MESH_SEED)
    rho, rf, chi = oracle.seeded_design(c, DESIGN_SEED, npass=3)
    out = {}
    for th in threads:
        oracle.set_threads(th)
        ch = oracle.Channel(c, chi, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
        t = time.perf_counter()
        its = [ch.step()[0] for _ in range(steps)]
        dt = time.perf_counter() - t
        out[str(th)] = {"total_s": round(dt, 3), "s_per_step": round(dt / steps, 5),
                        "pressure_iters_first_last": [its[0][0], its[-1][0]]}
    return out


def cpu_sample_32(N, lx, cores):
    """The bounded sample at the workload's size: the FP64 oracle's Pn/Pn
    step with every solve capped at one PCG iteration (all the residual,
    advection and forcing work of a step, measured) and 10 iterations each of
    the pressure-form and Helmholtz-form PCG (the cost of one iteration,
    measured)."""
    import oracle
    oracle.set_threads(cores)
    c = oracle.build_case(N, N, N, lx, jitter=0.05, seed=MESH_SEED)
    ch = oracle.Channel(c, np.zeros(c.n), abstol_p=0.0, abstol_v=0.0, max_iter=1)
    t = time.perf_counter()
    ch.step()
    t_capped = time.perf_counter() - t
    del ch
    mask = oracle.face_mask(N, N, N, lx, oracle.PRS_DIRICHLET)
    f = c.gs(c.B * np.sin(np.pi * c.x))
    f[mask] = 0.0
    one = np.ones(c.n)
    k = 10
    t = time.perf_counter()
    c.cg(f, mask, one, abstol=0.0, max_iter=k, dinv=c.jacobi(one))
    t_itp = (time.perf_counter() - t) / k
    vmask = oracle.face_mask(N, N, N, lx, oracle.VEL_DIRICHLET)
    h2 = np.full(c.n, 1500.0)
    t = time.perf_counter()
    c.cg(f, vmask, 0.01 * one, h2, True, abstol=0.0, max_iter=k,
         dinv=c.jacobi(0.01 * one, h2, True))
    t_itv = (time.perf_counter() - t) / k
    return c, {"capped_step_s": t_capped, "pcg_iter_s": t_itp, "helmholtz_iter_s": t_itv}


def cpu_baseline(N, lx, fp64_iters, cores):
    """CPU baseline on the host cores (FP64 oracle, OpenMP): config 1 run in
    full; at the workload's size a bounded sample (residual work of a step
    plus the cost of one PCG iteration, both measured) scaled to the FP64
    GPU leg's median step (its pressure and velocity iteration counts; the
    oracle's PCG takes the same counts to +-1).  The sample's scaling is
    checked against a full measured step by bench.py --cpu-full-step."""
    threads = sorted({1, cores})
    t0 = time.time()
    c1 = config1_cpu(threads)
    c, smp = cpu_sample_32(N, lx, cores)
    rates = cpu_kernel_rates(c, threads)
    del c
    steps = sorted(fp64_iters, key=lambda s: sum(s))
    med = steps[len(steps) // 2]
    resid = max(smp["capped_step_s"] - smp["pcg_iter_s"] - 3 * smp["helmholtz_iter_s"], 0.0)
    est = resid + med[0] * smp["pcg_iter_s"] + sum(med[1:]) * smp["helmholtz_iter_s"]
    log(f"cpu baseline leg took {time.time() - t0:.1f} s")
    return {"value": round(est, 3), "unit": "s/timestep", "cores": cores, "kind": "port",
            "sample": (f"FP64 oracle (oracle/*.c, gcc -O3 -march=x86-64-v3, OpenMP {cores} "
                       f"threads) at {N}^3 p={lx - 1}: one Pn/Pn step with every solve capped at "
                       f"1 PCG iteration ({smp['capped_step_s']:.2f} s measured), 10 pressure "
                       f"and 10 Helmholtz PCG iterations ({smp['pcg_iter_s']:.4f} / "
                       f"{smp['helmholtz_iter_s']:.4f} s each, measured); value = residual "
                       f"work {resid:.2f} s + the FP64 leg's median step (pressure {med[0]}, "
                       f"velocity {sum(med[1:])} iterations) at those rates"),
            "cpu_model": cpu_model(), "host_cores": host_cores(),
            "config1_full_run": {"workload": "8x4x4 p=5 Brinkman channel, seeded filtered "
                                             "design, 50 steps, Jacobi-PCG to 1e-8",
                                 "threads": c1},
            "kernels_at_workload_size": rates}


def cpu_full_step(N, lx, cores, steps=1, start=None):
    """Complete FP64 oracle steps at the workload's size on the host cores,
    every solve to 1e-8, timed end to end, next to the bounded sample's
    estimate of the same steps from their iteration counts.  The steps are
    the first ones of the bench's timed window: the FP64 GPU channel runs the
    `start` steps before it (its state is pinned to the oracle's to 1e-11),
    and its state (u, v, w, p, the lags, the step counter) is moved into the
    oracle, which then runs the next `steps` steps."""
    import oracle
    from nekoamd import device
    from nekoamd.fortran import Channel
    start = PREFILL + 2 if start is None else start
    device.init(0)
    ch = Channel(N, N, N, lx, jitter=0.05, seed=MESH_SEED, re=100.0, dt=1e-3,
                 abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
    ch.design(DESIGN_SEED, npass=3, f_min=1e4, f_max=0.0, q=0.1)
    for k in range(start):
        it, _ = ch.step(1)
        log(f"gpu fp64 step {k + 1}: iters {it[0].tolist()}")
    state = {k: ch.get(k) for k in ("u", "v", "w", "p", "ulag", "vlag", "wlag",
                                    "fxlag", "fylag", "fzlag")}
    chi, ts = ch.get("chi"), ch.tstep()
    ch.free()
    oracle.set_threads(cores)
    c, smp = cpu_sample_32(N, lx, cores)
    ref = oracle.Channel(c, chi, abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
    for k in ("u", "v", "w", "p"):
        getattr(ref, k)[:] = state[k]
    for a, k in zip(ref.lags, ("ulag", "vlag", "wlag", "fxlag", "fylag", "fzlag")):
        a[:] = state[k]
    ref.s.tstep = ts
    os.environ["OR_PROGRESS"] = "1"
    rows = []
    for k in range(steps):
        t = time.perf_counter()
        it, res = ref.step()
        dt = time.perf_counter() - t
        resid = max(smp["capped_step_s"] - smp["pcg_iter_s"] - 3 * smp["helmholtz_iter_s"], 0.0)
        est = resid + it[0] * smp["pcg_iter_s"] + sum(it[1:]) * smp["helmholtz_iter_s"]
        rows.append({"step": ts + k + 1, "measured_s": round(dt, 2), "iters": it,
                     "sample_estimate_s": round(est, 2)})
        log(f"cpu full step {ts + k + 1}: {dt:.1f} s, iters {it}, sample estimate {est:.1f} s")
    return {"cpu_model": cpu_model(), "host_cores": host_cores(), "threads": cores,
            "workload": f"{N}^3 p={lx - 1} Brinkman channel, seeded filtered design, FP64 "
                        f"Jacobi-PCG to 1e-8, steps {ts + 1}..{ts + steps} from the FP64 GPU "
                        f"run's state after step {ts}",
            "sample": smp, "steps": rows}


def config4(N, lx, harvest_steps=200, eval_steps=200, epochs=20, every=16):
    """The learned-aid protocol on the 32^3 p=7 channel: FP64 Jacobi-PCG for
    steps 1..harvest_steps, harvesting the (residual, correction) pair of
    every pressure solve (one per step: 200 pairs; ten per step, as the
    protocol's 2000, would hold 206 GB of BF16 features and FP32 targets),
    training the network on them, then steps harvest_steps+1 ..
    harvest_steps+eval_steps three times from the same checkpointed state:
    plain FP64 PCG; the learned initial guess; and the learned guess plus the
    network's correction inside flexible CG every `every`-th iteration.  Per
    variant: pressure iterations, time per step and the largest final
    residual."""
    import tempfile
    from nekoamd import device
    from nekoamd.fortran import Channel
    t0 = time.time()
    ch = Channel(N, N, N, lx, jitter=0.05, seed=MESH_SEED, re=100.0, dt=1e-3,
                 abstol_p=1e-8, abstol_v=1e-8, max_iter=5000)
    ch.design(DESIGN_SEED, npass=3, f_min=1e4, f_max=0.0, q=0.1)
    ch.learned(Channel.LEARNED_HARVEST, capacity=harvest_steps, batch=LEARN_BATCH, seed=17)
    hit = []
    for k in range(harvest_steps):
        it, _ = ch.step(1)
        hit.append(int(it[0][0]))
        if k % 10 == 9:
            log(f"[config4] harvest step {k + 1}: pressure iterations {hit[-1]}")
    device.sync()
    t_h = time.time() - t0
    t = time.time()
    l0, l1, npairs = ch.train(epochs=epochs, batch=LEARN_BATCH, lr=1e-3, seed=23)
    device.sync()
    t_train = time.time() - t
    log(f"[config4] trained on {npairs} pairs in {t_train:.1f} s: loss {l0:.3e} -> {l1:.3e}")
    ch.learned(Channel.LEARNED_OFF)
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "config4.chkp")
    ch.checkpoint(path)
    out = {}
    for name in ("plain_pcg", "learned_guess", "learned_guess+fcg_correction"):
        ch.restart(path)
        if name == "plain_pcg":
            ch.pressure_solver("cg", "jacobi")
            ch.learned(Channel.LEARNED_OFF)
            ch.learned_correction(0)
        elif name == "learned_guess":
            ch.pressure_solver("cg", "jacobi")
            ch.learned(Channel.LEARNED_APPLY)
            ch.learned_correction(0)
        else:
            ch.pressure_solver("fcg", "jacobi")
            ch.learned(Channel.LEARNED_APPLY)
            ch.learned_correction(every)
        its, ms, res = [], [], []
        for k in range(eval_steps):
            t = time.perf_counter()
            it, r = ch.step(1)
            device.sync()
            ms.append((time.perf_counter() - t) * 1e3)
            its.append(int(it[0][0]))
            res.append(float(max(r[0])))
            if k % 20 == 19:
                log(f"[config4] {name} step {harvest_steps + k + 1}: pressure iterations "
                    f"{its[-1]}, {ms[-1]:.1f} ms")
        out[name] = {"pressure_iters_total": sum(its), "pressure_iters_per_step": its,
                     "ms_per_step_mean": round(sum(ms) / len(ms), 2),
                     "ms_per_step_median": round(statistics.median(ms), 2),
                     "max_final_residual": max(res)}
    ch.free()
    os.remove(path)
    os.rmdir(tmp)
    return {"workload": f"{N}^3 p={lx - 1} Brinkman channel, seeded filtered design, "
                        f"FP64 solves to 1e-8",
            "harvest": {"steps": harvest_steps, "pairs": npairs,
                        "pressure_iters_per_step": hit, "seconds": round(t_h, 1)},
            "training": {"epochs": epochs, "batch": LEARN_BATCH, "loss_first_epoch": l0,
                         "loss_last_epoch": l1, "seconds": round(t_train, 1)},
            "correction_every": every,
            "evaluation_steps": [harvest_steps + 1, harvest_steps + eval_steps],
            "variants": out}


def config5(nx=32, ny=16, nz=16, lx=8, iters=20, nfwd=200, nadj=200, dt=1e-4):
    """The optimisation loop of the topology-optimisation workload: a
    nx*ny*nz channel of order lx-1 with per-GLL design, `iters` design
    iterations of a forward Pn/Pn solve (nfwd steps from the initial state),
    objective, adjoint solve (nadj steps), sensitivity, chain rule and an MMA
    update with one volume constraint (fraction 0.4).  Run twice: the fast
    solver (FP32-iteration velocity PCG with FP64 reliable updates, GMRES +
    hybrid Schwarz multigrid pressure) and the FP64 solver (FP64 velocity
    PCG, the same FP64 GMRES + HSMG pressure, whose steps are pinned to the
    oracle's); the objective, volume and sensitivity of the fast loop are
    compared with the FP64 loop's at every iteration.  dt = 1e-4, the time
    step of the reference's own optimisation case (rugby_ball.case:10): at
    the channels' 1e-3 the explicit advection becomes unstable once the
    optimiser opens narrow passages (the velocity there reaches ~60; the
    oracle's loop on a 4x3x2 p=5 channel diverges in its second forward
    solve, this loop at its ninth), at 1e-4 the oracle's 20 iterations
    lower F monotonically.  The loop stops at the first non-finite
    objective or KKT residual."""
    from nekoamd import device
    from nekoamd.fortran import TopOpt
    runs = {}
    for name, mixed in (("fp64", False), ("fast", True)):
        T = TopOpt(nx, ny, nz, lx, jitter=0.05, seed=MESH_SEED, dt=dt, abstol_p=1e-8,
                   abstol_v=1e-8, max_iter=5000, mixed=mixed, design_seed=DESIGN_SEED, npass=3,
                   v_max=0.4)
        T.pressure_solver("gmres", "hsmg")
        rows, sens = [], []
        t_all = time.time()
        for k in range(iters):
            device.sync()
            t = time.perf_counter()
            a = T.iterate(nfwd, nadj)
            device.sync()
            a["seconds"] = round(time.perf_counter() - t, 3)
            rows.append(a)
            sens.append(T.get("dF_drho"))
            log(f"[config5 {name}] design iteration {k + 1}: F {a['F']:.10e}, volume "
                f"{a['volume']:.6f}, KKT {a['residunorm']:.3e}, pressure iterations "
                f"{a['p_iters_fwd']}/{a['p_iters_adj']}, {a['seconds']} s")
            if not (np.isfinite(a["F"]) and np.isfinite(a["residunorm"])):
                log(f"[config5 {name}] non-finite objective or KKT residual: stopping")
                break
        x = T.get("x")
        T.free()
        runs[name] = {"rows": rows, "sens": sens, "x": x, "total_s": round(time.time() - t_all, 1)}
    f64, fast = runs["fp64"], runs["fast"]
    cmp = []
    for k in range(min(len(f64["rows"]), len(fast["rows"]))):
        a, b = fast["rows"][k], f64["rows"][k]
        sb = f64["sens"][k]
        cmp.append({"iteration": k + 1,
                    "F_rel_err": abs(a["F"] - b["F"]) / abs(b["F"]),
                    "volume_abs_err": abs(a["volume"] - b["volume"]),
                    "sensitivity_rel_err": float(np.abs(fast["sens"][k] - sb).max()
                                                 / np.abs(sb).max())})
    return {"workload": f"{nx}x{ny}x{nz} p={lx - 1} Brinkman channel, per-GLL design, {iters} "
                        f"design iterations of {nfwd} forward and {nadj} adjoint steps "
                        f"(dt = {dt:g}), MMA with one volume constraint (0.4); solves to 1e-8",
            "fast_solver": "FP32-iteration velocity PCG (FP64 reliable updates) + FP64 GMRES "
                           "and hybrid Schwarz multigrid pressure",
            "fp64_solver": "FP64 velocity PCG + FP64 GMRES and hybrid Schwarz multigrid "
                           "pressure",
            "design_variables": int(f64["x"].size),
            "seconds": {"fp64": f64["total_s"], "fast": fast["total_s"]},
            "per_iteration": {name: [{k: r[k] for k in ("F", "volume", "residunorm",
                                                         "p_iters_fwd", "p_iters_adj",
                                                         "seconds")} for r in runs[name]["rows"]]
                              for name in runs},
            "fast_vs_fp64": cmp,
            "final_design_max_abs_diff": float(np.abs(fast["x"] - f64["x"]).max())}


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


def run_channel(N, lx, solver, steps, warmup, total=0):
    """Time `steps` Pn/Pn steps of the channel with the given solver (see
    SOLVERS: 'fp64' is FP64 Jacobi-PCG from a zero initial guess; 'mixed'
    runs FP32 iterations with FP64 reliable updates; '+proj' adds the
    previous-solution projection of the pressure solve; '+learned' harvests
    the pressure (residual, correction) pairs of LEARN_PAIRS steps, trains
    the element network on them and then adds its Galerkin-scaled direction
    to every pressure guess).  Every leg first runs the same PREFILL steps
    (the projection basis fills, the learned leg harvests and trains), then
    `warmup` untimed steps, then the `steps` timed ones, so every leg is
    timed on the same steps.  Each step is timed on its own (device sync on
    both sides).  total > 0 keeps stepping (each step timed) until `total`
    steps have run, for the reference's per-run step statistic.  Every
    solve stops on its FP64 residual < 1e-8."""
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
    if proj:
        # the mixed legs keep the basis in FP32 (nekoamd_proj_set_fp32): at
        # 32^3 p=7 with GMRES + HSMG 21.8-22.0 against 22.8-22.9 ms per step
        # with the FP64 basis, 108 against 105 pressure iterations over the
        # 20 timed steps
        ch.projection(proj, PROJ_L_VEL, activ_step=PROJ_ACTIV, keep_prs=PROJ_KEEP,
                      fp32=mixed)
    device.sync()
    log(f"[{solver}] channel built and design mapped: n={ch.n} in {time.time() - t0:.1f}s")
    all_ms, all_it, all_res, gammas = [], [], [], []

    def one(tag):
        t = time.perf_counter()
        it, res = ch.step(1)
        device.sync()
        all_ms.append((time.perf_counter() - t) * 1e3)
        all_it.append(it[0].tolist())
        all_res.append(res[0].tolist())
        if learned and tag == "timed":
            gammas.append(ch.gamma)
        log(f"[{solver}] {tag} step {len(all_ms)}: {all_ms[-1]:.1f} ms, iters "
            f"{it[0].tolist()}, res {['%.1e' % r for r in res[0]]}"
            + (f", gamma {gammas[-1]:.3e}" if learned and tag == "timed" else ""))

    if learned:
        for _ in range(PROJ_ACTIV):
            one("prefill")
        ch.learned(Channel.LEARNED_HARVEST, capacity=LEARN_PAIRS, batch=LEARN_BATCH, seed=17)
        for _ in range(LEARN_PAIRS):
            one("harvest")
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
    while len(all_ms) < PREFILL:
        one("prefill")
    for _ in range(warmup):
        one("warmup")
    t_first = len(all_ms)
    for _ in range(steps):
        one("timed")
    while len(all_ms) < total:
        one("extra")
    win = slice(t_first, t_first + steps)
    step_ms, iters, finals = all_ms[win], all_it[win], all_res[win]
    u = ch.get("u")
    assert np.all(np.isfinite(u)), "non-finite velocity"
    assert all(r < 1e-8 for f in all_res for r in f), "a solve missed 1e-8"
    n = ch.n
    ch.free()
    if learned:
        info["gamma_per_step"] = [float("%.4g" % g) for g in gammas]
    out = {"ms": sum(step_ms) / len(step_ms), "step_ms": step_ms, "iters": iters, "n": n,
           "prefill": PREFILL, "warmup": warmup, "first_timed_step": t_first + 1,
           "max_final_res": max(max(f) for f in finals), "learned": info or None}
    if total:
        # the reference's "Step time" statistic (simulation_adjoint.f90:107-130):
        # the median wall clock over steps 11..total of one run
        out["median_ms_steps_11_to_total"] = statistics.median(all_ms[10:total])
        out["total_steps"] = total
        out["pressure_iters_all"] = [s[0] for s in all_it[:total]]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=PROJ_L,
                    help="timed steps (default: two restart cycles of the pressure basis)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nel", type=int, default=32, help="elements per direction")
    ap.add_argument("--lx", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--config5", type=int, default=0, metavar="K",
                    help="only the optimisation loop on the 32x16x16 p=7 channel, K design "
                         "iterations of 200 forward and 200 adjoint steps, fast and FP64 "
                         "solvers; prints JSON")
    ap.add_argument("--config4", action="store_true",
                    help="only the learned-aid protocol (200 harvest steps, training, "
                         "steps 201-400 plain / learned guess / guess + FCG correction); "
                         "prints JSON")
    ap.add_argument("--cpu-full-step", type=int, default=0, metavar="K",
                    help="only time K complete FP64 oracle steps at the workload's size on the "
                         "host cores (no GPU work) next to the bounded sample's estimate of "
                         "them; prints JSON")
    ap.add_argument("--config2-steps", type=int, default=CONFIG2_STEPS,
                    help="steps of the plain FP64 leg (the reference's step-time statistic is "
                         "the median over steps 11..this); 0 runs it like the other legs")
    ap.add_argument("--no-kernels", action="store_true", help="skip the kernel-rate leg")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the kernel-rate leg (ax_helm, gs_op, copy); prints its rates as "
                         "JSON (used for counter runs)")
    ap.add_argument("--solvers", default="fp64,mixed,mixed+proj,mixed+proj+learned,gmres+hsmg,"
                                         "mixed+proj+learned+hsmg,mixed+proj+hsmg",
                    help="comma list of channel solvers to time: " + ", ".join(SOLVERS))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    N, lx = args.nel, args.lx
    if args.config5:
        from nekoamd import device
        device.init(0)
        print(json.dumps({"config5": config5(iters=args.config5)}), flush=True)
        return
    if args.config4:
        from nekoamd import device
        device.init(0)
        print(json.dumps({"config4": config4(N, lx)}), flush=True)
        return
    if args.cpu_full_step:
        cores = min(16, host_cores())
        print(json.dumps({"cpu_full_step": cpu_full_step(N, lx, cores, args.cpu_full_step)}),
              flush=True)
        return

    from nekoamd import device

    device.init(0)
    dname = device.name()
    log("device:", dname)
    if args.kernels_only:
        kr = kernel_rates(N, lx)
        print(json.dumps({"kernel_rates": kr,
                          "elements": N ** 3, "lx": lx}), flush=True)
        return
    legs = {}
    for solver in args.solvers.split(","):
        legs[solver] = run_channel(N, lx, solver, args.steps, args.warmup,
                                   total=args.config2_steps if solver == "fp64" else 0)
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
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(N, lx, base["iters"], min(16, host_cores()))
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC,
        "value": round(ms / 1e3, 5),
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
                   "timing": f"mean wall clock over {args.steps} consecutive timed steps "
                             f"(steps {head['first_timed_step']}.."
                             f"{head['first_timed_step'] + args.steps - 1}), each between "
                             f"device syncs, after {PREFILL} prefill steps (every leg; the "
                             f"projection basis fills, the learned leg harvests and trains) "
                             f"and {args.warmup} warm-up steps; the pressure projection basis "
                             f"restarts every {PROJ_L - PROJ_KEEP} steps",
                   "prefill_steps": PREFILL},
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
                        "first_timed_step": v["first_timed_step"], "learned": v["learned"]}
                    for k, v in legs.items()},
        "speedup_vs_fp64": (round(legs["fp64"]["ms"] / ms, 3) if "fp64" in legs else None),
        "config2_fp64": (None if "median_ms_steps_11_to_total" not in base else {
            "workload": f"plain FP64 Jacobi-PCG, {base['total_steps']} steps from the start",
            "median_s_steps_11_to_end": round(base["median_ms_steps_11_to_total"] / 1e3, 5),
            "pressure_iters_per_step": base["pressure_iters_all"],
            "note": "the reference's Step time statistic (simulation_adjoint.f90:107-130)"}),
        "ax_helm": None if not kr else {
            "f64_GBps": round(kr["ax_gbs"], 1), "f64_TFLOPs": round(kr["ax_tflops"], 2),
            "f64_velocity_GBps": round(kr["axv_gbs"], 1), "f32_GBps": round(kr["ax32_gbs"], 1)},
        "gs_op_GBps": None if not kr else round(kr["gs_gbs"], 1),
        "pcg_iteration": None if not kr else {
            "fp64_us": round(kr["cg64_us"], 1),
            "fp64_GBps": round((152 + 20 * kr["gs_occ_per_point"]) * n / kr["cg64_us"] / 1e3, 1),
            "fp32_us": round(kr["cg32_us"], 1),
            "fp32_GBps": round((84 + 12 * kr["gs_occ_per_point"]) * n / kr["cg32_us"] / 1e3, 1),
            "bytes_per_point": "operator + update: FP64 152, FP32 84 (FP32 vectors and "
                               "geometry, FP64 x); gs: 20 (FP64) or "
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
