"""Time steps per second of the Bluebottle step on one MI355X.

Workload: particle sedimentation in a 256^3 triply-periodic box with 1000
spheres of radius 6 dx (volume fraction 5.4 %), density ratio 2, Re_p = 10,
Lamb order L = 4 with the 26-node rule, pressure Poisson solve at relative
residual 1e-8 (Jacobi-preconditioned BiCGSTAB), synthetic seeded particle
positions, quiescent start.  One step is the full src/bluebottle.c:598-711
sequence (forcing, cages, the Lamb sub-iteration loop of predictor /
Poisson solve / projection / pressure update / particle BCs / quadrature,
particle motion, history shift, next dt), driven through the C ABI.

Both solver forms run on the same case, one after the other: the plain
FP64 BiCGSTAB and the mixed form (FP32 BiCGSTAB inside FP64 iterative
refinement).  `value` is the mixed form's rate.  Per-kernel durations come
from HIP events recorded on the library stream around every launch inside
the timed region; the roofline entry is the kernel with the largest total
time, priced at its algorithmic bytes per launch.

The CPU baseline times the serial FP64 oracle on this host over a bounded
sample (one pass of every stage of a Lamb sub-iteration plus a fixed number
of BiCGSTAB iterations at full size) and extrapolates to a step with the
Lamb sub-iteration and Krylov iteration counts the GPU's FP64 run measured.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("timesteps/s @256³+1000 spheres, learned+mixed vs plain FP64; "
          "HBM GB/s")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def kernel_bytes(n, xn, yn, zn):
    """Algorithmic bytes per launch for the single-launch kernel kinds:
    every input and output array once (FP64 8 B, FP32 4 B, int32 4 B,
    packed operator flags 1 B)."""
    nf = (xn + 1) * yn * zn + xn * (yn + 1) * zn + xn * yn * (zn + 1)
    return {
        # 13 FP64 reads + 1 int, 9 FP64 writes per face (3 components)
        "u_star": 180 * n,
        # u*, v*, w*, phase, phase_shell, flag_u/v/w, phi -> rhs, x, flags
        "pp_setup": (24 + 8 + 12 + 8 + 8 + 8 + 1) * n,
        # three sweeps per iteration (FP64, and FP32 by default):
        # r, p, v, r*, flags -> p, v  (the p update folded into the sweep)
        "spmv_p": (6 * 8 + 1) * n, "spmv_p32": (6 * 4 + 1) * n,
        # r, v, flags -> t
        "spmv_s": (3 * 8 + 1) * n, "spmv_s32": (3 * 4 + 1) * n,
        # x, p, r, v, t, r*, flags -> x, r
        "update_xr": (8 * 8 + 1) * n, "update_xr32": (8 * 4 + 1) * n,
        # the two-sweep FP32 form (--inner-sweeps 2):
        # sweep A: r, p, v, t, x, r*, flags -> r', x, p', v'
        "sweep_a32": (10 * 4 + 1) * n,
        # sweep B: r', v', r*, flags -> t
        "sweep_b32": (4 * 4 + 1) * n,
        # b, x, flags -> r (FP64)
        "resid64": (3 * 8 + 1) * n,
        # x, e -> x
        "correct64": (2 * 8 + 4) * n,
        # u*, v*, w*, phi, flag_u/v/w -> u, v, w
        "project": 68 * n,
        # p0, phi, phase -> p (fused sum and subtract)
        "update_pressure": 44 * n,
        "find_dt": 8 * nf,
    }


def run_mode(s, mode, steps, warmup, prof_steps, runtime):
    """Warm-up steps, then `steps` timed steps with no instrumentation (the
    rate), then `prof_steps` more steps with HIP events recorded on the
    library stream around every launch (the per-kernel table and the
    roofline)."""
    host, hip = s.host, s.hip
    s.set_solver(mode)
    for i in range(warmup):
        t0 = time.perf_counter()
        it = s.step()
        hip.bb_synchronize()
        log("mode %d warmup step %d: %d Lamb iterations, %.3f s" % (
            mode, i, it, time.perf_counter() - t0))
    import ctypes as C
    tot = C.c_long.in_dll(host, "bb_rec_total_iters")
    calls = C.c_int.in_dll(host, "bb_rec_calls")
    it0, c0 = tot.value, calls.value
    hip.bb_prof_enable(0)
    lamb_iters = 0
    hip.bb_synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        lamb_iters += s.step()
        log("mode %d timed step %d done (%.3f s elapsed)" % (
            mode, i, time.perf_counter() - t0))
    hip.bb_synchronize()
    elapsed = time.perf_counter() - t0
    it1, c1 = tot.value, calls.value
    # profiled steps: per-kernel event timing
    hip.bb_prof_reset()
    hip.bb_prof_enable(1)
    hip.bb_synchronize()
    t1 = time.perf_counter()
    pl = 0
    for i in range(prof_steps):
        pl += s.step()
        log("mode %d profiled step %d done" % (mode, i))
    hip.bb_synchronize()
    prof_elapsed = time.perf_counter() - t1
    hip.bb_prof_enable(0)
    prof = {}
    for k, name in enumerate(runtime.PROF_KINDS):
        ms = C.c_double()
        cnt = C.c_long()
        hip.bb_prof_read(k, C.byref(ms), C.byref(cnt))
        if cnt.value:
            prof[name] = (ms.value, cnt.value)
    return {
        "seconds": elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "timesteps_per_s": steps / elapsed,
        "lamb_iters_per_step": lamb_iters / steps,
        "krylov_iters_per_solve": (it1 - it0) / max(c1 - c0, 1),
        "solves": c1 - c0,
        "krylov_iters": it1 - it0,
        "profiled_steps": prof_steps,
        "profiled_ms_per_step": (1e3 * prof_elapsed / prof_steps
                                 if prof_steps else None),
        "profiled_krylov_iters_per_solve": (
            (tot.value - it1) / max(calls.value - c1, 1)),
        "prof": prof,
    }


def prepare_learned(s, warm_steps, iters, lr, capacity):
    """In-situ training of the learned preconditioner on this state: mixed
    steps keep the (residual, correction) pairs of their Jacobi inner
    solves, then Adam trains the network on them."""
    import ctypes as C
    hip = s.hip
    hip.bb_learn_collect(capacity)
    s.set_solver(1)
    t0 = time.perf_counter()
    for i in range(warm_steps):
        s.step()
        log("learned: collection step %d done, %d pairs" % (
            i, hip.bb_learn_pairs()))
    hip.bb_synchronize()
    t_collect = time.perf_counter() - t0
    npairs = hip.bb_learn_pairs()
    hip.bb_unet_init(20251202)
    base = [hip.bb_learn_pair_loss(k) for k in range(npairs)]
    t0 = time.perf_counter()
    loss = hip.bb_learn_train(iters, lr, 7)
    hip.bb_synchronize()
    t_train = time.perf_counter() - t0
    after = [hip.bb_learn_pair_loss(k) for k in range(npairs)]
    log("learned: %d pairs, %d Adam steps in %.2f s, pair loss %.3e -> %.3e" % (
        npairs, iters, t_train, sum(base) / max(len(base), 1),
        sum(after) / max(len(after), 1)))
    hip.bb_learn_collect(0)
    return {"pairs": npairs, "collection_steps": warm_steps,
            "collection_seconds": round(t_collect, 3),
            "adam_steps": iters, "lr": lr,
            "train_seconds": round(t_train, 3),
            "pair_loss_jacobi": round(sum(base) / max(len(base), 1), 6),
            "pair_loss_trained": round(sum(after) / max(len(after), 1), 6),
            "last_tenth_loss": round(loss, 6)}


def time_network(s, reps, runtime):
    """HIP-event times of the network's application on this grid (random
    input, all-fluid flags): the whole application and its full-resolution
    8->8 convolutions (two per application)."""
    import ctypes as C
    import numpy as np
    hip = s.hip
    d = s.dom
    n = d.xn * d.yn * d.zn
    r = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    fl = np.full(n, 63, np.uint8)
    dr = hip.bb_device_alloc(r.nbytes)
    df = hip.bb_device_alloc(fl.nbytes)
    do = hip.bb_device_alloc(r.nbytes)
    hip.bb_copy_to_device(dr, r.ctypes.data, r.nbytes)
    hip.bb_copy_to_device(df, fl.ctypes.data, fl.nbytes)
    hip.bb_unet_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    hip.bb_unet_apply(dr, df, do)
    hip.bb_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.bb_unet_apply(dr, df, do)      # returns when done
    app_s = (time.perf_counter() - t0) / reps
    hip.bb_prof_reset()
    hip.bb_prof_enable(1)
    for _ in range(reps):
        hip.bb_unet_apply(dr, df, do)
    hip.bb_synchronize()
    hip.bb_prof_enable(0)
    ms = C.c_double()
    cnt = C.c_long()
    hip.bb_prof_read(runtime.PROF_KINDS.index("unet_conv0"), C.byref(ms),
                     C.byref(cnt))
    for p in (dr, df, do):
        hip.bb_device_free(p)
    conv_ms, conv_cnt = ms.value, cnt.value
    avg = conv_ms / max(conv_cnt, 1) * 1e-3
    flops = 2.0 * 27 * 8 * 8 * n
    nbytes = (16 + 16) * n
    return {
        "application_ms": round(app_s * 1e3, 4),
        "conv0_kernel": "k_conv3<8,8> (full resolution, BF16 MFMA)",
        "conv0_avg_us": round(avg * 1e6, 2),
        "bound": "mfma", "achieved": round(flops / avg / 1e12, 2),
        "peak": 2500.0, "unit": "TFLOP/s",
        "frac": round(flops / avg / 1e12 / 2500.0, 4),
        "flops_per_launch": flops,
        "hbm_GBs": round(nbytes / avg / 1e9, 1),
        "hbm_frac": round(nbytes / avg / 1e9 / HBM_PEAK_GBS, 4),
    }


def roofline(prof, nbytes):
    best = None
    for name, (ms, cnt) in prof.items():
        if name not in nbytes:
            continue
        if best is None or ms > best[1]:
            best = (name, ms, cnt)
    if best is None:
        return None, {}
    table = {}
    for name, (ms, cnt) in prof.items():
        if name in nbytes:
            avg = ms / cnt
            gbs = nbytes[name] / (avg * 1e-3) / 1e9
            table[name] = {"launches": cnt, "avg_us": round(avg * 1e3, 2),
                           "total_ms": round(ms, 2),
                           "GB/s": round(gbs, 1),
                           "frac": round(gbs / HBM_PEAK_GBS, 3)}
    name, ms, cnt = best
    avg = ms / cnt
    achieved = nbytes[name] / (avg * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_us": round(avg * 1e3, 2),
            "bytes_per_launch": nbytes[name]}, table


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE,
    calibrated on the library's copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "sediment256_hbm_traffic.json")
    if workload != "sediment" or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get("bytes_per_launch", {}).get(kernel)
    return rec["total"] if rec else None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(s, fp64, krylov_sample, desc):
    """Serial FP64 oracle on this host, bounded sample, extrapolated."""
    import ctypes as C
    from oracle.pyoracle import OracleState
    s.pull()
    o = OracleState.from_sim(s)
    lib = o.lib
    lib.orc_set_threads(1)
    nparts = o.nparts

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    # one Lamb sub-iteration without the solve
    stage = 0.0
    stage += timed(lambda: o.call("U_star_2"))
    if nparts:
        stage += timed(lambda: o.call("part_BC_star"))
    stage += timed(lambda: o.call("dom_BC_star"))
    stage += timed(lambda: o.call("solvability"))
    if nparts:
        stage += timed(lambda: o.call("part_BC_star"))
    stage += timed(lambda: o.call("dom_BC_star"))
    # rhs (and the particle mask), then a fixed number of iterations
    stage += timed(lambda: o.call("PP_rhs"))
    if nparts:
        stage += timed(lambda: o.call("part_BC_p"))
    n = o.s.dom.Gcc.s3
    import numpy as np
    b = np.array(o.array("rhs_p")[:n])
    nrm = float(np.sqrt(np.dot(b, b))) or 1.0
    b /= nrm
    x = np.zeros(n)
    resid = C.c_double()
    conv = C.c_int()
    t_solve = timed(lambda: lib.orc_bicgstab(
        o.ptr, b.ctypes.data_as(C.POINTER(C.c_double)),
        x.ctypes.data_as(C.POINTER(C.c_double)), krylov_sample, 0.0,
        C.byref(resid), C.byref(conv)))
    per_iter = t_solve / krylov_sample
    stage += timed(lambda: o.call("dom_BC_phi"))
    stage += timed(lambda: o.call("project"))
    if nparts:
        stage += timed(lambda: o.call("part_BC"))
    stage += timed(lambda: o.call("dom_BC"))
    stage += timed(lambda: o.call("update_p"))
    if nparts:
        stage += timed(lambda: o.call("part_BC"))
        stage += timed(lambda: o.call("part_p_fill"))
    stage += timed(lambda: o.call("dom_BC_p"))
    if nparts:
        stage += timed(lambda: o.call("move_parts_sub"))
        stage += timed(lambda: o.call("Lamb"))
        stage += timed(lambda: lib.orc_lamb_err(o.ptr))
    outer = 0.0
    outer += timed(lambda: o.call("build_cages"))
    if nparts:
        outer += timed(lambda: o.call("move_parts"))
        outer += timed(lambda: o.call("parts_internal"))
    outer += timed(lambda: o.call("dom_BC"))
    outer += timed(lambda: o.call("store_u"))
    if nparts:
        outer += timed(lambda: o.call("store_coeffs"))
    outer += timed(lambda: lib.orc_find_dt(o.ptr))
    li = fp64["lamb_iters_per_step"]
    ki = fp64["krylov_iters_per_solve"]
    step_s = outer + li * (stage + ki * per_iter)
    sample_s = stage + outer + t_solve
    del o
    return {
        "value": round(1.0 / step_s, 6), "unit": "timesteps/s", "cores": 1,
        "kind": "port", "host_nproc": os.cpu_count(),
        "host_cpu": cpu_model(),
        "sample": ("serial FP64 oracle on the full state of the workload (%s): "
                   "one pass of each sub-iteration stage (%.1f s), the step's "
                   "outer stages (%.1f s) and %d BiCGSTAB iterations "
                   "(%.2f s/iteration), %.1f s measured in all; extrapolated "
                   "to %.2f Lamb sub-iterations x %.1f Krylov iterations per "
                   "step as the GPU FP64 run measured" % (
                       desc, stage, outer, krylov_sample, per_iter, sample_s,
                       li, ki)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nparts", type=int, default=1000)
    ap.add_argument("--workload", choices=("sediment", "tg", "channel"),
                    default="sediment",
                    help="sediment: 256^3 with --nparts spheres (the metric's "
                         "configuration); tg: the same box as a sphere-free "
                         "Taylor-Green flow; channel: the dense sheared "
                         "channel, 384x192x384 with 8000 spheres")
    ap.add_argument("--modes", default="fp64,mixed,learned")
    ap.add_argument("--learn-collect-steps", type=int, default=3,
                    help="mixed steps whose inner solves supply the training "
                         "pairs of the learned form")
    ap.add_argument("--learn-iters", type=int, default=400)
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--prof-steps", type=int, default=1,
                    help="steps after the timed ones with per-launch event "
                         "timing (the kernel table and the roofline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--krylov-sample", type=int, default=10)
    ap.add_argument("--mixed-inner-tol", type=float, default=0.0,
                    help="relative tolerance of each FP32 inner solve of the "
                         "mixed form (default: the library's)")
    ap.add_argument("--inner-sweeps", type=int, default=3, choices=(2, 3),
                    help="sweeps per iteration of the FP32 inner BiCGSTAB of "
                         "the mixed and learned forms")
    ap.add_argument("--poll", type=int, default=0,
                    help="BiCGSTAB iterations between host convergence checks "
                         "(default: the library's)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the library's plain HBM copy of BYTES "
                         "(10 launches) and print its rate: the reference for "
                         "calibrating FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not "
              "supported (the multi-GPU decomposition is a design, not an "
              "implementation)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from bluebottle_amd import configs, runtime, sim
    if args.calibrate_copy:
        _, hip = runtime.load()
        gbs = hip.bb_copy_bandwidth(args.calibrate_copy, 10)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "launches": 10,
                          "GB/s": round(gbs, 1)}), flush=True)
        return
    workload = args.workload
    if workload == "sediment":
        case = configs.config3(n=args.n, nparts=args.nparts)
        desc = ("sedimentation %dx%dx%d periodic, %d spheres r=6dx, "
                "rho_p/rho_f=2, Re_p=10, L=4, 26 nodes, Poisson rel tol 1e-8"
                % (args.n, args.n, args.n, args.nparts))
    elif workload == "channel":
        case = configs.config5()
        desc = ("dense sheared channel 384x192x384, y walls at -U/2 and "
                "+U/2, periodic x and z, Re_bulk=2000, %d spheres r=4dx, "
                "rho_p/rho_f=1.5, L=4, Poisson rel tol 1e-8"
                % len(case.particles))
    else:
        case = configs.taylor_green(n=args.n, pp_residual=1e-8)
        desc = ("Taylor-Green %dx%dx%d periodic unit box, Re=100, no spheres, "
                "Poisson rel tol 1e-8"
                % (args.n, args.n, args.n))
    xn, yn, zn = case.n
    ncell = xn * yn * zn
    nbytes = kernel_bytes(ncell, xn, yn, zn)
    modes = {}
    names = {"fp64": 0, "mixed": 1, "learned": 2}
    learn = None
    net_roof = None
    cpu = None
    copy_gbs = None
    for m in args.modes.split(","):
        # every solver form starts from the same freshly set-up state
        log("setting up %s" % desc)
        t0 = time.perf_counter()
        s = sim.Simulation(case)
        if args.mixed_inner_tol > 0:
            s.hip.bb_set_mixed_params(args.mixed_inner_tol, 0, 0)
        if args.poll > 0:
            s.hip.bb_set_poll_interval(args.poll)
        s.hip.bb_set_inner_sweeps(args.inner_sweeps)
        s.to_device()
        s.prepare()
        s.hip.bb_synchronize()
        log("set-up done in %.1f s on %s" % (time.perf_counter() - t0,
                                            s.hip.bb_device_name().decode()))
        if copy_gbs is None:
            # the achievable streaming rate on this device: a plain 1 GiB copy
            copy_gbs = s.hip.bb_copy_bandwidth(1 << 30, 10)
            log("plain HBM copy: %.0f GB/s" % copy_gbs)
        if m == "learned":
            learn = prepare_learned(s, args.learn_collect_steps,
                                    args.learn_iters, args.learn_lr, 32)
            s.hip.bb_learn_reset_stats()
        modes[m] = run_mode(s, names[m], args.steps, args.warmup,
                            args.prof_steps, runtime)
        if m == "learned":
            import ctypes as C
            st = [C.c_long() for _ in range(5)]
            ts = C.c_double()
            s.hip.bb_learn_stats(*[C.byref(x) for x in st], C.byref(ts))
            learn.update({"network_applications": st[0].value,
                          "fallbacks": st[1].value,
                          "flexible_iters": st[2].value,
                          "jacobi_iters_after_fallback": st[3].value,
                          "inner_solves_skipping_network": st[4].value})
            modes[m]["learned"] = learn
            net_roof = time_network(s, 10, runtime)
            log("network: %.3f ms per application, conv0 %.1f us "
                "(%.1f TFLOP/s, %.0f GB/s)" % (
                    net_roof["application_ms"], net_roof["conv0_avg_us"],
                    net_roof["achieved"], net_roof["hbm_GBs"]))
        log("%s: %.4f timesteps/s, %.2f Lamb it/step, %.1f Krylov it/solve" % (
            m, modes[m]["timesteps_per_s"], modes[m]["lamb_iters_per_step"],
            modes[m]["krylov_iters_per_solve"]))
        if m == "fp64" and not args.no_cpu:
            log("timing the CPU oracle sample")
            cpu = cpu_baseline(s, modes["fp64"], args.krylov_sample, desc)
            log("cpu sample done")
    head = ("learned" if "learned" in modes else
            "mixed" if "mixed" in modes else list(modes)[0])
    H = modes[head]
    roof, table = roofline(H["prof"], nbytes)
    if roof is not None and args.n == 256 and args.nparts == 1000:
        roof["traffic"] = measured_traffic(roof["kernel"], workload)
        if roof["traffic"] is not None:
            roof["traffic_source"] = "profiles/sediment256_hbm_traffic.md"
    out = {
        "metric": METRIC,
        "value": round(H["timesteps_per_s"], 5),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"learned": "f64+f32+bf16", "mixed": "f64+f32"}.get(head,
                                                                   "f64"),
        "data": "synthetic",
        "config": {
            "workload": desc,
            "solver": {
                "learned": "FP64 refinement around FP32 flexible BiCGSTAB "
                           "preconditioned by the in-situ trained U-Net, "
                           "Jacobi fallback",
                "mixed": "FP64 refinement around FP32 Jacobi BiCGSTAB",
                "fp64": "FP64 Jacobi BiCGSTAB"}[head],
            "lamb_max_iter": case.lamb_max_iter,
        },
        "modes": {m: {k: (round(v, 5) if isinstance(v, float) else v)
                      for k, v in d.items() if k != "prof"}
                  for m, d in modes.items()},
        "roofline": roof,
        "roofline_mfma": net_roof,
        "hbm_copy_GBs": round(copy_gbs, 1),
        "kernels": table,
        "cpu_baseline": cpu,
    }
    for m in ("mixed", "learned"):
        if "fp64" in modes and m in modes:
            out[m + "_over_fp64"] = round(
                modes[m]["timesteps_per_s"] / modes["fp64"]["timesteps_per_s"],
                3)
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
