"""Capture kernel launches of one case from the host emulation build
(test infrastructure): every launch of the first few calls of each kernel
is written as launch shape, kernarg bytes and the device memory image
before and after, for replay against the gfx950 machine code (replay.py).

    python tests/isa/capture.py <case> <outdir> [steps] [fp64|mixed] [implicit]

Cases are the small parity cases of tests/test_gpu_parity.py (and
`precursor` / `precursor_x`, the cases of tests/test_gpu_precursor.py:
planes pushed, cuda_dom_BC, cuda_dom_BC_star, cuda_yank_turb_planes; and
`turb`, the precursor domain: set-up, one step with its linear forcing,
cuda_compute_energy and the three colocations, or `turb_forcing`, the
same without the set-up and step; `particle_kernels`, three spheres with
a pair 0.3 r apart: the cages and face lists of the set-up, one
cuda_move_parts_sub (bins) and cuda_Lamb then cuda_lamb_err); the optional
fourth argument selects the Poisson solver form and `implicit` the
Crank-Nicolson (Helmholtz) predictor."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    case_name, out = sys.argv[1], sys.argv[2]
    steps = int(sys.argv[3]) if len(sys.argv) > 3 else 1
    mode = {"fp64": 0, "mixed": 1}[sys.argv[4] if len(sys.argv) > 4 else "fp64"]
    implicit = len(sys.argv) > 5 and sys.argv[5] == "implicit"
    os.makedirs(out, exist_ok=True)
    os.environ["BB_EMU"] = "1"
    os.environ["BB_EMU_ISA_DUMP"] = out
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import conftest  # noqa: F401  (points the loader at the emulation build)
    from test_gpu_parity import CASES
    import dataclasses
    from bluebottle_amd import runtime, sim
    if case_name in ("precursor", "precursor_x"):
        # the PRECURSOR ghost fill and the plane extraction on random
        # fields and planes (tests/test_gpu_precursor.py's cases)
        import ctypes as C
        import numpy as np
        from test_gpu_precursor import CASES as PCASES
        from test_oracle_precursor import configs18
        case = PCASES["all" if case_name == "precursor" else "inflow_x"]()
        s = sim.Simulation(case)
        rng = np.random.default_rng(5)
        for nm in ("u", "v", "w", "u_star", "v_star", "w_star", "p"):
            a = s.array(nm)
            a[:] = rng.standard_normal(a.size)
        for nm in s.PLANES:
            a = s.array(nm)
            a[:] = rng.standard_normal(a.size)
        s.to_device()
        cfg = configs18(case)
        s.hip.cuda_dom_turb_planes_push(cfg)
        s.hip.cuda_dom_BC()
        s.hip.cuda_dom_BC_star()
        d = s.dom
        pos = (C.c_double * 9)(d.xs + 0.37 * d.xl, 0, 0, 0,
                               d.ys + 0.61 * d.yl, 0, 0, 0,
                               d.zs + 0.43 * d.zl)
        vel = (C.c_double * 18)(*[0.01 * (q + 1) for q in range(18)])
        s.hip.cuda_yank_turb_planes(cfg, pos, vel)
        s.pull()
        s.close()
        print("captured", len([f for f in os.listdir(out) if f.endswith(".json")]),
              "launches of", case_name, flush=True)
        return
    if case_name in ("turb", "turb_forcing"):
        import ctypes as C
        from test_oracle_turb import turb_case
        case = turb_case(n=(12, 10, 8), hi=(1.0, 0.8, 0.7), turbA=0.6, seed=4)
        s = sim.Simulation(case)
        s.to_device()
        if case_name == "turb":
            s.prepare()
            s.step()
        else:
            s.hip.cuda_compute_turb_forcing()
        s.hip.cuda_compute_energy()
        host = s.host
        dst = C.POINTER(C.POINTER(C.c_double)).in_dll(host, "_phi")[0]
        ddom = C.POINTER(C.c_void_p).in_dll(host, "_dom")[0]
        for comp, fn in enumerate(("cuda_colocate_Gfx", "cuda_colocate_Gfy",
                                   "cuda_colocate_Gfz")):
            src = C.POINTER(C.POINTER(C.c_double)).in_dll(host,
                                                          "_" + "uvw"[comp])[0]
            getattr(s.hip, fn)(C.cast(src, C.c_void_p), C.cast(dst, C.c_void_p),
                               ddom)
        s.pull()
        s.close()
        print("captured", len([f for f in os.listdir(out) if f.endswith(".json")]),
              "launches of", case_name, flush=True)
        return
    if case_name == "particle_kernels":
        a = 0.15
        parts = [sim.Particle(x=0.3, y=0.5, z=0.5, r=a),
                 sim.Particle(x=0.3 + 2.3 * a, y=0.5, z=0.5, r=a),
                 sim.Particle(x=0.75, y=0.2, z=0.8, r=a)]
        case = sim.Case(n=(12, 12, 12), nu=0.01, init="taylor_green", seed=3,
                        noise=1e-3, particles=parts, lamb_max_iter=1)
        s = sim.Simulation(case)
        s.to_device()
        s.prepare()
        s.hip.cuda_move_parts_sub()
        s.hip.cuda_Lamb()
        s.hip.cuda_lamb_err()
        s.pull()
        s.close()
        print("captured", len([f for f in os.listdir(out) if f.endswith(".json")]),
              "launches of", case_name, flush=True)
        return
    if case_name == "copy":
        # the HBM copy reference kernel: 64 KiB, two launches
        _, hip = runtime.load()
        hip.bb_copy_bandwidth(1 << 16, 2)
        print("captured", len([f for f in os.listdir(out) if f.endswith(".json")]),
              "launches of", case_name, flush=True)
        return
    extra = {
        # 16 x 64 x 16: eight SpMV tiles and 16 element-wise tiles, so
        # the XCD block remap of the Krylov sweeps takes its permuting
        # branch (as at 256^3)
        "tg_remap": lambda: sim.Case(n=(16, 64, 16), nu=0.01, CFL=0.3,
                                     pp_max_iter=500, pp_residual=1e-11),
    }
    case = (extra.get(case_name) or CASES[case_name])()
    if implicit:
        case = dataclasses.replace(case, implicit=True)
    s = sim.Simulation(case)
    s.to_device()
    s.set_solver(mode)
    if os.environ.get("BB_CAPTURE_INNER_SWEEPS"):
        # the mixed form's inner solve in two sweeps per iteration
        s.hip.bb_set_inner_sweeps(int(os.environ["BB_CAPTURE_INNER_SWEEPS"]))
    s.prepare()
    if os.environ.get("BB_CAPTURE_HELPERS") == "1":
        # the stand-alone entry points the step does not call
        import ctypes as C
        import numpy as np
        from oracle import pyoracle
        lib = pyoracle.load()
        th, ph, wt = np.zeros(26), np.zeros(26), np.zeros(26)
        P = C.POINTER(C.c_double)
        lib.orc_lebedev_nodes(th.ctypes.data_as(P), ph.ctypes.data_as(P),
                              wt.ctypes.data_as(P))
        hip = s.hip
        dt_, dp_ = hip.bb_device_alloc(208), hip.bb_device_alloc(208)
        hip.bb_copy_to_device(dt_, th.ctypes.data, 208)
        hip.bb_copy_to_device(dp_, ph.ctypes.data, 208)
        n = 26 * max(len(case.particles), 1)
        outs = [hip.bb_device_alloc(8 * n) for _ in range(4)]
        hip.cuda_quad_check_nodes(0, dt_, dp_, 26)
        hip.cuda_quad_interp(0, dt_, dp_, 26, *outs)
        hip.cuda_part_BC_p(0)
        for d in [dt_, dp_] + outs:
            hip.bb_device_free(d)
        steps = 0
    for _ in range(steps):
        s.step()
    s.pull()
    s.close()
    print("captured", len([f for f in os.listdir(out) if f.endswith(".json")]),
          "launches of", case_name, flush=True)


if __name__ == "__main__":
    main()
