"""Throughput of the plain FP64 RHS + SSP-RK3 path on one MI355X.

Workload: decaying compressible isotropic turbulence, periodic [0, 2pi]^3,
256^3 cells, gamma = 1.4, Pr = 0.72, Sutherland mu, Re = 100, Ma_t = 0.6,
WCNS5 interpolation + HLLC flux, CFL 0.5, seeded synthetic initial state.
One step = record U0 + CFL time step + three RK3 stages, entirely on the
device (no host synchronisation inside the timed region).

Prints one JSON line; see --help.  --gpus above 1 is rejected: this build
runs one GPU per process and measures one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-RK3-steps/s @256^3: FP64 plain vs mixed-FP32 at stated err; adjoint/fwd"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# compulsory FP64 state traffic per cell and RK3 step: stage 1 reads U^n and
# writes U^(1) (80 B); stages 2 and 3 read U^(s), U^n and write U^(s+1) (120 B each)
B_ALG_STEP = 320.0
PROFILE_DIR = os.path.join(ROOT, "profiles")


def kernel_flops_per_cell(f, cls):
    """Algorithmic FP operations per cell of one launch of a kernel class."""
    if cls == "pencil_x":
        return f["face"] + f["diff"] / 3 + f["visc_div"] / 3 + f["update"]
    if cls in ("pencil_y", "pencil_z"):
        return f["face"] + f["diff"] / 3 + f["visc_div"] / 3
    if cls == "viscous":
        return f["visc_flux"]
    if cls == "cfl":
        return f["cfl"]
    return 0.0


def profiled_traffic(cls, n):
    """HBM bytes per launch of a kernel class at n^3 from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, calibrated on a plain
    copy kernel), or None."""
    path = os.path.join(PROFILE_DIR, f"{cls}_wcns5_hllc_{n}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", default="wcns5", help="wcns5 | weno5_js | mdcd | mdcd_linear")
    ap.add_argument("--riemann", default="HLLC")
    ap.add_argument("--mach-t", type=float, default=0.6)
    ap.add_argument("--mixed", type=int, default=1, help="also time the mixed method and report its error")
    ap.add_argument("--err-steps", type=int, default=20,
                    help="steps of the mixed-vs-FP64 error comparison from the same state")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU checker on a bounded sample")
    ap.add_argument("--cpu-threads", type=int, default=16)
    return ap.parse_args()


def cpu_baseline(prim, grid, params, threads):
    """One full RK3 step of the same workload through the CPU checker (the
    restatement of the reference's pass structure), on `threads` host cores."""
    from oracle.ns3d_oracle import OracleField, set_threads

    used = set_threads(threads)
    ref = OracleField(grid, params)
    ref.initialize_primitive(*prim)
    t0 = time.perf_counter()
    dt = ref.compute_timestep()
    ref.record_conserved_to0()
    ref.runge_kutta_3(dt)
    el = time.perf_counter() - t0
    cells = grid.global_nx * grid.global_ny * grid.global_nz
    ref.close()
    return {"value": cells / el / 1e6, "unit": "Mcell-RK3-steps/s", "cores": used, "kind": "port",
            "sample": f"1 RK3 step (CFL + 3 stages) of the {grid.global_nx}^3 workload, {el:.2f} s wall"}


def mixed_report(args, prim, G, P, ms_plain, timed_run):
    """Throughput of the mixed method on the same workload, and its error
    against the plain FP64 method after --err-steps steps from the same
    state (per conserved field, kinetic energy, TKE, dissipation, E(k) per
    shell)."""
    from ns3d_wcns_amd import PRECISION_FP64, PRECISION_MIXED, DeviceField
    from ns3d_wcns_amd.observables import compare_runs

    ms, kt, _, finite = timed_run(PRECISION_MIXED)
    cells = G.global_nx * G.global_ny * G.global_nz
    snaps = []
    for prec in (PRECISION_MIXED, PRECISION_FP64):
        d = DeviceField(G, P, precision=prec)
        d.initialize_primitive(*prim)
        d.advance_steps(args.err_steps)
        snaps.append(d.snapshot())
        d.close()
    rep = compare_runs(snaps[0], snaps[1], P, G.dx * G.dy * G.dz)
    return {
        "value": round(cells / (ms * 1e-3) / 1e6, 2),
        "ms_per_step": round(ms, 4),
        "speedup_vs_fp64": round(ms_plain / ms, 3),
        "dtype": "f32 face/viscous arithmetic, f64 state and accumulation",
        "kernel_ms_per_launch": {k: round(t / c, 4) if c else 0.0 for k, (t, c) in kt.items()},
        "state_finite": finite,
        "error_vs_fp64": {
            "steps": args.err_steps,
            "fields": {k: float(f"{v:.3e}") for k, v in rep["fields"].items()},
            "kinetic_energy": float(f"{rep['kinetic_energy']:.3e}"),
            "tke": float(f"{rep['tke']:.3e}"),
            "dissipation": float(f"{rep['dissipation']:.3e}"),
            "spectrum_max_shell": float(f"{rep['spectrum_max']:.3e}"),
            "spectrum_per_shell": [float(f"{e:.2e}") for e in rep["spectrum_per_shell"]],
        },
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from ns3d_wcns_amd import PRECISION_FP64, DeviceField, RiemannSolver, SCHEME_ALIASES, SolverParams, periodic_box
    from ns3d_wcns_amd._lib import algorithmic_flops
    from ns3d_wcns_amd.isotropic import isotropic_state

    torch.cuda.set_device(0)
    n = args.n
    prim, ma = isotropic_state(n, mach_t=args.mach_t, seed=12345)
    P = SolverParams(Ma=ma, Re=100.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES[args.scheme],
                     riemann_solver=RiemannSolver[args.riemann]).derive()
    G = periodic_box(n)
    stream = torch.cuda.current_stream()
    cells = n ** 3

    def timed_run(precision):
        """W untimed + K timed steps of one method from the initial state;
        returns (ms per step, per-class kernel times, wall s, state finite)."""
        dev = DeviceField(G, P, precision=precision)
        dev.set_stream(stream.cuda_stream)
        dev.initialize_primitive(*prim)
        dts = torch.zeros(args.warmup + args.steps, dtype=torch.float64, device="cuda")
        base = dts.data_ptr()
        for s in range(args.warmup):
            dev.advance_steps(1, base + 8 * s)
        torch.cuda.synchronize()
        # HIP events bracket every kernel launch of the timed steps (on the
        # library's stream, which is this stream)
        dev.kernel_times()
        dev.set_kernel_timing(True)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record(stream)
        dev.advance_steps(args.steps, base + 8 * args.warmup)
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        dev.set_kernel_timing(False)
        kt = dev.kernel_times()
        finite = bool(np.isfinite(dev.interior(dev.download("rho"))).all())
        dev.close()
        return start.elapsed_time(stop) / args.steps, kt, wall, finite

    ms, ktimes, wall, finite = timed_run(PRECISION_FP64)
    value = cells / (ms * 1e-3) / 1e6
    flops = algorithmic_flops(P)
    per_launch = {k: (t / c if c else 0.0) for k, (t, c) in ktimes.items()}
    dom = max(("pencil_x", "pencil_y", "pencil_z", "viscous"), key=lambda k: ktimes[k][0])
    fl = kernel_flops_per_cell(flops, dom) * cells
    achieved = fl / (per_launch[dom] * 1e-3) / 1e12
    traffic = profiled_traffic(dom, n)
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
        "traffic": traffic,
        "kernel": dom,
        "kernel_ms": round(per_launch[dom], 4),
        "algorithmic_flops_per_launch": fl,
        "compute": "FP64 vector FMA/add/mul (v_fma_f64); no MFMA instruction is issued; the MI355X FP64 "
                   "vector and FP64 matrix peaks are both 78.6 TFLOP/s",
        "hbm_GBps_measured_traffic": round(traffic / (per_launch[dom] * 1e-3) / 1e9, 1) if traffic else None,
    }
    step_tflops = flops["per_step"] * cells / (ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"isotropic turbulence {n}^3 periodic, Ma_t={args.mach_t}, Re=100, "
                               f"{args.scheme.upper()}+{args.riemann}, CFL 0.5, plain FP64",
                   "cells": cells, "parallelism": "1 GPU"},
        "roofline": roofline,
        "step_roofline": {
            "algorithmic_flops_per_cell_step": flops["per_step"],
            "fp64_TFLOPs": round(step_tflops, 3),
            "fp64_frac": round(step_tflops / FP64_PEAK_TFLOPS, 4),
            "algorithmic_bytes_per_cell_step": B_ALG_STEP,
            "hbm_frac_algorithmic": round(B_ALG_STEP * cells / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        },
        "kernel_ms_per_launch": {k: round(v, 4) for k, v in per_launch.items()},
        "wall_s": round(wall, 3),
        "state_finite": finite,
    }
    if args.mixed:
        out["mixed"] = mixed_report(args, prim, G, P, ms, timed_run)
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prim, G, P, args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
