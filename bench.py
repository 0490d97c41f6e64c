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


def profiled_valu(cls, n):
    """VALU-busy percentage, VALU issue fraction and counted FP64 TFLOP/s of
    one launch of a kernel class at n^3 from the committed rocprofv3 --pmc
    summary (SQ_ACTIVE_INST_VALU, SQ_INSTS_VALU, SQ_INSTS_VALU_*_F64), or {}."""
    path = os.path.join(PROFILE_DIR, f"{cls}_valu_{n}.md")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        for key in ("valu_busy_pct", "valu_issue", "fp64_tflops_counted"):
            if line.startswith(key + ":"):
                out[key] = float(line.split(":", 1)[1])
    return out


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
    ap.add_argument("--adjoint", type=int, default=1,
                    help="also run the adjoint workload (128^3, 50 steps, checkpoint every 10, FD check)")
    ap.add_argument("--adj-n", type=int, default=128)
    ap.add_argument("--adj-steps", type=int, default=50)
    ap.add_argument("--sweep", type=int, default=1,
                    help="also time {WCNS5, WENO5-JS, MDCD} x {HLLC, Rusanov} on the same workload")
    ap.add_argument("--sweep-steps", type=int, default=5)
    ap.add_argument("--scale-mixed", type=int, default=0,
                    help="n of the at-scale mixed-vs-plain run (0: off; 512: Ma_t 0.3, Re 1600, 500 steps)")
    ap.add_argument("--scale-mixed-steps", type=int, default=500)
    ap.add_argument("--scale-adjoint", type=int, default=0,
                    help="n of the at-scale adjoint run (0: off; 512: Ma_t 0.5, Re 1600, 200 steps, mixed "
                         "forward, FP64 adjoint, J = time-integrated dissipation)")
    ap.add_argument("--scale-adjoint-steps", type=int, default=200)
    ap.add_argument("--scale-adjoint-snapshots", type=int, default=12,
                    help="binomial checkpoint snapshots (12 x 5.6 GB at 512^3)")
    ap.add_argument("--scale-adjoint-fd", type=int, default=4, help="FD directions")
    ap.add_argument("--scale-adjoint-fp64", type=int, default=1,
                    help="also the FP64-forward gradient of the same objective, for comparison")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU checker on a bounded sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the multi-core CPU baseline (0: the box's CPU share, OMP_NUM_THREADS or 16, "
                         "at most os.cpu_count())")
    ap.add_argument("--cpu-1core-n", type=int, default=128,
                    help="box of the single-core CPU baseline (the same workload on a smaller box)")
    return ap.parse_args()


def cpu_threads(requested):
    if requested > 0:
        return requested
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(share, os.cpu_count() or 1))


def cpu_baseline(prim, grid, params, threads, what="workload"):
    """One full RK3 step (CFL + 3 stages) through the CPU checker (the
    restatement of the reference's pass structure, which the reference runs
    on one MPI rank per core), on `threads` host threads."""
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
    return {"value": round(cells / el / 1e6, 4), "unit": "Mcell-RK3-steps/s", "cores": used, "kind": "port",
            "sample": f"1 RK3 step (CFL + 3 stages) of the {grid.global_nx}^3 {what}, {el:.2f} s wall"}


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
            "spectrum_max_shell_over": "shells k >= 1 holding at least 1e-6 of the largest shell energy",
            "spectrum_per_shell": [float(f"{e:.2e}") for e in rep["spectrum_per_shell"]],
        },
    }


def sweep_report(args, prim, ma, G, stream):
    """Throughput of each {WCNS5, WENO5-JS, MDCD hybrid} x {HLLC, Rusanov}
    combination on the same state (FP64, --sweep-steps timed steps after one
    warm-up step)."""
    import torch

    from ns3d_wcns_amd import DeviceField, RiemannSolver, SCHEME_ALIASES, SolverParams

    cells = G.global_nx * G.global_ny * G.global_nz
    res = {}
    for scheme in ("wcns5", "weno5_js", "mdcd"):
        for rs in ("HLLC", "Rusanov"):
            P = SolverParams(Ma=ma, Re=100.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES[scheme],
                             riemann_solver=RiemannSolver[rs]).derive()
            dev = DeviceField(G, P)
            dev.set_stream(stream.cuda_stream)
            dev.initialize_primitive(*prim)
            dev.advance_steps(1)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            dev.advance_steps(args.sweep_steps)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.sweep_steps
            dev.close()
            res[f"{scheme.upper()}+{rs}"] = {"ms_per_step": round(ms, 4), "Mcell_steps_per_s": round(cells / ms / 1e3, 1)}
    return res


def adjoint_report(args):
    """Discrete adjoint of an isotropic-turbulence run: --adj-n^3, Ma_t = 0.3,
    Re = 100, WCNS5 + HLLC, --adj-steps RK3 steps at the CFL-0.5 time steps of
    the forward run, checkpoint every 10 steps.  J = mean 0.5 rho |u|^2 at the
    final time; gradient with respect to the initial conserved fields.
    Reports the backward/forward cost ratio of the whole sweep (recomputation
    included; the uniform policy keeps the stage states of the recomputed
    segment) and per step (adjoint step with its stages recomputed, and
    with them stored), central-difference checks on 8 seeded random
    directions in (rho u, rho v, rho w, E) (unit Euclidean norm at eps =
    1e-6 and 1e-4, unit RMS per entry at eps = 1e-6), and the
    tangent/adjoint dot test of one step."""
    import torch

    from ns3d_wcns_amd import DeviceField, RiemannSolver, SCHEME_ALIASES, SolverParams, periodic_box
    from ns3d_wcns_amd.adjoint import central_difference, mean_kinetic_energy, rollout, state_tensor
    from ns3d_wcns_amd.isotropic import isotropic_state

    n, nsteps = args.adj_n, args.adj_steps
    prim, ma = isotropic_state(n, mach_t=0.3, seed=2024)
    P = SolverParams(Ma=ma, Re=100.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES["wcns5"],
                     riemann_solver=RiemannSolver.HLLC).derive()
    f = DeviceField(periodic_box(n), P)
    f.set_stream(torch.cuda.current_stream().cuda_stream)
    f.initialize_primitive(*prim)
    U = state_tensor(f)
    dts = []
    for _ in range(nsteps):
        dts.append(f.compute_timestep())
        f.step(dts[-1])
    policy = "uniform:10"
    # warm-up of both sweeps, then the measured one
    U0 = U.clone().requires_grad_(True)
    mean_kinetic_energy(rollout(U0, f, dts[:2], policy)).backward()
    stats = {}
    U0 = U.clone().requires_grad_(True)
    J = mean_kinetic_energy(rollout(U0, f, dts, policy, stats))
    J.backward()
    grad = U0.grad
    # per-step costs: forward step, adjoint step (state loaded, stages recomputed inside)
    adj = torch.zeros_like(U)
    adj[:, 3:-3, 3:-3, 3:-3] = 1e-3
    ptr = [adj[q].data_ptr() for q in range(5)]
    Up = [U[q].data_ptr() for q in range(5)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    f.set_state(Up)
    torch.cuda.synchronize()
    e0.record()
    for i in range(10):
        f.step(dts[i])
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 10
    f.set_state(Up)
    torch.cuda.synchronize()
    e0.record()
    for i in range(10):
        f.adjoint_step(dts[i], ptr)
    e1.record()
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1) / 10
    # the same with the stage states stored by the recomputation (what the
    # uniform checkpoint sweep does)
    S1, S2 = torch.empty_like(U), torch.empty_like(U)
    f.set_state(Up)
    f.step_stages(dts[0], [S1[q].data_ptr() for q in range(5)], [S2[q].data_ptr() for q in range(5)])
    torch.cuda.synchronize()
    e0.record()
    for i in range(10):
        f.adjoint_step_stages(dts[0], ptr, [S1[q].data_ptr() for q in range(5)], [S2[q].data_ptr() for q in range(5)])
    e1.record()
    torch.cuda.synchronize()
    adj_stored_ms = e0.elapsed_time(e1) / 10
    del S1, S2
    # central differences along 8 seeded random directions in (rho u, rho v, rho w, E): of unit Euclidean
    # norm over all entries (the configuration's statement) at eps = 1e-6 and 1e-4, and of unit RMS per
    # entry at eps = 1e-6 (a unit-norm direction is nearly orthogonal to the gradient and perturbs each
    # entry by ~3.5e-10 eps, so at eps = 1e-6 its difference sits on the rounding floor of the perturbed
    # trajectories, ~1e-16 J / eps; unit RMS scales the derivative by sqrt(entries))
    g = torch.Generator(device="cpu").manual_seed(7)
    gnorm = float(torch.linalg.vector_norm(grad[1:, 3:-3, 3:-3, 3:-3]))
    entries = 4 * n ** 3
    rel6, rel4, rms6, cs6 = [], [], [], []
    for k in range(8):
        d = torch.zeros_like(U)
        d[1:, 3:-3, 3:-3, 3:-3] = torch.randn((4, n, n, n), generator=g, dtype=torch.float64).cuda()
        d /= torch.linalg.vector_norm(d)
        ad = float((grad * d).sum())
        with torch.no_grad():
            fd6 = float(central_difference(f, U, d, dts, 1e-6))
            fd4 = float(central_difference(f, U, d, dts, 1e-4))
            dr = d * math.sqrt(entries)
            fdr = float(central_difference(f, U, dr, dts, 1e-6))
        adr = ad * math.sqrt(entries)
        rel6.append(float(f"{abs(fd6 - ad) / abs(ad):.2e}"))
        cs6.append(abs(fd6 - ad) / gnorm)
        rel4.append(float(f"{abs(fd4 - ad) / abs(ad):.2e}"))
        rms6.append(float(f"{abs(fdr - adr) / abs(adr):.2e}"))
    # dot test of one step
    gg = torch.Generator(device="cpu").manual_seed(8)
    dU = torch.zeros_like(U)
    lam = torch.zeros_like(U)
    dU[:, 3:-3, 3:-3, 3:-3] = 1e-6 * torch.randn((5, n, n, n), generator=gg, dtype=torch.float64).cuda()
    lam[:, 3:-3, 3:-3, 3:-3] = torch.randn((5, n, n, n), generator=gg, dtype=torch.float64).cuda()
    tl, ad = dU.clone(), lam.clone()
    f.set_state(Up)
    f.tangent_step(dts[0], [tl[q].data_ptr() for q in range(5)])
    f.set_state(Up)
    f.adjoint_step(dts[0], [ad[q].data_ptr() for q in range(5)])
    torch.cuda.synchronize()
    a, b = float((tl * lam).sum()), float((dU * ad).sum())
    f.close()
    return {
        "workload": f"isotropic {n}^3, Ma_t=0.3, Re=100, WCNS5+HLLC, {nsteps} RK3 steps, J = mean 0.5 rho |u|^2",
        "checkpoint": policy,
        "J": float(J.detach()),
        "forward_sweep_s": round(stats["forward_s"], 4),
        "backward_sweep_s": round(stats["backward_s"], 4),
        "sweep_cost_ratio": round(stats["backward_s"] / stats["forward_s"], 3),
        "forward_step_ms": round(fwd_ms, 4),
        "adjoint_step_ms": round(adj_ms, 4),
        "step_cost_ratio": round(adj_ms / fwd_ms, 3),
        "adjoint_step_stored_stages_ms": round(adj_stored_ms, 4),
        "step_cost_ratio_stored_stages": round(adj_stored_ms / fwd_ms, 3),
        "recomputed_steps": stats["recomputed_steps"],
        "fd_unit_norm_eps_1e-6_rel_err": rel6,
        "fd_unit_norm_eps_1e-6_max_err_over_grad_norm": float(f"{max(cs6):.3e}"),
        "fd_unit_norm_eps_1e-4_rel_err": rel4,
        "fd_unit_rms_eps_1e-6_rel_err": rms6,
        "fd_max_rel_err": {"unit_norm_eps_1e-6": max(rel6), "unit_norm_eps_1e-4": max(rel4),
                           "unit_rms_eps_1e-6": max(rms6)},
        "dot_test_rel_err": float(f"{abs(a - b) / abs(a):.3e}"),
    }


def progress(msg):
    print(f"[bench] {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def device_run(G, P, precision, prim, nsteps, label, chunk=50):
    """nsteps CFL-controlled steps on the device; returns (field, ms per step)."""
    import torch

    from ns3d_wcns_amd import DeviceField

    f = DeviceField(G, P, precision=precision)
    stream = torch.cuda.current_stream()
    f.set_stream(stream.cuda_stream)
    f.initialize_primitive(*prim)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms, done = 0.0, 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        e0.record(stream)
        f.advance_steps(k)
        e1.record(stream)
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
        done += k
        progress(f"{label}: {done}/{nsteps} steps, {ms / done:.2f} ms/step")
    return f, ms / nsteps


def scale_mixed_report(args):
    """Mixed vs plain at scale: isotropic --scale-mixed^3, Ma_t = 0.3,
    Re = 1600, WCNS5 + HLLC, --scale-mixed-steps CFL-0.5 steps of each method
    from the same state.  Errors of the mixed run against the FP64 run, all
    on the device: norm-wise per conserved field, kinetic energy, TKE, the
    spectral dissipation and E(k) per shell (hipFFT)."""
    import torch

    from ns3d_wcns_amd import PRECISION_FP64, PRECISION_MIXED, RiemannSolver, SCHEME_ALIASES, SolverParams, periodic_box
    from ns3d_wcns_amd.adjoint import interior, state_tensor
    from ns3d_wcns_amd.isotropic import isotropic_state

    n, nsteps = args.scale_mixed, args.scale_mixed_steps
    progress(f"scale-mixed: initial state {n}^3")
    prim, ma = isotropic_state(n, mach_t=0.3, seed=12345)
    P = SolverParams(Ma=ma, Re=1600.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES["wcns5"],
                     riemann_solver=RiemannSolver.HLLC).derive()
    G = periodic_box(n)
    res = {}
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        f, ms = device_run(G, P, prec, prim, nsteps, f"scale-mixed {name}")
        res[name] = {"ms": ms, "U": state_tensor(f), "stats": f.turbulence_statistics(),
                     "spec": f.energy_spectrum(n // 2 + 1), "ke": f.total_kinetic_energy()}
        f.close()
    a, b = res["mixed"], res["fp64"]
    fields = {}
    for q, name in enumerate(("rho", "rhou", "rhov", "rhow", "E")):
        x, y = interior(a["U"])[q], interior(b["U"])[q]
        fields[name] = float(f"{float((x - y).abs().max() / y.abs().max()):.3e}")
    rel = lambda x, y: float(f"{abs(x - y) / abs(y):.3e}")  # noqa: E731
    ea, eb = a["spec"], b["spec"]
    shells = [float(f"{abs(ea[k] - eb[k]) / eb[k]:.2e}") if eb[k] > 0 else 0.0 for k in range(1, n // 2 + 1)]
    big = [e for k, e in enumerate(shells, start=1) if eb[k] >= 1e-6 * eb[1:].max()]
    cells = n ** 3
    return {
        "workload": f"isotropic {n}^3, Ma_t=0.3, Re=1600, WCNS5+HLLC, {nsteps} RK3 steps at CFL 0.5, each method",
        "fp64": {"ms_per_step": round(b["ms"], 3), "Mcell_steps_per_s": round(cells / b["ms"] / 1e3, 1)},
        "mixed": {"ms_per_step": round(a["ms"], 3), "Mcell_steps_per_s": round(cells / a["ms"] / 1e3, 1)},
        "speedup": round(b["ms"] / a["ms"], 3),
        "error_vs_fp64": {
            "fields": fields,
            "kinetic_energy": rel(a["ke"], b["ke"]),
            "tke": rel(a["stats"]["tke"], b["stats"]["tke"]),
            "dissipation": rel(a["stats"]["dissipation"], b["stats"]["dissipation"]),
            "spectrum_max_shell": max(big) if big else 0.0,
            "spectrum_max_shell_over": "shells k >= 1 holding at least 1e-6 of the largest shell energy",
            "spectrum_per_shell": shells,
        },
    }


def scale_adjoint_report(args):
    """Adjoint at scale: isotropic --scale-adjoint^3, Ma_t = 0.5, Re = 1600,
    WCNS5 + HLLC, --scale-adjoint-steps steps at the CFL-0.5 time steps of a
    first forward run.  Forward in the mixed FP32/FP64 method, discrete
    adjoint in FP64 about its trajectory, binomial checkpointing with
    --scale-adjoint-snapshots snapshots.  J = sum_n dt_n eps(U_{n+1}), the
    time-integrated volume-mean viscous dissipation (its gradient formed
    slab by slab); the gradient with respect to the initial velocity at
    fixed rho and p (dJ/du_c = rho dJ/d(rho u_c) + rho u_c dJ/dE), checked
    by central differences of J along --scale-adjoint-fd seeded velocity
    directions of unit RMS per cell.  The forward's FP32 rounding makes J
    noisy at about 1e-7 relative, so the differences use eps = 1e-4 (the
    first direction also 1e-3, where the O(eps^2) truncation shows); the gradient of the same objective with an
    FP64 forward is reported beside it (--scale-adjoint-fp64 1), which
    separates the adjoint's accuracy from the forward's rounding.  Costs:
    the adjoint step against the forward step (per step, events) and the
    whole backward sweep against the forward sweep."""
    import torch

    from ns3d_wcns_amd import PRECISION_MIXED, DeviceField, RiemannSolver, SCHEME_ALIASES, SolverParams, periodic_box
    from ns3d_wcns_amd.adjoint import (binomial_beta, dissipation_rate, dissipation_rate_grad, interior,
                                       state_tensor, time_integral, time_integral_value)
    from ns3d_wcns_amd.isotropic import isotropic_state

    n, nsteps, snaps = args.scale_adjoint, args.scale_adjoint_steps, args.scale_adjoint_snapshots
    progress(f"scale-adjoint: initial state {n}^3")
    prim, ma = isotropic_state(n, mach_t=0.5, seed=12345)
    P = SolverParams(Ma=ma, Re=1600.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES["wcns5"],
                     riemann_solver=RiemannSolver.HLLC).derive()
    G = periodic_box(n)
    f = DeviceField(G, P, precision=PRECISION_MIXED)
    stream = torch.cuda.current_stream()
    f.set_stream(stream.cuda_stream)
    f.initialize_primitive(*prim)
    del prim
    U = state_tensor(f)
    dts = []
    for i in range(nsteps):
        dts.append(f.compute_timestep())
        f.step(dts[-1])
        if (i + 1) % 50 == 0:
            progress(f"scale-adjoint: time steps {i + 1}/{nsteps}")
    density = lambda X: dissipation_rate(X, P, G)  # noqa: E731
    density_grad = lambda X: dissipation_rate_grad(X, P, G)  # noqa: E731
    policy = f"binomial:{snaps}"

    def gradient(field, label):
        st = {}
        U0 = U.clone().requires_grad_(True)
        progress(f"scale-adjoint: {label} forward sweep")
        J = time_integral(U0, field, dts, density, policy, st, density_grad=density_grad)
        progress(f"scale-adjoint: {label} forward {st['forward_s']:.1f} s, backward sweep")
        J.backward()
        progress(f"scale-adjoint: {label} backward {st['backward_s']:.1f} s")
        Jv, gU = float(J.detach()), U0.grad
        del J, U0
        I = interior(U)
        gv = torch.stack([I[0] * interior(gU)[1 + c] + I[1 + c] * interior(gU)[4] for c in range(3)])
        del gU
        torch.cuda.empty_cache()
        return Jv, gv, st

    J, gvel, stats = gradient(f, "mixed")
    # per-step costs
    ptr = lambda t: [t[q].data_ptr() for q in range(5)]  # noqa: E731
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    f.set_state(ptr(U))
    e0.record(stream)
    for i in range(5):
        f.step(dts[i])
    e1.record(stream)
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 5
    lam = torch.zeros_like(U)
    lam[:, 3:-3, 3:-3, 3:-3] = 1e-3
    f.set_state(ptr(U))
    e0.record(stream)
    for i in range(5):
        f.adjoint_step(dts[i], ptr(lam))
    e1.record(stream)
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1) / 5
    del lam
    progress(f"scale-adjoint: forward step {fwd_ms:.2f} ms, adjoint step {adj_ms:.2f} ms")
    # central differences along velocity directions (rho, p fixed), unit RMS per cell
    g = torch.Generator(device="cpu").manual_seed(99)
    gam = P.gamma
    I = interior(U)
    rho = I[0]
    checks = []
    for k in range(args.scale_adjoint_fd):
        d = torch.randn((3, n, n, n), generator=g, dtype=torch.float64).cuda()
        ad = float((gvel * d).sum())
        row = {"adjoint": ad}
        for eps in ((1e-4, 1e-3) if k == 0 else (1e-4,)):
            vals = []
            for sgn in (1.0, -1.0):
                V = U.clone()
                Iv = interior(V)
                vel = [I[1 + c] / rho + sgn * eps * d[c] for c in range(3)]
                p = (gam - 1.0) * (I[4] - 0.5 * (I[1] ** 2 + I[2] ** 2 + I[3] ** 2) / rho)
                for c in range(3):
                    Iv[1 + c] = rho * vel[c]
                Iv[4] = p / (gam - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
                del vel, p
                vals.append(time_integral_value(f, V, dts, density))
                del V, Iv
            fd = (vals[0] - vals[1]) / (2 * eps)
            row[f"fd_eps_{eps:g}"] = fd
            row[f"rel_err_eps_{eps:g}"] = float(f"{abs(fd - ad) / abs(ad):.3e}")
        checks.append(row)
        progress(f"scale-adjoint: FD direction {k}: {row}")
        del d
    f.close()
    torch.cuda.empty_cache()
    fp64_cmp = None
    if args.scale_adjoint_fp64:
        from ns3d_wcns_amd import PRECISION_FP64

        f64 = DeviceField(G, P, precision=PRECISION_FP64)
        f64.set_stream(stream.cuda_stream)
        J64, g64, st64 = gradient(f64, "FP64")
        f64.close()
        fp64_cmp = {
            "J_fp64_forward": J64,
            "J_rel_diff": float(f"{abs(J - J64) / abs(J64):.3e}"),
            "grad_rel_diff_norm": float(f"{float(torch.linalg.vector_norm(gvel - g64) / torch.linalg.vector_norm(g64)):.3e}"),
            "grad_rel_diff_max": float(f"{float((gvel - g64).abs().max() / g64.abs().max()):.3e}"),
            "forward_sweep_s": round(st64["forward_s"], 2),
            "backward_sweep_s": round(st64["backward_s"], 2),
        }
        del g64
    t = 1
    while binomial_beta(snaps, t) < nsteps:
        t += 1
    return {
        "workload": f"isotropic {n}^3, Ma_t=0.5, Re=1600, WCNS5+HLLC, {nsteps} RK3 steps at CFL 0.5; mixed forward, "
                    "FP64 adjoint; J = sum_n dt_n <tau:grad u>(U_n+1); gradient w.r.t. initial velocity (rho, p fixed)",
        "checkpoint": policy,
        "checkpoint_GB": round(snaps * 5 * (n + 6) ** 3 * 8 / 1e9, 1),
        "binomial_sweeps_bound": t,
        "J": J,
        "forward_step_ms": round(fwd_ms, 3),
        "adjoint_step_ms": round(adj_ms, 3),
        "step_cost_ratio": round(adj_ms / fwd_ms, 3),
        "forward_sweep_s": round(stats["forward_s"], 2),
        "backward_sweep_s": round(stats["backward_s"], 2),
        "sweep_cost_ratio": round(stats["backward_s"] / stats["forward_s"], 3),
        "recomputed_steps": stats["recomputed_steps"],
        "fd_direction": "velocity, unit RMS per cell",
        "fd": checks,
        "vs_fp64_forward": fp64_cmp,
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
        """W untimed + K timed steps of one method from the initial state,
        then K more with HIP events around every kernel launch (on the
        library's stream, which is this stream) for the per-class kernel
        times, so that the events cost the timed steps nothing; returns (ms
        per step, per-class kernel times, wall s, state finite)."""
        dev = DeviceField(G, P, precision=precision)
        dev.set_stream(stream.cuda_stream)
        dev.initialize_primitive(*prim)
        dts = torch.zeros(args.warmup + 2 * args.steps, dtype=torch.float64, device="cuda")
        base = dts.data_ptr()
        for s in range(args.warmup):
            dev.advance_steps(1, base + 8 * s)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record(stream)
        dev.advance_steps(args.steps, base + 8 * args.warmup)
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        dev.kernel_times()
        dev.set_kernel_timing(True)
        dev.advance_steps(args.steps, base + 8 * (args.warmup + args.steps))
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
    valu = profiled_valu(dom, n)
    roofline = {
        "bound": "fp64_valu",
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
        # the same kernel against the HBM roof on its measured bytes: the x pencil runs on memory (cutting a
        # sixth of its VALU instructions left its time unchanged), so this is the roof it presses on
        "hbm_frac_measured_traffic": round(traffic / (per_launch[dom] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4) if traffic else None,
        # rocprofv3 counters of the same kernel (profiles/<kernel>_valu_<n>.md): share of cycles the VALU
        # is busy, fraction of the one-wave-instruction-per-CU-per-cycle issue rate, counted FP64 rate
        "valu_busy_pct": valu.get("valu_busy_pct"),
        "valu_issue_frac": valu.get("valu_issue"),
        "fp64_TFLOPs_counted": valu.get("fp64_tflops_counted"),
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
    if args.sweep:
        out["sweep"] = sweep_report(args, prim, ma, G, stream)
    if args.adjoint:
        out["adjoint"] = adjoint_report(args)
    if args.scale_mixed:
        out["mixed_scale"] = scale_mixed_report(args)
    if args.scale_adjoint:
        out["adjoint_scale"] = scale_adjoint_report(args)
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prim, G, P, cpu_threads(args.cpu_threads))
        # one core (the analogue of one MPI rank of the reference) on a smaller box of the same workload:
        # the per-cell work of a step does not depend on the box size
        m = args.cpu_1core_n
        prim1, _ = isotropic_state(m, mach_t=args.mach_t, seed=12345)
        out["cpu_baseline_1core"] = cpu_baseline(prim1, periodic_box(m), P, 1, "box of the same workload")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
