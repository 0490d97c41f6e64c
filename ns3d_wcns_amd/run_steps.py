"""Run RK3 steps of the synthetic isotropic-turbulence box on the device.

    python -m ns3d_wcns_amd.run_steps --n 256 --steps 1 [--scheme wcns5] [--riemann HLLC] [--mixed]
                                      [--adjoint-steps K]

Initialises the box, records U0 once (five plain device copies of known size,
the calibration point for byte counters), then advances --steps steps
entirely on the device; with --adjoint-steps K, then K discrete-adjoint steps
(fixed dt) from the final state.  Used as the program under rocprofv3.
"""
import argparse
import os
import sys

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--scheme", default="wcns5")
    ap.add_argument("--riemann", default="HLLC")
    ap.add_argument("--mach-t", type=float, default=0.6)
    ap.add_argument("--char", action="store_true")
    ap.add_argument("--mixed", action="store_true", help="the mixed FP32/FP64 method")
    ap.add_argument("--adjoint-steps", type=int, default=0)
    args = ap.parse_args(argv)
    if args.adjoint_steps:
        import torch

        torch.cuda.set_device(0)  # torch's runtime first, before the library touches the device
    from ns3d_wcns_amd import (PRECISION_FP64, PRECISION_MIXED, DeviceField, RiemannSolver, SCHEME_ALIASES,
                               SolverParams, periodic_box)
    from ns3d_wcns_amd.isotropic import isotropic_state

    prim, ma = isotropic_state(args.n, mach_t=args.mach_t, seed=12345)
    P = SolverParams(Ma=ma, Re=100.0, Pr=0.72, cfl=0.5, interpolation=SCHEME_ALIASES[args.scheme],
                     riemann_solver=RiemannSolver[args.riemann], char_recon=args.char).derive()
    dev = DeviceField(periodic_box(args.n), P, precision=PRECISION_MIXED if args.mixed else PRECISION_FP64)
    dev.initialize_primitive(*prim)
    dev.recordConservedTo0()
    dev.advance_steps(args.steps)
    dev.synchronize()
    print(f"{args.steps} step(s) of {args.n}^3 done")
    if args.adjoint_steps:
        import torch

        from ns3d_wcns_amd.adjoint import state_tensor

        dev.set_stream(torch.cuda.current_stream().cuda_stream)
        U = state_tensor(dev)
        adj = torch.zeros_like(U)
        adj[:, 3:-3, 3:-3, 3:-3] = 1.0 / args.n ** 3
        dt = dev.compute_timestep()
        ptr = [adj[q].data_ptr() for q in range(5)]
        for _ in range(args.adjoint_steps):
            dev.set_state([U[q].data_ptr() for q in range(5)])
            dev.adjoint_step(dt, ptr)
        torch.cuda.synchronize()
        print(f"{args.adjoint_steps} adjoint step(s) done, |adj| = {float(adj.abs().max()):.3e}")


if __name__ == "__main__":
    main()
