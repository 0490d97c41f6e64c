"""Run RK3 steps of the synthetic isotropic-turbulence box on the device.

    python -m ns3d_wcns_amd.run_steps --n 256 --steps 1 [--scheme wcns5] [--riemann HLLC] [--mixed]

Initialises the box, records U0 once (five plain device copies of known size,
the calibration point for byte counters), then advances --steps steps
entirely on the device.  Used as the program under rocprofv3 counter passes.
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
    args = ap.parse_args(argv)
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


if __name__ == "__main__":
    main()
