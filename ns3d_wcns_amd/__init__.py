"""MI355X (gfx950) RHS + SSP-RK3 hot path of the WCNS compressible
Navier-Stokes solver for one periodic box (drop-in for wylx-2/ns3d_wcns's
runge_kutta_3 / compute_rhs / compute_timestep path)."""
from .params import (BCType, GridDesc, Interpolation, RiemannSolver, SCHEME_ALIASES, SolverParams,
                     periodic_box, read_solver_params_from_file)
from ._lib import NS3DError, PRECISION_FP64, PRECISION_MIXED, load_library
from .field import DeviceField

__all__ = [
    "BCType", "GridDesc", "Interpolation", "RiemannSolver", "SCHEME_ALIASES", "SolverParams", "periodic_box",
    "read_solver_params_from_file", "NS3DError", "PRECISION_FP64", "PRECISION_MIXED", "load_library", "DeviceField",
]
