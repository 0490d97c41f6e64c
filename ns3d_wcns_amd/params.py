"""Solver parameters and grid description of one periodic box.

Mirrors ``SolverParams`` / ``GridDesc`` (field_structures.h:33-186) and the
``solver.in`` reader ``read_solver_params_from_file`` (initialization.cpp:36-200),
including its derived values (initialization.cpp:161-197).
"""
from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field


class Interpolation(enum.IntEnum):
    """``SolverParams::Interpolation`` (field_structures.h:127-132) plus the
    Jiang-Shu WENO5 weights of ``weno5_reconstruction`` (reconstruction.cpp:69-100)."""

    WENO5 = 0  # the reference's "weno5" key: WCNS5 weights 1/16, 5/8, 5/16
    ZERO = 1  # rejected by the device library
    MDCD_LINEAR = 2
    MDCD_HYBRID = 3
    WENO5_JS = 4


# Scheme names used by the throughput sweep: "WCNS5" is the reference's
# interpolation = WENO5, "WENO5" the Jiang-Shu weights, "MDCD" the hybrid.
SCHEME_ALIASES = {
    "wcns5": Interpolation.WENO5,
    "weno5": Interpolation.WENO5,
    "weno5_js": Interpolation.WENO5_JS,
    "mdcd": Interpolation.MDCD_HYBRID,
    "mdcd_hybrid": Interpolation.MDCD_HYBRID,
    "mdcd_linear": Interpolation.MDCD_LINEAR,
    "zero": Interpolation.ZERO,
}


class RiemannSolver(enum.IntEnum):
    """``SolverParams::RiemannSolver`` (field_structures.h:134-141)."""

    Roe = 0
    Rusanov = 1
    HLLC = 2
    HLL = 3
    HLLC_p = 4
    AUSM = 5


class BCType(enum.IntEnum):
    Periodic = 0
    Wall = 1
    Symmetry = 2
    Outflow = 3
    Inflow = 4


class CParams(ctypes.Structure):
    """``ns3d_params`` of include/ns3d_hip.h (and the oracle's ``orc_params``)."""

    _fields_ = [
        ("gamma", ctypes.c_double),
        ("Pr", ctypes.c_double),
        ("Ma", ctypes.c_double),
        ("Re", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("S_ref", ctypes.c_double),
        ("Cv", ctypes.c_double),
        ("Cp", ctypes.c_double),
        ("Rgas", ctypes.c_double),
        ("mdcd_diss", ctypes.c_double),
        ("mdcd_disp", ctypes.c_double),
        ("interpolation", ctypes.c_int),
        ("riemann_solver", ctypes.c_int),
        ("char_recon", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("pad_", ctypes.c_int),
    ]


class CGrid(ctypes.Structure):
    """``ns3d_grid`` of include/ns3d_hip.h."""

    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("ng", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("dz", ctypes.c_double),
    ]


@dataclass
class SolverParams:
    """Numerical and control fields of ``SolverParams`` with the reference's
    defaults (field_structures.h:86-186)."""

    gamma: float = 1.4
    Pr: float = 0.71
    Ma: float = 0.4
    Re: float = 500.0
    cfl: float = 0.1
    dt_fixed: float = -1.0
    S_ref: float = 110.4 / 288.15
    interpolation: Interpolation = Interpolation.WENO5
    riemann_solver: RiemannSolver = RiemannSolver.Roe
    char_recon: bool = False
    mdcd_diss: float = 0.01
    mdcd_disp: float = 0.0463783
    stencil: int = 6
    ghost_layers: int = 3
    bc: list = field(default_factory=lambda: [BCType.Periodic] * 6)
    max_steps: int = 1
    monitor_Stepfreq: int = 1
    output_Timefreq: float = 1.0
    TotalTime: float = 1.0
    monitor_res: bool = True
    monitor_energy: bool = True
    post_basicfield: bool = True
    isotropic_analyse: bool = True
    restart: bool = False
    restart_file: str = "field.h5"
    # derived (set by derive(); struct defaults follow field_structures.h:93-115)
    mu: float = 1.0 / 500.0
    Cv: float = 1.0 / (1.4 * 0.4 * 0.4 * 0.4)
    Cp: float = 1.4 / (1.4 * 0.4 * 0.4 * 0.4)
    Rgas: float = 1.0 / (0.4 * 0.4 * 1.4)

    def derive(self) -> "SolverParams":
        """Derived values exactly as initialization.cpp:175-197 sets them."""
        if self.interpolation == Interpolation.ZERO:
            self.ghost_layers, self.stencil = 1, 2
        else:
            self.ghost_layers, self.stencil = 3, 6
        self.Cv = 1.0 / (self.gamma * (self.gamma - 1.0) * self.Ma * self.Ma)
        self.Cp = self.Cv * self.gamma
        self.Rgas = 1.0 / (self.Ma * self.Ma * self.gamma)
        self.mu = 1.0 / self.Re
        return self

    def to_c(self) -> CParams:
        """The C-ABI parameter block, boundary types included (an axis with
        one Periodic and one other face is refused by the library)."""
        if len(self.bc) != 6:
            raise ValueError(f"six boundary types expected (got {len(self.bc)})")
        return CParams(
            self.gamma, self.Pr, self.Ma, self.Re, self.cfl, self.mu, self.S_ref, self.Cv, self.Cp,
            self.Rgas, self.mdcd_diss, self.mdcd_disp, int(self.interpolation), int(self.riemann_solver),
            int(bool(self.char_recon)), (ctypes.c_int * 6)(*[int(b) for b in self.bc]), 0,
        )

    @property
    def periodic(self) -> bool:
        """True when all six faces are periodic."""
        return all(b == BCType.Periodic for b in self.bc)


@dataclass
class GridDesc:
    """``GridDesc`` (field_structures.h:33-38) of the whole periodic box."""

    global_nx: int = 16
    global_ny: int = 16
    global_nz: int = 16
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0

    def derive(self) -> "GridDesc":
        """dx = L / N (initialization.cpp:170-172)."""
        self.dx = self.Lx / self.global_nx
        self.dy = self.Ly / self.global_ny
        self.dz = self.Lz / self.global_nz
        return self

    def to_c(self, ng: int = 3) -> CGrid:
        return CGrid(self.global_nx, self.global_ny, self.global_nz, ng, self.dx, self.dy, self.dz)


def periodic_box(n, length=2.0 * math.pi) -> GridDesc:
    """A periodic cube (or box, for a 3-tuple n) of side ``length``."""
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    return GridDesc(nx, ny, nz, Lx=length, Ly=length, Lz=length).derive()


def _parse_bool(v: str) -> bool:
    return v.lower() in ("true", "ture", "yes", "1", "on")


def _flag(v: str) -> bool:
    """The reference's boolean spelling set (initialization.cpp:30-34)."""
    return v in ("true", "ture", "yes", "1", "on")


def read_solver_params_from_file(fname: str):
    """Parse a ``solver.in`` file: ``key = value`` lines, keys and values
    case-insensitive, ``#`` comments, unknown keys ignored, unknown BC strings
    map to Periodic (initialization.cpp:36-200).  Returns (params, grid,
    periods) or raises OSError like the reference's failed open (which
    returns false)."""
    P = SolverParams()
    G = GridDesc()
    with open(fname, "r") as fh:
        lines = fh.read().splitlines()
    bcs = {"bc_xmin": 0, "bc_xmax": 1, "bc_ymin": 2, "bc_ymax": 3, "bc_zmin": 4, "bc_zmax": 5}
    bcmap = {"inflow": BCType.Inflow, "wall": BCType.Wall, "symmetry": BCType.Symmetry, "outflow": BCType.Outflow}
    for raw in lines:
        line = raw.strip()
        if not line or line.startswith("#") or "=" not in line:
            continue
        key, val = line.split("=", 1)
        key, val = key.strip(), val.strip()
        k, v = key.lower(), val.lower()
        if k == "gamma": P.gamma = float(val)
        elif k == "pr": P.Pr = float(val)
        elif k == "ma": P.Ma = float(val)
        elif k == "re": P.Re = float(val)
        elif k == "cfl": P.cfl = float(val)
        elif k == "dt_fixed": P.dt_fixed = float(val)
        elif k == "interpolation":
            m = {"mdcd_hybrid": Interpolation.MDCD_HYBRID, "mdcd_linear": Interpolation.MDCD_LINEAR,
                 "weno5": Interpolation.WENO5, "zero": Interpolation.ZERO, "weno5_js": Interpolation.WENO5_JS}
            if v in m:
                P.interpolation = m[v]
        elif k == "char_recon": P.char_recon = v in ("yes", "true")
        elif k == "riemann_solver":
            m = {"roe": RiemannSolver.Roe, "rusanov": RiemannSolver.Rusanov, "hllc": RiemannSolver.HLLC,
                 "hll": RiemannSolver.HLL, "hllc_p": RiemannSolver.HLLC_p, "ausm": RiemannSolver.AUSM}
            if v in m:
                P.riemann_solver = m[v]
        elif k == "mdcd_diss": P.mdcd_diss = float(val)
        elif k == "mdcd_disp": P.mdcd_disp = float(val)
        elif k == "global_nx": G.global_nx = int(val)
        elif k == "global_ny": G.global_ny = int(val)
        elif k == "global_nz": G.global_nz = int(val)
        elif k == "x0": G.x0 = float(val)
        elif k == "y0": G.y0 = float(val)
        elif k == "z0": G.z0 = float(val)
        elif k == "lx": G.Lx = float(val)
        elif k == "ly": G.Ly = float(val)
        elif k == "lz": G.Lz = float(val)
        elif k == "max_steps": P.max_steps = int(val)
        elif k == "monitor_stepfreq": P.monitor_Stepfreq = int(val)
        elif k == "output_timefreq": P.output_Timefreq = float(val)
        elif k == "totaltime": P.TotalTime = float(val)
        elif k == "restart": P.restart = _flag(v)
        elif k == "restart_file": P.restart_file = val
        elif k == "post_basicfield": P.post_basicfield = _flag(v)
        elif k == "isotropicanalyse": P.isotropic_analyse = _flag(v)
        elif k == "monitor_res": P.monitor_res = _flag(v)
        elif k == "monitor_energy": P.monitor_energy = _flag(v)
        if k in bcs:
            P.bc[bcs[k]] = bcmap.get(v, BCType.Periodic)
    periods = [int(P.bc[2 * a] == BCType.Periodic and P.bc[2 * a + 1] == BCType.Periodic) for a in range(3)]
    G.derive()
    P.derive()
    return P, G, periods
