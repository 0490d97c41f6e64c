"""The C++ drop-in (include/ns3d_dropin.h, libns3d_dropin.so): the
reference's hot-path names and signatures over the C-ABI, its solver.in
reader, and a time-advance program written against it
(ns3d_wcns_amd/lib/ns3d_time_advance) against the CPU checker.

- parser: every parsed and derived value equal to the Python reader's;
- time advance (faithful and device-resident mirror): the dt sequence and
  the final conserved fields within 1e-10 of the checker's time_advance
  (norm-wise relative), the monitored kinetic energy within 1e-12.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import box, params, state
from ns3d_wcns_amd import read_solver_params_from_file

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ns3d_wcns_amd", "lib")
DRIVER = os.path.join(LIB, "ns3d_time_advance")
DROPIN = os.path.join(LIB, "libns3d_dropin.so")

SOLVER_IN = """# isotropic box
gamma = 1.4
Pr = 0.72
Ma = {ma!r}
Re = 100
cfl = 0.5
interpolation = weno5
riemann_solver = HLLC
char_recon = false
global_nx = {nx}
global_ny = {ny}
global_nz = {nz}
Lx = 6.283185307179586
Ly = 6.283185307179586
Lz = 6.283185307179586
max_steps = {steps}
monitor_Stepfreq = 1
output_Timefreq = 1e30
TotalTime = 1e30
"""


def test_dropin_exports_reference_signatures():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", DROPIN], capture_output=True, text=True, check=True).stdout
    for sig in ("runge_kutta_3(Field3D&, CartDecomp&, GridDesc&, SolverParams&, HaloRequests&, double)",
                "compute_rhs(Field3D&, CartDecomp&, GridDesc&, SolverParams&, HaloRequests&)",
                "compute_timestep(Field3D&, GridDesc const&, SolverParams const&)",
                "compute_invis_flux(Field3D&, SolverParams const&)",
                "compute_invis_dflux(Field3D&, SolverParams const&, GridDesc const&)",
                "apply_boundary(Field3D&, GridDesc&, CartDecomp&, SolverParams const&)",
                "apply_boundary_halfnode_flux(Field3D&, GridDesc const&, CartDecomp&, SolverParams const&)",
                "compute_gradients(Field3D&, GridDesc const&)",
                "compute_viscous_flux(Field3D&, SolverParams const&)",
                "compute_vis_flux(Field3D&, GridDesc const&)",
                "exchange_halos_viscous_flux(Field3D&, CartDecomp&, LocalDesc&, HaloRequests&)",
                "read_solver_params_from_file(std::__cxx11::basic_string<char, std::char_traits<char>, "
                "std::allocator<char> > const&, SolverParams&, GridDesc&, CartDecomp&)",
                "Field3D::conservedToPrimitive(SolverParams const&)",
                "Field3D::primitiveToConserved(SolverParams const&)",
                "Field3D::recordConservedTo0()"):
        assert sig in out, sig


def test_dropin_parser_matches_python_reader(tmp_path):
    f = tmp_path / "solver.in"
    f.write_text("""# comment
Re = 1600
Ma=0.3
gamma = 1.4
interpolation = MDCD_HYBRID
Riemann_Solver = Rusanov
char_recon = True
global_nx = 32
global_ny = 24
global_nz = 20
Lx = 6.283185307179586
Ly = 3.0
cfl = 0.4
dt_fixed = 0.001
max_steps = 7
TotalTime = 2.5
bc_xmin = Outflow
bc_ymin = nonsense
unknown_key = 5
""")
    out = subprocess.run([DRIVER, str(f), "--print-params"], capture_output=True, text=True, check=True).stdout
    got = dict(line.split("=", 1) for line in out.split())
    P, G, periods = read_solver_params_from_file(str(f))
    for k in ("gamma", "Pr", "Ma", "Re", "cfl", "dt_fixed", "mu", "Cv", "Cp", "Rgas", "mdcd_diss", "mdcd_disp",
              "output_Timefreq", "TotalTime"):
        assert float(got[k]) == float(getattr(P, k)), k
    for k in ("interpolation", "riemann_solver", "char_recon", "ghost_layers", "stencil", "max_steps",
              "monitor_Stepfreq"):
        assert int(got[k]) == int(getattr(P, k)), k
    for k in ("global_nx", "global_ny", "global_nz"):
        assert int(got[k]) == getattr(G, k)
    for k in ("dx", "dy", "dz"):
        assert float(got[k]) == getattr(G, k)
    assert [int(x) for x in got["periods"].split(",")] == periods


def write_case(tmp_path, n, steps, seed):
    prim, ma = state(n, 0.3, noise=0.01, seed=seed)
    nx, ny, nz = n
    (tmp_path / "solver.in").write_text(SOLVER_IN.format(ma=ma, nx=nx, ny=ny, nz=nz, steps=steps))
    with open(tmp_path / "initial.bin", "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        for a in prim:
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
    return prim, ma


def read_final(path, n, steps):
    nx, ny, nz = n
    cells = (nx + 6) * (ny + 6) * (nz + 6)
    raw = np.fromfile(path, dtype=np.uint8)
    nsteps = int(np.frombuffer(raw[:4].tobytes(), dtype=np.int32)[0])
    d = np.frombuffer(raw[4:].tobytes(), dtype=np.float64)
    dts = d[:nsteps]
    fields = d[nsteps:nsteps + 5 * cells].reshape(5, nz + 6, ny + 6, nx + 6)
    energy = d[nsteps + 5 * cells:]
    return dts, fields, energy


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["faithful", "resident"])
def test_dropin_time_advance_matches_checker(gpu, tmp_path, mode):
    from oracle.ns3d_oracle import OracleField

    n, steps = (16, 12, 10), 4
    prim, ma = write_case(tmp_path, n, steps, seed=7)
    args = [DRIVER, str(tmp_path / "solver.in"), str(tmp_path / "initial.bin"), str(tmp_path / "final.bin")]
    if mode == "resident":
        args.append("resident")
    print("running", " ".join(args), flush=True)
    subprocess.run(args, check=True, timeout=120)
    dts, fields, energy = read_final(tmp_path / "final.bin", n, steps)
    P = params(ma)
    G = box(n)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    rdts = ref.time_advance(steps)
    assert len(dts) == len(rdts) == steps
    assert np.max(np.abs(dts - rdts) / rdts) < 1e-12, (dts, rdts)
    for q, name in enumerate(("rho", "rhou", "rhov", "rhow", "E")):
        a = fields[q][3:-3, 3:-3, 3:-3]
        b = ref.interior(ref.view(name))
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < 1e-10, (name, err)
    ke = ref.total_kinetic()
    assert abs(energy[-1] - ke) <= 1e-12 * abs(ke), (energy[-1], ke)
    ref.close()


def test_python_params_carry_boundary_types():
    from ns3d_wcns_amd import SolverParams
    from ns3d_wcns_amd.params import BCType

    P = SolverParams().derive()
    P.bc[2] = P.bc[3] = BCType.Outflow
    assert list(P.to_c().bc) == [0, 0, 3, 3, 0, 0]
    assert not P.periodic


def test_driver_parses_physical_faces(tmp_path):
    """solver.in's bc_* keys reach the driver: an axis is periodic only when
    both of its faces are (initialization.cpp:160-168)."""
    f = tmp_path / "solver.in"
    f.write_text("global_nx = 8\nglobal_ny = 8\nglobal_nz = 8\nbc_xmin = Wall\nbc_xmax = Outflow\n"
                 "bc_ymin = Periodic\nbc_ymax = Periodic\nbc_zmin = Inflow\nbc_zmax = Symmetry\n")
    r = subprocess.run([DRIVER, str(f), "--print-params"], capture_output=True, text=True, check=True)
    assert "periods=0,1,0" in r.stdout


REFERENCE_DEFAULT = """# the reference's default run (2D Riemann problem, Outflow faces), on a smaller grid
Re = 500000
Ma = 0.5
gamma = 1.4
Pr = 0.71
cfl = 0.2
dt_fixed = -0.0005
interpolation = WENO5
char_recon = False
Riemann_Solver = Roe
mdcd_disp = 0.0463783
mdcd_diss = 0.020
global_nx = 32
global_ny = 32
global_nz = 6
bc_xmin = Outflow
bc_xmax = Outflow
bc_ymin = Outflow
bc_ymax = Outflow
bc_zmin = Outflow
bc_zmax = Outflow
Lx = 1.0
Ly = 1.0
Lz = 1.0
max_steps = {steps}
monitor_Stepfreq = 1
output_Timefreq = 0.1
TotalTime = 0.1
restart = False
"""


@pytest.mark.gpu
def test_dropin_reference_default_run(gpu, tmp_path):
    """The driver with initialize_riemann_2d (the reference's built-in
    initial state) and Outflow on all six faces against the checker."""
    from oracle.ns3d_oracle import OracleField
    from test_gpu_boundaries import riemann_2d_case

    n, steps = (32, 32, 6), 10
    (tmp_path / "solver.in").write_text(REFERENCE_DEFAULT.format(steps=steps))
    args = [DRIVER, str(tmp_path / "solver.in"), "riemann2d", str(tmp_path / "final.bin")]
    print("running", " ".join(args), flush=True)
    subprocess.run(args, check=True, timeout=120)
    dts, fields, energy = read_final(tmp_path / "final.bin", n, steps)
    P, G, prim = riemann_2d_case(n)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    rdts = ref.time_advance(steps, -1.0, 0.1, 0.1)
    assert len(dts) == len(rdts) == steps
    assert np.max(np.abs(dts - rdts) / rdts) < 1e-12
    from test_gpu_boundaries import state_errors

    names = ("rho", "rhou", "rhov", "rhow", "E")
    for name, err in state_errors(lambda q: fields[names.index(q)], ref.view).items():
        assert err < 1e-10, (name, err)
    ref.close()
