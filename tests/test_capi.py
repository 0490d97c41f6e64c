"""The C-ABI library loads without a GPU and exports every entry point that
include/ns3d_hip.h declares; host-side logic (parameters, parser, flop count)."""
import os
import re
import subprocess

import pytest

from ns3d_wcns_amd import BCType, Interpolation, RiemannSolver, SolverParams, read_solver_params_from_file
from ns3d_wcns_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ns3d_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ns3d_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    assert declared_functions() == _lib.exported_symbols()


def test_library_exports_every_declared_symbol():
    lib = _lib.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (ns3d_\w+)", out.stdout))
    for name in declared_functions():
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_derived_parameters():
    """initialization.cpp:194-197."""
    P = SolverParams(gamma=1.4, Ma=0.5, Re=200).derive()
    assert P.Cv == pytest.approx(1.0 / (1.4 * 0.4 * 0.25))
    assert P.Cp == pytest.approx(1.4 * P.Cv)
    assert P.Rgas == pytest.approx(1.0 / (0.25 * 1.4))
    assert P.mu == 1.0 / 200


def test_solver_in_parser(tmp_path):
    f = tmp_path / "solver.in"
    f.write_text("""# comment
Re = 1600
Ma=0.3
gamma = 1.4
interpolation = MDCD_HYBRID
Riemann_Solver = Rusanov
char_recon = True
global_nx = 32
Lx = 6.283185307179586
bc_xmin = Outflow
bc_ymin = nonsense
unknown_key = 5
""")
    P, G, periods = read_solver_params_from_file(str(f))
    assert P.Re == 1600 and P.Ma == 0.3
    assert P.interpolation == Interpolation.MDCD_HYBRID and P.riemann_solver == RiemannSolver.Rusanov
    assert P.char_recon
    assert G.global_nx == 32 and G.dx == pytest.approx(6.283185307179586 / 32)
    assert P.bc[0] == BCType.Outflow and P.bc[2] == BCType.Periodic
    assert periods == [0, 1, 1]
    assert P.mu == 1.0 / 1600


def test_missing_file_raises():
    with pytest.raises(OSError):
        read_solver_params_from_file("/nonexistent/solver.in")


def test_algorithmic_flops_host_side():
    f = _lib.algorithmic_flops(SolverParams(Ma=0.2, riemann_solver=RiemannSolver.HLLC).derive())
    assert f["per_step"] == 3 * f["per_stage"] + f["cfl"]
    assert 700 < f["face"] < 1000
    fc = _lib.algorithmic_flops(SolverParams(Ma=0.2, riemann_solver=RiemannSolver.HLLC, char_recon=True).derive())
    assert fc["face"] > f["face"]
    fr = _lib.algorithmic_flops(SolverParams(Ma=0.2, riemann_solver=RiemannSolver.Rusanov).derive())
    assert fr["face"] < f["face"]


def test_boundary_combinations_checked_before_device_work():
    """ns3d_create validates the boundary types before any HIP call: a
    Periodic face opposite a non-periodic one, an unknown type, and the mixed
    method on a box with physical faces are refused (NS3D_E_UNSUPPORTED /
    NS3D_E_ARG), so this runs without a GPU."""
    import ctypes

    from ns3d_wcns_amd import periodic_box

    lib = _lib.load_library()
    grid = periodic_box(8).to_c(3)
    cases = [
        ([0, 3, 0, 0, 0, 0], _lib.PRECISION_FP64, -2, "Periodic face opposite"),
        ([9, 9, 0, 0, 0, 0], _lib.PRECISION_FP64, -1, "unknown boundary type"),
        ([3, 3, 0, 0, 0, 0], _lib.PRECISION_MIXED, -2, "periodic boxes only"),
    ]
    for bc, prec, code, msg in cases:
        P = SolverParams(Ma=0.3, Re=100.0).derive()
        cp = P.to_c()
        for f in range(6):
            cp.bc[f] = bc[f]
        h = ctypes.c_void_p()
        rc = lib.ns3d_create(ctypes.byref(h), ctypes.byref(grid), ctypes.byref(cp), prec)
        assert rc == code, (bc, rc)
        assert msg in lib.ns3d_last_error().decode(), lib.ns3d_last_error()
