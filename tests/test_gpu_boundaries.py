"""Boxes with physical faces (Wall, Symmetry, Outflow, Inflow; boundary.cpp,
ViscousFlux.cpp:15-50) on the device through the C-ABI against the CPU
checker.

- apply_boundary + primitiveToConserved: every cell of rho, u, v, w, p, T and
  the conserved fields bit for bit (ghost layers, edges and corners
  included);
- the stepwise entry points: face fluxes in the reference's face layout
  (boundary faces and ghost faces included), gradients, viscous fluxes and
  the RHS, 1e-12 norm-wise relative;
- one RK3 step, 1e-12 on every cell;
- the reference's default run (initialize_riemann_2d with Outflow on all six
  faces, Roe flux, WCNS5, CFL 0.2; solver.in, main.cpp:85) on 64 x 64 x 6
  for 30 CFL steps: dt sequence 1e-12, fields 1e-10 (the momenta against the
  largest momentum component: rho w vanishes there up to rounding);
- Sod's shock tube (Outflow x faces, 200 cells) to t = 0.2: the checker's
  run to 1e-10 and the exact solution to 6e-3 in L1(rho);
- the derivative entry points refuse a box with physical faces.
"""
import numpy as np
import pytest

from cases import box, params, rel_err, state
from ns3d_wcns_amd import DeviceField, GridDesc, Interpolation, NS3DError, RiemannSolver, SolverParams
from ns3d_wcns_amd.params import BCType
from oracle.ns3d_oracle import OracleField, set_threads

pytestmark = pytest.mark.gpu

W, S, O, I, P_ = BCType.Wall, BCType.Symmetry, BCType.Outflow, BCType.Inflow, BCType.Periodic
BC_SETS = {
    "wall_outflow_x": [W, O, P_, P_, P_, P_],
    "symmetry_y_inflow_z": [P_, P_, S, S, I, O],
    "all_outflow": [O] * 6,
    "mixed": [I, O, W, S, O, W],
}
SCHEMES = [(Interpolation.WENO5, RiemannSolver.HLLC, False), (Interpolation.MDCD_HYBRID, RiemannSolver.Roe, True),
           (Interpolation.WENO5_JS, RiemannSolver.AUSM, False), (Interpolation.MDCD_LINEAR, RiemannSolver.HLLC_p, False)]
N = (12, 10, 8)
CONS = ("rho", "rhou", "rhov", "rhow", "E")
PRIMS = ("u", "v", "w", "p", "T")


def say(msg):
    print(msg, flush=True)


def state_errors(get_dev, get_ref):
    """Norm-wise errors per conserved field; the momenta against the largest
    momentum component (a momentum that vanishes up to rounding, like rho w
    of a two-dimensional flow, has no scale of its own)."""
    mom = max(float(np.max(np.abs(get_ref(q)))) for q in ("rhou", "rhov", "rhow"))
    out = {}
    for q in CONS:
        a, b = get_dev(q), get_ref(q)
        scale = mom if q in ("rhou", "rhov", "rhow") else float(np.max(np.abs(b)))
        out[q] = float(np.max(np.abs(a - b))) / scale
    return out


def pair(n, bc, interp=Interpolation.WENO5, rs=RiemannSolver.HLLC, char=False, seed=4):
    prim, ma = state(n, 0.3, noise=0.02, seed=seed)
    P = params(ma, interp, rs, char=char)
    P.bc = list(bc)
    G = box(n)
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    return dev, ref


@pytest.mark.parametrize("name", list(BC_SETS))
def test_apply_boundary_bit_exact(gpu, name):
    say(f"boundary fill {name}")
    dev, ref = pair(N, BC_SETS[name])
    for q in CONS + PRIMS:
        assert np.array_equal(dev.download(q), ref.view(q)), (name, q)
    dev.close()
    ref.close()


def face_shape(f, d):
    return [(f.sz, f.sy, f.sx - 1), (f.sz, f.sy - 1, f.sx), (f.sz - 1, f.sy, f.sx)][d]


def interior_lines(a, f, d):
    """The interior lines of a face array (all faces along d)."""
    sl = [slice(3, 3 + f.nz), slice(3, 3 + f.ny), slice(3, 3 + f.nx)]
    sl[2 - d] = slice(None)
    return a.reshape(face_shape(f, d))[tuple(sl)]


@pytest.mark.parametrize("name", list(BC_SETS))
@pytest.mark.parametrize("interp,rs,char", SCHEMES, ids=lambda x: getattr(x, "name", str(x)))
def test_stepwise_parity(gpu, name, interp, rs, char):
    say(f"stepwise {name} {interp.name}+{rs.name} char={char}")
    dev, ref = pair(N, BC_SETS[name], interp, rs, char)
    comps = ("mass", "momx", "momy", "momz", "E")
    dev.compute_invis_flux()
    dev.apply_boundary_halfnode_flux()
    ref.compute_invis_flux()
    ref.apply_boundary_halfnode_flux()
    for d, tag in enumerate(("fx", "fy", "fz")):
        for c, comp in enumerate(comps):
            a = interior_lines(dev.download_face_flux(d, c), dev, d)
            b = interior_lines(ref.view(f"flux_{tag}_{comp}"), dev, d)
            assert rel_err(a, b) < 1e-12, (tag, comp, rel_err(a, b))
    dev.compute_gradients()
    ref.compute_gradients()
    names = ("du_dx", "du_dy", "du_dz", "dv_dx", "dv_dy", "dv_dz", "dw_dx", "dw_dy", "dw_dz", "dT_dx", "dT_dy",
             "dT_dz")
    for q, g in enumerate(names):
        a = dev.interior(dev.download_gradient(q))
        b = ref.interior(ref.view(g))
        assert rel_err(a, b) < 1e-12, (g, rel_err(a, b))
    dev.compute_viscous_flux()
    ref.compute_viscous_flux()
    ref.exchange_halos_viscous_flux()
    for d, tag in enumerate(("Fvflux", "Hvflux", "Gvflux")):
        for c in range(1, 5):
            a = dev.download_viscous_flux(d, c)
            b = ref.view(f"{tag}_{comps[c]}")
            assert rel_err(dev.interior(a), ref.interior(b)) < 1e-12, (tag, c)
    dev.compute_rhs()
    ref.compute_rhs()
    for q in CONS:
        e = rel_err(dev.interior(dev.download("rhs_" + q)), ref.interior(ref.view("rhs_" + q)))
        assert e < 1e-12, (q, e)
    dev.close()
    ref.close()


@pytest.mark.parametrize("name", list(BC_SETS))
@pytest.mark.parametrize("interp,rs,char", SCHEMES[:3], ids=lambda x: getattr(x, "name", str(x)))
def test_rk3_step(gpu, name, interp, rs, char):
    say(f"RK3 step {name} {interp.name}+{rs.name} char={char}")
    dev, ref = pair(N, BC_SETS[name], interp, rs, char)
    dt = ref.compute_timestep()
    assert dev.compute_timestep() == dt
    dev.recordConservedTo0()
    dev.runge_kutta_3(dt)
    ref.record_conserved_to0()
    ref.runge_kutta_3(dt)
    for q in CONS + PRIMS:
        e = rel_err(dev.download(q), ref.view(q))
        assert e < 1e-12, (q, e)
    dev.close()
    ref.close()


def riemann_2d_case(n):
    """The reference's default run (solver.in): 2D Riemann problem,
    Outflow on every face, unit box, Roe + WCNS5, Re 5e5, Ma 0.5, Pr 0.71,
    CFL 0.2; initialize_riemann_2d (initialization.cpp:516-568)."""
    P = SolverParams(Ma=0.5, Re=500000.0, Pr=0.71, cfl=0.2, interpolation=Interpolation.WENO5,
                     riemann_solver=RiemannSolver.Roe).derive()
    P.bc = [O] * 6
    G = GridDesc(n[0], n[1], n[2], Lx=1.0, Ly=1.0, Lz=1.0).derive()
    x = (np.arange(n[0]) + 0.5) * G.dx
    y = (np.arange(n[1]) + 0.5) * G.dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    right, top = X >= 0.5, Y >= 0.5
    same = right == top
    shape = (n[2], n[1], n[0])
    rho = np.broadcast_to(np.where(same, 1.1, 0.5065), shape).copy()
    u = np.broadcast_to(np.where(right, 0.0, 0.8939), shape).copy()
    v = np.broadcast_to(np.where(top, 0.0, 0.8939), shape).copy()
    p = np.broadcast_to(np.where(same, 1.1, 0.35), shape).copy()
    return P, G, (rho, u, v, np.zeros(shape), p)


def test_reference_default_run_riemann_2d(gpu):
    n, steps = (64, 64, 6), 30
    P, G, prim = riemann_2d_case(n)
    say(f"2D Riemann, Outflow faces, {n} x {steps} steps")
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    dts = dev.time_advance(steps)
    set_threads(16)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    rdts = ref.time_advance(steps)
    assert len(dts) == len(rdts) == steps
    e_dt = float(np.max(np.abs(dts - rdts) / rdts))
    errs = state_errors(dev.download, ref.view)
    say(f"  dt {e_dt:.1e}, fields {errs}")
    assert e_dt < 1e-12
    for q, e in errs.items():
        assert e < 1e-10, (q, e)
    # stays two-dimensional
    r = dev.interior(dev.download("rho"))
    assert np.ptp(r, axis=0).max() < 1e-12
    dev.close()
    ref.close()


def test_sod_shock_tube(gpu):
    from test_oracle_boundaries import sod_exact

    nx = 200
    P = SolverParams(Ma=0.5, Re=1e8, Pr=0.72, cfl=0.5, interpolation=Interpolation.WENO5,
                     riemann_solver=RiemannSolver.HLLC).derive()
    P.bc = [O, O, P_, P_, P_, P_]
    G = GridDesc(nx, 6, 6, Lx=1.0, Ly=6.0 / nx, Lz=6.0 / nx).derive()
    x = (np.arange(nx) + 0.5) * G.dx
    shape = (6, 6, nx)
    left = np.broadcast_to(x < 0.5, shape)
    prim = (np.where(left, 1.0, 0.125), np.zeros(shape), np.zeros(shape), np.zeros(shape), np.where(left, 1.0, 0.1))
    say("Sod shock tube, 200 cells, to t = 0.2")
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    dts = dev.time_advance(10000, total_time=0.2)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    rdts = ref.time_advance(10000, -1.0, 0.2, 1e30)
    assert len(dts) == len(rdts)
    assert float(np.max(np.abs(dts - rdts) / rdts)) < 1e-12
    for q, e in state_errors(dev.download, ref.view).items():
        assert e < 1e-10, (q, e)
    rho = dev.interior(dev.download("rho"))[0, 0]
    err = float(np.mean(np.abs(rho - sod_exact(x, 0.2))))
    say(f"  {len(dts)} steps, L1(rho) = {err:.2e}")
    assert err < 6e-3
    dev.close()
    ref.close()


def test_derivatives_refuse_physical_faces(gpu):
    import torch

    from ns3d_wcns_amd.adjoint import state_tensor

    dev, ref = pair(N, BC_SETS["all_outflow"])
    U = state_tensor(dev)
    ptr = [U[q].data_ptr() for q in range(5)]
    with pytest.raises(NS3DError, match="periodic boxes only"):
        dev.adjoint_step(1e-3, ptr)
    with pytest.raises(NS3DError, match="periodic boxes only"):
        dev.tangent_rhs(ptr, [torch.zeros_like(U)[q].data_ptr() for q in range(5)])
    dev.close()
    ref.close()
