"""Pin the CPU checker's non-periodic boundaries (boundary.cpp,
ViscousFlux.cpp:15-50) without running the reference.  No GPU.

- ghost fill (apply_boundary, boundary.cpp:14-205): brute force over every
  ghost cell of boxes mixing Wall, Symmetry, Outflow, Inflow and periodic
  axes, against the reference's rules applied face by face in order xmin ..
  zmax over full face planes, after the exchange that zeroes the face ghosts
  of a non-periodic axis;
- ghost faces (apply_boundary_halfnode_flux, boundary.cpp:207-300): zero on
  Wall / Symmetry / Inflow faces, the boundary face on Outflow faces, the
  periodic images on a periodic axis;
- one-sided viscous orders (choose_scheme): gradients and the viscous
  divergence are exact for quadratics on every cell; on cubics the first
  two cells next to a physical face are not and every other cell is; on
  quintics the third cell is not either, and the rest are.  On the last
  interior cell of a physical axis the reference's derivative is -df/dx (its
  backward one-sided difference carries the forward one's sign), and that
  is what is checked there;
- a uniform flow with Outflow faces has zero RHS (free-stream preservation);
- mirror equivariance of the inviscid boundary treatment: with the same
  physical type on both faces of an axis, mirroring a nearly inviscid state
  (Re = 1e14) about the box centre (normal velocity negated) mirrors the RHS;
- Sod's shock tube with Outflow x faces (initialize_sod_shock_tube,
  initialization.cpp:570-599) at t = 0.2 against the exact Riemann solution
  (Toro, ch. 4): L1 error of the density below 6e-3 on 200 cells.
"""
import math

import numpy as np
import pytest

from cases import box, params, state
from ns3d_wcns_amd import Interpolation, RiemannSolver, SolverParams
from ns3d_wcns_amd.params import BCType
from oracle.ns3d_oracle import NG, OracleField

W, S, O, I, P_ = BCType.Wall, BCType.Symmetry, BCType.Outflow, BCType.Inflow, BCType.Periodic
PRIMS = ("rho", "u", "v", "w", "p")
CONS = ("rho", "rhou", "rhov", "rhow", "E")
RHS = ("rhs_rho", "rhs_rhou", "rhs_rhov", "rhs_rhow", "rhs_E")

BC_SETS = {
    "wall_outflow_x": [W, O, P_, P_, P_, P_],
    "symmetry_y_inflow_z": [P_, P_, S, S, I, O],
    "all_outflow": [O] * 6,
    "mixed": [I, O, W, S, O, W],
}


def bounded(n, bc, mach_t=0.3, noise=0.02, seed=4, **kw):
    prim, ma = state(n, mach_t, noise=noise, seed=seed)
    P = params(ma, **kw)
    P.bc = list(bc)
    F = OracleField(box(n), P)
    F.initialize_primitive(*prim)
    return F, prim, P


# ---------------------------------------------------------------- ghost fill
def expected_fill(F, bc, vals):
    """The reference's apply_boundary on encoded values, restated face by face."""
    out = {k: v.copy() for k, v in vals.items()}
    s = (F.sx, F.sy, F.sz)
    n = (F.nx, F.ny, F.nz)
    per = [bc[2 * a] == P_ and bc[2 * a + 1] == P_ for a in range(3)]
    # exchange: face ghosts over the interior tangential range
    for k in range(F.sz):
        for j in range(F.sy):
            for i in range(F.sx):
                c = (i, j, k)
                inside = [NG <= c[a] < NG + n[a] for a in range(3)]
                if sum(inside) != 2:
                    continue
                a = inside.index(False)
                src = list(c)
                src[a] = (c[a] - NG) % n[a] + NG
                for q in PRIMS:
                    out[q][k, j, i] = out[q][src[2], src[1], src[0]] if per[a] else 0.0
    # physical faces, in order, over full planes
    for face in range(6):
        a, hi, t = face // 2, face & 1, bc[face]
        if per[a] or t == P_:
            continue
        rng = range(s[a] - NG, s[a]) if hi else range(NG)
        for g in rng:
            sl = [slice(None)] * 3
            sl[2 - a] = g  # arrays are (z, y, x)
            src = [slice(None)] * 3
            if t == I:
                for q, v in zip(PRIMS, (1.0, 1.0, 0.0, 0.0, 1.0)):
                    out[q][tuple(sl)] = v
                continue
            if t == O:
                src[2 - a] = s[a] - NG - 1 if hi else NG
            else:
                src[2 - a] = 2 * (s[a] - NG) - 1 - g if hi else 2 * NG - 1 - g
            for qi, q in enumerate(PRIMS):
                v = out[q][tuple(src)].copy()
                out[q][tuple(sl)] = -v if (t in (W, S) and qi == 1 + a) else v
    return out


@pytest.mark.parametrize("name", list(BC_SETS))
def test_boundary_fill_brute_force(name):
    bc = BC_SETS[name]
    F, _, _ = bounded((12, 10, 8), bc)
    rng = np.random.default_rng(3)
    vals = {q: rng.standard_normal(F.shape) + 10 * qi for qi, q in enumerate(PRIMS)}
    for q in PRIMS:
        F.view(q)[...] = vals[q]
    F.apply_boundary()
    want = expected_fill(F, bc, vals)
    for q in PRIMS:
        assert np.array_equal(F.view(q), want[q]), (name, q)
    F.close()


# ---------------------------------------------------------------- ghost faces
@pytest.mark.parametrize("name", list(BC_SETS))
def test_ghost_faces(name):
    bc = BC_SETS[name]
    F, _, _ = bounded((12, 10, 8), bc)
    F.compute_invis_flux()
    F.apply_boundary_halfnode_flux()
    n = (F.nx, F.ny, F.nz)
    shapes = [(F.sz, F.sy, F.sx - 1), (F.sz, F.sy - 1, F.sx), (F.sz - 1, F.sy, F.sx)]
    for d, tag in enumerate(("fx", "fy", "fz")):
        per = bc[2 * d] == P_
        for comp in ("mass", "momx", "momy", "momz", "E"):
            A = F.view(f"flux_{tag}_{comp}").reshape(shapes[d])
            A = np.moveaxis(A, 2 - d, 0)  # face index first
            inner = [slice(NG, NG + n[a]) for a in range(3) if a != d][::-1]  # tangential interior (slower first)
            B = A[(slice(None),) + tuple(inner)]
            for side in (0, 1):
                ghosts = (0, 1) if side == 0 else (NG + n[d], NG + n[d] + 1)
                bface = NG - 1 if side == 0 else NG + n[d] - 1
                for q in ghosts:
                    if per:
                        want = B[q + n[d]] if side == 0 else B[q - n[d]]
                    elif bc[2 * d + side] == O:
                        want = B[bface]
                    else:
                        want = np.zeros_like(B[q])
                    assert np.array_equal(B[q], want), (name, tag, comp, q)
    F.close()


# ---------------------------------------------------------------- one-sided viscous orders
def poly_field(F, axis, coeffs):
    h = (F.grid.dx, F.grid.dy, F.grid.dz)[axis]
    x = (np.arange((F.sx, F.sy, F.sz)[axis]) - NG + 0.5) * h
    f = sum(c * x ** m for m, c in enumerate(coeffs))
    df = sum(m * c * x ** (m - 1) for m, c in enumerate(coeffs) if m > 0)
    shape = [1, 1, 1]
    shape[2 - axis] = -1
    return np.broadcast_to(f.reshape(shape), F.shape).copy(), np.broadcast_to(df.reshape(shape), F.shape).copy()


def sign_last(d, axis):
    """The reference's derivative on the last interior cell of a physical
    axis is -df/dx: choose_scheme returns -1 there and diff_2nd_forward(f, i,
    h, -1) = (-3 f_i + 4 f_{i-1} - f_{i-2}) / (2 h), the backward difference
    with its sign flipped (ViscousFlux.cpp:15-17, 38-40)."""
    d = d.copy()
    idx = [slice(None)] * 3
    idx[2 - axis] = -1
    d[tuple(idx)] *= -1.0
    return d


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_one_sided_gradient_orders(axis):
    bc = [P_] * 6
    bc[2 * axis] = W
    bc[2 * axis + 1] = O
    F, _, _ = bounded((12, 12, 12), bc)
    n = 12
    for degree, bad in ((2, set()), (3, {0, 1, n - 2, n - 1}), (5, {0, 1, 2, n - 3, n - 2, n - 1})):
        coeffs = [0.3 * (m + 1) * (-1) ** m for m in range(degree + 1)]
        f, df = poly_field(F, axis, coeffs)
        F.view("u")[...] = f
        F.compute_gradients()
        g = F.interior(F.view(("du_dx", "du_dy", "du_dz")[axis]))
        want = sign_last(F.interior(df), axis)
        e = np.abs(g - want)
        e = np.moveaxis(e, 2 - axis, 0).reshape(n, -1).max(axis=1)
        scale = np.abs(df).max()
        for c in range(n):
            if c in bad:
                assert e[c] > 1e-8 * scale, (degree, c, e[c])
            else:
                assert e[c] < 1e-10 * scale, (degree, c, e[c])
        # the viscous divergence uses the same stencils (ViscousFlux.cpp:210-255)
        for name in RHS:
            F.view(name)[...] = 0.0
        F.view(("Fvflux_momx", "Hvflux_momx", "Gvflux_momx")[axis])[...] = f
        F.compute_vis_flux()
        r = np.moveaxis(np.abs(F.interior(F.view("rhs_rhou")) - want), 2 - axis, 0).reshape(n, -1).max(1)
        for c in range(n):
            assert (r[c] > 1e-8 * scale) if c in bad else (r[c] < 1e-10 * scale), (degree, c, r[c])
    F.close()


# ---------------------------------------------------------------- invariants
def test_uniform_flow_with_outflow_faces_has_zero_rhs():
    n = (12, 10, 8)
    P = params(0.3)
    P.bc = [O] * 6
    F = OracleField(box(n), P)
    F.initialize_primitive(np.full(n[::-1], 1.3), np.full(n[::-1], 0.4), np.full(n[::-1], -0.2),
                           np.full(n[::-1], 0.1), np.full(n[::-1], 20.0))
    F.compute_rhs()
    for name in RHS:
        assert np.max(np.abs(F.interior(F.view(name)))) < 1e-12, name
    F.close()


def mirror(a, axis):
    return np.flip(a, axis=2 - axis).copy()


@pytest.mark.parametrize("axis,bctype", [(0, S), (1, W), (2, O)])
@pytest.mark.parametrize("interp,rs", [(Interpolation.WENO5, RiemannSolver.HLLC),
                                       (Interpolation.MDCD_HYBRID, RiemannSolver.Rusanov),
                                       (Interpolation.WENO5_JS, RiemannSolver.Roe)], ids=lambda x: x.name)
def test_mirror_equivariance(axis, bctype, interp, rs):
    n = (12, 10, 8)
    bc = [P_] * 6
    bc[2 * axis] = bc[2 * axis + 1] = bctype
    # nearly inviscid: the reference's viscous terms are not mirror symmetric
    # next to a physical face (the sign of its last-cell derivative)
    F, prim, P = bounded(n, bc, interp=interp, rs=rs, re=1e14)
    F.compute_rhs()
    r1 = {q: F.interior(F.view(q)).copy() for q in RHS}
    vel = ("u", "v", "w")[axis]
    m = [mirror(a, axis) for a in prim]
    m[1 + axis] = -m[1 + axis]
    G = OracleField(box(n), P)
    G.initialize_primitive(*m)
    G.compute_rhs()
    for q in RHS:
        want = mirror(r1[q], axis)
        if q == ("rhs_rhou", "rhs_rhov", "rhs_rhow")[axis]:
            want = -want
        got = G.interior(G.view(q))
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), (q, vel)
    F.close()
    G.close()


# ---------------------------------------------------------------- Sod's shock tube
def sod_exact(x, t, x0=0.5, g=1.4):
    """Density of the exact solution of Sod's problem (Toro, ch. 4)."""
    rl, ul, pl, rr, ur, pr = 1.0, 0.0, 1.0, 0.125, 0.0, 0.1
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)

    def fk(p, rk, pk, ak):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-8, 5.0
    for _ in range(200):  # bisection on f(p) = fL + fR + du
        mid = 0.5 * (lo + hi)
        if fk(mid, rl, pl, al) + fk(mid, rr, pr, ar) + (ur - ul) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, ar) - fk(ps, rl, pl, al))
    rsl = rl * (ps / pl) ** (1 / g)  # rarefaction on the left
    rsr = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))  # shock on the right
    asl = al * (ps / pl) ** ((g - 1) / (2 * g))
    sh = ur + ar * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
    out = np.empty_like(x)
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s < ul - al:
            out[i] = rl
        elif s < us - asl:
            c = 2 / (g + 1) * (al + (g - 1) / 2 * (ul - s))
            out[i] = rl * (c / al) ** (2 / (g - 1))
        elif s < us:
            out[i] = rsl
        elif s < sh:
            out[i] = rsr
        else:
            out[i] = rr
    return out


def test_sod_shock_tube_against_exact_solution():
    nx = 200
    P = SolverParams(Ma=0.5, Re=1e8, Pr=0.72, cfl=0.5, interpolation=Interpolation.WENO5,
                     riemann_solver=RiemannSolver.HLLC).derive()
    P.bc = [O, O, P_, P_, P_, P_]
    from ns3d_wcns_amd import GridDesc

    G = GridDesc(nx, 6, 6, Lx=1.0, Ly=6.0 / nx, Lz=6.0 / nx).derive()
    F = OracleField(G, P)
    x = (np.arange(nx) + 0.5) * G.dx
    shape = (6, 6, nx)
    left = np.broadcast_to(x < 0.5, shape)
    F.initialize_primitive(np.where(left, 1.0, 0.125), np.zeros(shape), np.zeros(shape), np.zeros(shape),
                           np.where(left, 1.0, 0.1))
    dts = F.time_advance(10000, -1.0, 0.2, 1e30)
    assert abs(dts.sum() - 0.2) < 1e-12
    rho = F.interior(F.view("rho"))
    assert np.ptp(rho, axis=(0, 1)).max() < 1e-12  # stays one-dimensional
    err = np.mean(np.abs(rho[0, 0] - sod_exact(x, 0.2)))
    print(f"Sod: {len(dts)} steps, L1(rho) = {err:.2e}")
    assert err < 6e-3
    F.close()
