"""Pin the CPU checker's scalar kernels (interpolation, Riemann solvers,
characteristic matrices) with closed forms and invariants that a plausible
mistake (wrong weight, sign, index or transposed operand) fails.

What these checks leave open is listed in the checker's header and in the
design notes: the MDCD-hybrid nonlinear weights beyond quadratic exactness and
the branch threshold, and which cell's eigenvectors the characteristic-wise
interpolation uses.
"""
import math

import numpy as np
import pytest

from ns3d_wcns_amd import Interpolation, RiemannSolver, SolverParams
from oracle import ns3d_oracle as orc

G = 1.4


def P(interp=Interpolation.WENO5, rs=RiemannSolver.HLLC, diss=0.02, disp=0.0463783):
    return SolverParams(Ma=0.3, Re=100.0, interpolation=interp, riemann_solver=rs, mdcd_diss=diss,
                        mdcd_disp=disp).derive()


# ---------------------------------------------------------------- interpolation
def _stencils(fun, h, x0s):
    """Point values f(x0 + (m - 2) h), m = 0..5 (cells i-2..i+3); target x0 + h/2."""
    m = np.arange(6) - 2
    return np.array([[fun(x + k * h) for k in m] for x in x0s])


def _cell_averages(F, h, x0s):
    """Exact cell averages of f = F' over [x-h/2, x+h/2] for cells i-2..i+3."""
    m = np.arange(6) - 2
    return np.array([[(F(x + k * h + h / 2) - F(x + k * h - h / 2)) / h for k in m] for x in x0s])


@pytest.mark.parametrize("interp", [Interpolation.WENO5, Interpolation.MDCD_LINEAR, Interpolation.MDCD_HYBRID,
                                    Interpolation.WENO5_JS])
def test_interpolation_preserves_constants(interp):
    st = np.full((5, 6), 3.25) * np.arange(1, 6)[:, None]
    for flag in (1.0, -1.0):
        out = orc.interpolate_batch(P(interp), st, flag)
        assert np.allclose(out, st[:, 0], rtol=1e-15, atol=0)


@pytest.mark.parametrize("interp", [Interpolation.WENO5, Interpolation.MDCD_LINEAR, Interpolation.MDCD_HYBRID])
def test_interpolation_exact_for_quadratics(interp):
    """All WCNS5 candidates (and the MDCD-WENO candidates) are 3-point
    interpolants, exact for quadratics; MDCD-linear is exact for cubics."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b, c = rng.uniform(-3, 3, 3)
        f = lambda x: a * x * x + b * x + c
        st = np.array([[f(m) for m in range(-2, 4)]])
        assert abs(orc.interpolate_batch(P(interp), st, 1.0)[0] - f(0.5)) < 1e-12
        # right state: the mirrored stencil targets the same face
        assert abs(orc.interpolate_batch(P(interp), st, -1.0)[0] - f(0.5)) < 1e-12


def test_mdcd_linear_exact_for_cubics_any_parameters():
    rng = np.random.default_rng(1)
    for _ in range(20):
        diss, disp = rng.uniform(0, 0.05, 2)
        a, b, c, d = rng.uniform(-2, 2, 4)
        f = lambda x: ((a * x + b) * x + c) * x + d
        st = np.array([[f(m) for m in range(-2, 4)]])
        out = orc.interpolate_batch(P(Interpolation.MDCD_LINEAR, diss=diss, disp=disp), st, 1.0)[0]
        assert abs(out - f(0.5)) < 1e-12


def test_mdcd_linear_dissipation_is_the_antisymmetric_part():
    """diss only enters an antisymmetric stencil: with diss = 0 the left and
    right states coincide; the left-right split grows linearly with diss."""
    st = np.array([[0.0, 1.0, 0.0, 1.0, 0.0, 1.0]])  # grid-scale mode
    split = []
    for diss in (0.0, 0.01, 0.02):
        pp = P(Interpolation.MDCD_LINEAR, diss=diss)
        split.append(orc.interpolate_batch(pp, st, 1.0)[0] - orc.interpolate_batch(pp, st, -1.0)[0])
    assert abs(split[0]) < 1e-15
    assert abs(split[1]) > 1e-3
    assert abs(split[2] - 2 * split[1]) < 1e-14


def _order(errs, hs):
    return [math.log(errs[i] / errs[i + 1]) / math.log(hs[i] / hs[i + 1]) for i in range(len(errs) - 1)]


@pytest.mark.parametrize("flag", [1.0, -1.0])
def test_wcns5_fifth_order(flag):
    """Point-value interpolation to x_{i+1/2}: 5th order for smooth data
    (linear weights 1/16, 5/8, 5/16 give the 5-point Lagrange value); a wrong
    weight drops the order to 3."""
    hs = [0.2, 0.1, 0.05]
    errs = []
    for h in hs:
        x0 = np.linspace(0.1, 3.0, 7)
        st = _stencils(np.sin, h, x0)
        out = orc.interpolate_batch(P(Interpolation.WENO5), st, flag)
        errs.append(np.max(np.abs(out - np.sin(x0 + h / 2))))
    assert min(_order(errs, hs)) > 4.5, errs


def test_wcns5_linear_weights_give_lagrange():
    """Pin d_k directly: eps = 1e-6 dominates when every beta_k << 1e-6, so
    tiny-amplitude quartic data returns the 5-point Lagrange interpolant at
    x = 1/2 (the unique linear combination of the candidates exact for
    quartics)."""
    xs = np.arange(-2, 3, dtype=float)
    amp = 1e-8
    for coeffs in ([amp, 0, 0, 0, 0], [0, amp, 0.2 * amp, 0, 0], [0.3 * amp, -0.1 * amp, 0, 0.4 * amp, 0]):
        f = np.polyval(coeffs, np.arange(-2, 4, dtype=float))
        lag = np.polyval(np.polyfit(xs, f[:5], 4), 0.5)
        out = orc.interpolate_batch(P(Interpolation.WENO5), f[None, :], 1.0)[0]
        assert abs(out - lag) < 1e-6 * amp


def test_weno5_js_fifth_order_reconstruction():
    """Jiang-Shu WENO5 reconstructs the interface value from cell averages
    to 5th order."""
    hs = [0.2, 0.1, 0.05]
    errs = []
    for h in hs:
        x0 = np.linspace(0.1, 3.0, 7)
        st = _cell_averages(lambda x: -np.cos(x), h, x0)
        out = orc.interpolate_batch(P(Interpolation.WENO5_JS), st, 1.0)
        errs.append(np.max(np.abs(out - np.sin(x0 + h / 2))))
    assert min(_order(errs, hs)) > 4.5, errs


def test_wcns5_nonoscillatory_at_a_step():
    """At a jump the nonlinear weights select the smooth side: no overshoot."""
    st = np.array([[0, 0, 0, 1, 1, 1], [1, 1, 1, 0, 0, 0]], dtype=float)
    for interp in (Interpolation.WENO5, Interpolation.WENO5_JS):
        out = orc.interpolate_batch(P(interp), st, 1.0)
        assert np.all(out > -1e-4) and np.all(out < 1 + 1e-4)
        assert abs(out[0]) < 1e-3 and abs(out[1] - 1) < 1e-3


def test_mdcd_hybrid_branches():
    """Smooth data: sensor > 0.4 -> exactly the linear MDCD value.  A jump:
    the WENO branch, non-oscillatory."""
    x0 = np.linspace(0.0, 2.0, 9)
    st = _stencils(np.sin, 0.05, x0)
    hyb = orc.interpolate_batch(P(Interpolation.MDCD_HYBRID), st, 1.0)
    lin = orc.interpolate_batch(P(Interpolation.MDCD_LINEAR), st, 1.0)
    assert np.array_equal(hyb, lin)
    step = np.array([[0, 0, 0, 1, 1, 1]], dtype=float)
    h = orc.interpolate_batch(P(Interpolation.MDCD_HYBRID), step, 1.0)[0]
    lnr = orc.interpolate_batch(P(Interpolation.MDCD_LINEAR), step, 1.0)[0]
    assert h != lnr and -1e-3 < h < 1 + 1e-3


# ---------------------------------------------------------------- Riemann solvers
def cons(rho, u, v, w, p, g=G):
    return np.array([rho, rho * u, rho * v, rho * w, p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def phys_flux(U, dim, g=G):
    rho = U[0]
    vel = U[1:4] / rho
    p = (g - 1) * (U[4] - 0.5 * rho * vel @ vel)
    un = vel[dim]
    F = np.array([rho * un, *(rho * vel * un), (U[4] + p) * un])
    F[1 + dim] += p
    return F


def random_states(n, seed, mach=0.5, g=G):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rho = rng.uniform(0.5, 2.0)
        p = rng.uniform(0.5, 2.0)
        c = math.sqrt(g * p / rho)
        u, v, w = rng.uniform(-mach, mach, 3) * c
        out.append(cons(rho, u, v, w, p))
    return np.array(out)


@pytest.mark.parametrize("rs", list(RiemannSolver))
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_riemann_consistency(rs, dim):
    """F(U, U) is the physical flux for every solver."""
    U = random_states(16, 3 + dim)
    F = orc.riemann_batch(P(rs=rs), U, U, dim)
    ref = np.array([phys_flux(u, dim) for u in U])
    assert np.allclose(F, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("rs", [RiemannSolver.HLLC, RiemannSolver.HLL, RiemannSolver.HLLC_p])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_supersonic_upwinding(rs, dim):
    """SL >= 0: the left flux; SR <= 0: the right flux, exactly."""
    UL = random_states(8, 10, mach=0.3)
    UR = random_states(8, 11, mach=0.3)
    for sign in (+1, -1):
        a, b = UL.copy(), UR.copy()
        for U in (a, b):
            rho = U[:, 0]
            p = (G - 1) * (U[:, 4] - 0.5 * np.sum(U[:, 1:4] ** 2, axis=1) / rho)
            c = np.sqrt(G * p / rho)
            du = sign * 3.0 * c  # Mach 3 along the normal
            U[:, 4] += 0.5 * rho * ((U[:, 1 + dim] / rho + du) ** 2 - (U[:, 1 + dim] / rho) ** 2)
            U[:, 1 + dim] += rho * du
        F = orc.riemann_batch(P(rs=rs), a, b, dim)
        src = a if sign > 0 else b
        ref = np.array([phys_flux(u, dim) for u in src])
        assert np.array_equal(F, F) and np.allclose(F, ref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("rs", [RiemannSolver.HLLC, RiemannSolver.HLLC_p])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_hllc_resolves_contacts(rs, dim):
    """A contact (equal normal velocity and pressure, density and tangential
    velocity jump) is an exact solution: HLLC returns the upwind physical
    flux.  HLL smears it, so this pins the star state."""
    rng = np.random.default_rng(20 + dim)
    for _ in range(10):
        un = rng.uniform(-0.5, 0.5)
        p = rng.uniform(0.5, 2)
        vl, vr = np.zeros(3), np.zeros(3)
        vl[dim] = vr[dim] = un
        t = [d for d in range(3) if d != dim]
        vl[t] = rng.uniform(-1, 1, 2)
        vr[t] = rng.uniform(-1, 1, 2)
        UL = cons(rng.uniform(0.5, 1), *vl, p)
        UR = cons(rng.uniform(1.5, 3), *vr, p)
        F = orc.riemann_batch(P(rs=rs), UL[None], UR[None], dim)[0]
        ref = phys_flux(UL if un >= 0 else UR, dim)
        assert np.allclose(F, ref, rtol=1e-12, atol=1e-12), (F, ref)
    # HLL does not keep the contact
    F = orc.riemann_batch(P(rs=RiemannSolver.HLL), UL[None], UR[None], dim)[0]
    assert not np.allclose(F, phys_flux(UL if un >= 0 else UR, dim), rtol=1e-6)


def test_rusanov_is_hll_with_symmetric_speeds():
    """Rusanov = HLL with SL = -smax, SR = smax (textbook identity)."""
    UL, UR = random_states(12, 30), random_states(12, 31)
    F = orc.riemann_batch(P(rs=RiemannSolver.Rusanov), UL, UR, 0)
    for k in range(12):
        fl, fr = phys_flux(UL[k], 0), phys_flux(UR[k], 0)
        s = 0.0
        for U in (UL[k], UR[k]):
            rho = U[0]
            u = U[1] / rho
            p = (G - 1) * (U[4] - 0.5 * np.sum(U[1:4] ** 2) / rho)
            s = max(s, abs(u) + math.sqrt(G * p / rho))
        hll = (s * fl + s * fr - s * s * (UR[k] - UL[k])) / (2 * s)
        assert np.allclose(F[k], hll, rtol=1e-13, atol=1e-13)


def euler_jacobian(U, dim, h=1e-6):
    J = np.zeros((5, 5))
    for c in range(5):
        e = np.zeros(5)
        e[c] = h * max(1.0, abs(U[c]))
        J[:, c] = (phys_flux(U + e, dim) - phys_flux(U - e, dim)) / (2 * e[c])
    return J


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_eigen_matrices_diagonalise_the_jacobian(dim):
    """L R = I and R diag(lambda) L = dF/dU at a single state (Roe average of
    identical states)."""
    for U in random_states(6, 40 + dim):
        rho = U[0]
        vel = U[1:4] / rho
        p = (G - 1) * (U[4] - 0.5 * rho * vel @ vel)
        prim = np.array([rho, *vel, p])
        L, R, lam = orc.eigen(prim, prim, dim, G)
        assert np.allclose(L @ R, np.eye(5), atol=1e-12)
        A = R @ np.diag(lam) @ L
        assert np.allclose(A, euler_jacobian(U, dim), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_roe_property(dim):
    """Roe average: A~ (UR - UL) = FR - FL exactly (Roe's property), which pins
    the averaged state and the eigen-decomposition together; without entropy
    fix the Roe flux is FL + sum_{lambda<0} lambda alpha r."""
    UL, UR = random_states(8, 50 + dim, mach=0.9), random_states(8, 60 + dim, mach=0.9)
    for a, b in zip(UL, UR):
        def prim(U):
            rho = U[0]
            vel = U[1:4] / rho
            return np.array([rho, *vel, (G - 1) * (U[4] - 0.5 * rho * vel @ vel)])
        L, R, lam = orc.eigen(prim(a), prim(b), dim, G)
        A = R @ np.diag(lam) @ L
        assert np.allclose(A @ (b - a), phys_flux(b, dim) - phys_flux(a, dim), rtol=1e-11, atol=1e-11)
        F = orc.riemann_batch(P(rs=RiemannSolver.Roe), a[None], b[None], dim)[0]
        if np.min(np.abs(lam)) >= 0.1 * np.max(np.abs(lam)):  # entropy fix inactive
            alpha = L @ (b - a)
            ref = phys_flux(a, dim) + R @ (np.minimum(lam, 0) * alpha)
            assert np.allclose(F, ref, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("rs", [r for r in RiemannSolver if r != RiemannSolver.AUSM])
def test_riemann_axis_permutation(rs):
    """Permuting the velocity components permutes the fluxes: the y/z faces
    are the x face in a rotated frame.  (AUSM's pressure term uses the x
    velocity on every face, so it is excluded.)"""
    UL, UR = random_states(10, 70), random_states(10, 71)
    Fx = orc.riemann_batch(P(rs=rs), UL, UR, 0)
    for dim in (1, 2):
        perm = [0, 1, 2, 3, 4]
        # put the x velocity into slot 1 + dim, keep the others cyclic
        order = [dim, (dim + 1) % 3, (dim + 2) % 3]
        A, B = UL.copy(), UR.copy()
        for q in range(3):
            A[:, 1 + order[q]] = UL[:, 1 + q]
            B[:, 1 + order[q]] = UR[:, 1 + q]
        Fd = orc.riemann_batch(P(rs=rs), A, B, dim)
        for q in range(3):
            perm[1 + q] = 1 + order[q]
        assert np.allclose(Fd[:, perm], Fx, rtol=1e-13, atol=1e-13)


def test_ausm_uses_x_velocity_difference():
    """The AUSM pressure-diffusion term of the reference reads (u_R - u_L) of
    the x velocity on every face (InviscousFlux.cpp:823): changing only the
    x velocity jump changes the y-face momentum flux."""
    UL = cons(1.0, 0.1, 0.2, 0.0, 1.0)
    UR1 = cons(1.0, 0.1, 0.25, 0.0, 1.0)
    UR2 = cons(1.0, 0.4, 0.25, 0.0, 1.0)
    F1 = orc.riemann_batch(P(rs=RiemannSolver.AUSM), UL[None], UR1[None], 1)[0]
    F2 = orc.riemann_batch(P(rs=RiemannSolver.AUSM), UL[None], UR2[None], 1)[0]
    assert F1[2] != F2[2]


# ---------------------------------------------------------------- face flux
@pytest.mark.parametrize("char", [False, True])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_face_flux_uniform_state(char, dim):
    U = random_states(4, 80 + dim)
    Ut = np.repeat(U[:, :, None], 6, axis=2)
    prim = []
    for u in U:
        rho = u[0]
        vel = u[1:4] / rho
        prim.append([rho, *vel, (G - 1) * (u[4] - 0.5 * rho * vel @ vel)])
    ut = np.repeat(np.array(prim)[:, :, None], 6, axis=2)
    pp = P(Interpolation.WENO5, RiemannSolver.HLLC)
    pp.char_recon = char
    F = orc.face_flux_batch(pp, Ut, ut, dim)
    ref = np.array([phys_flux(u, dim) for u in U])
    assert np.allclose(F, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_characteristic_equals_componentwise_for_linear_schemes(dim):
    """With a linear interpolation the projection L (fixed per side)
    commutes with the interpolation, so R interp(L U) = interp(U)."""
    rng = np.random.default_rng(90 + dim)
    U = random_states(6, 91 + dim)
    Ut = U[:, :, None] * (1 + 0.05 * rng.uniform(-1, 1, (6, 5, 6)))
    ut = np.zeros_like(Ut)
    for n in range(6):
        for m in range(6):
            u = Ut[n, :, m]
            rho = u[0]
            vel = u[1:4] / rho
            ut[n, :, m] = [rho, *vel, (G - 1) * (u[4] - 0.5 * rho * vel @ vel)]
    pa = P(Interpolation.MDCD_LINEAR)
    pb = P(Interpolation.MDCD_LINEAR)
    pb.char_recon = True
    Fa = orc.face_flux_batch(pa, Ut, ut, dim)
    Fb = orc.face_flux_batch(pb, Ut, ut, dim)
    assert np.allclose(Fa, Fb, rtol=1e-12, atol=1e-12)
    # with the nonlinear WCNS5 weights they differ
    pc = P(Interpolation.WENO5)
    pd = P(Interpolation.WENO5)
    pd.char_recon = True
    assert not np.allclose(orc.face_flux_batch(pc, Ut, ut, dim), orc.face_flux_batch(pd, Ut, ut, dim), rtol=1e-10)


def test_sutherland():
    """mu(T) = mu_ref T^1.5 (1 + S)/(T + S); mu(1) = 1/Re."""
    pp = P()
    assert orc.sutherland(pp, 1.0) == pytest.approx(1.0 / 100.0, rel=1e-15)
    S = 110.4 / 288.15
    for T in (0.5, 0.9, 1.3, 2.0):
        assert orc.sutherland(pp, T) == pytest.approx(0.01 * T ** 1.5 * (1 + S) / (T + S), rel=1e-14)
