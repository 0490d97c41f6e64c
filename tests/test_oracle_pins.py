"""Checks of the CPU checker's scheme internals against what the mathematics
fixes, through the checker's test hooks (orc_mdcd_parts, orc_hllc_parts,
orc_ausm_split).  No GPU.

MDCD hybrid (interpolation.cpp:135-209):
- beta0..beta2 are the Jiang-Shu indicators of the three 3-cell candidates
  (the definition sum_l h^(2l-1) int (p^(l))^2 over the stencil's third
  cell, p the polynomial with the given cell averages), derived here
  symbolically and compared bit for bit on integer stencils;
- beta3 is the 6-cell indicator from the same definition, plus the
  reference's extra 20 f0^2 / 120960 (its f0^2 coefficient reads 271799
  where the definition gives 271779), bit for bit on integer stencils;
- q0..q3 are exact for quadratics on their own three cells (which fixes a
  3-point midpoint interpolant), tau6 is |beta3 - (beta0 + 4 beta1 +
  beta2) / 6| (the 6th-order reference indicator of WENO-CU6, Hu, Wang &
  Adams 2010), and the branch value is sum c_k q_k / sum c_k with
  c_k = d_k (20 + tau6 / beta_k)^2 (C = 20, q = 2 of the same family);
- the linear weights: sum d_k = 1 for any (disp, diss), central
  (0, 1/2, 1/2, 0) at disp = diss = 0, mirror symmetric (d3(disp, diss) =
  d0(disp, -diss), d2(disp, diss) = d1(disp, -diss)) and linear in (disp,
  diss).  Their slopes (3/2 and 9/2) are the published MDCD-WENO weights:
  with the finite-volume (WENO-JS) candidates they reproduce the six-point
  linear MDCD reconstruction identically (Sun et al. 2011), which has no
  other solution.  With the WCNS candidates the reference uses,
  the smooth limit sum d_k q_k is then exact only through quadratics
  (cubic error -9 diss / 4), unlike the linear branch mdcd_linear (exact
  through cubics); the weights that would reproduce mdcd_linear are
  (disp + diss, 1/2 - disp + 5 diss, 1/2 - disp - 5 diss, disp - diss).
HLLC and HLLC_p (InviscousFlux.cpp:449-717), on generic subsonic states:
- Davis wave speeds (HLLC) and Toro's PVRS pressure with shock factors
  (HLLC_p), computed here from the textbook formulas;
- the star states satisfy the integral consistency condition
  (SM - SL) U*L + (SR - SM) U*R = SR UR - SL UL - (FR - FL), i.e. the
  Rankine-Hugoniot conditions across all three waves, with F from the
  Euler flux computed here; both star states move with SM and keep the
  tangential velocity of their side, and the contact pressures
  p_K + rho_K (S_K - u_K)(SM - u_K) seen from the two sides agree;
- the flux is F*L when SM >= 0 and F*R otherwise;
- |den| <= 1e-12 (near-vacuum densities) falls back to the Roe-averaged
  normal velocity.
Roe (InviscousFlux.cpp:313-382) with Harten's entropy fix at sonic and
  stagnation states: 1/2 (FL + FR) - 1/2 |A|_fix (UR - UL), A the Euler
  Jacobian (derived symbolically here) at the Roe-averaged state, |A|_fix
  formed from its spectral projectors (eigenvalues from numpy) with |lambda|
  replaced by (lambda^2 + delta^2) / (2 delta) below delta = 0.1 max |lambda|;
  independent of the checker's eigenvector matrices.
AUSM (InviscousFlux.cpp:719-846) split functions: M+ + M- = M and
  P+ + P- = 1 for every M, mirror symmetry, continuity at |M| = 1, and
  Liou's polynomials 1/4 (M +- 1)^2 and 1/4 (M +- 1)^2 (2 -+ M) inside
  (the beta = 0, alpha = 0 members of the AUSM+ family, although the
  reference names them M4 and P5).
"""
import ctypes
import functools
from fractions import Fraction

import numpy as np
import pytest
import sympy as sp

from ns3d_wcns_amd import Interpolation, SolverParams
from oracle.ns3d_oracle import load, riemann_batch

DP = ctypes.POINTER(ctypes.c_double)
GAMMA = 1.4


def _dp(a):
    return a.ctypes.data_as(DP)


def _lib():
    lib = load()
    from ns3d_wcns_amd.params import CParams

    lib.orc_mdcd_parts.argtypes = [DP, ctypes.POINTER(CParams), DP]
    lib.orc_hllc_parts.argtypes = [ctypes.c_double, DP, DP, ctypes.c_int, ctypes.c_int, DP]
    lib.orc_ausm_split.argtypes = [ctypes.c_double, DP]
    return lib


def mdcd_parts(f, disp=0.0463783, diss=0.02):
    P = SolverParams(Ma=0.3, Re=100.0, interpolation=Interpolation.MDCD_HYBRID, mdcd_disp=disp,
                     mdcd_diss=diss).derive()
    cp = P.to_c()
    s = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(16)
    _lib().orc_mdcd_parts(_dp(s), ctypes.byref(cp), _dp(out))
    return {"sai": out[0], "beta": out[1:5], "tau6": out[5], "d": out[6:10], "q": out[10:14], "weno": out[14],
            "value": out[15]}


# ---------------------------------------------------------------- indicators from their definition
X = sp.Symbol("x")
F = sp.symbols("f0:6")


@functools.lru_cache(maxsize=None)
def js_indicator(first: int, count: int):
    """sum_{l=1}^{count-1} int_{cell 2} (p^(l))^2 dx for the polynomial p of
    degree count-1 whose averages over cells first..first+count-1 (cell j
    is [j - 5/2, j - 3/2] for j = 0..5, so cell 2 is [-1/2, 1/2]) are
    f_first..; unit spacing."""
    c = sp.symbols(f"c0:{count}")
    p = sum(c[k] * X ** k for k in range(count))
    eqs = [sp.integrate(p, (X, j - 2 - sp.Rational(1, 2), j - 2 + sp.Rational(1, 2))) - F[j]
           for j in range(first, first + count)]
    P = p.subs(sp.solve(eqs, c))
    return sp.expand(sum(sp.integrate(sp.diff(P, X, l) ** 2, (X, -sp.Rational(1, 2), sp.Rational(1, 2)))
                         for l in range(1, count)))


def exact(expr, f):
    """Value of a rational polynomial at integer f, correctly rounded."""
    v = sp.Rational(expr.subs({F[k]: int(f[k]) for k in range(6)}))
    return float(Fraction(int(v.p), int(v.q)))


def integer_stencils(n=200, seed=3):
    rng = np.random.default_rng(seed)
    return rng.integers(-40, 41, size=(n, 6)).astype(np.float64)


def test_three_cell_indicators_from_definition():
    """13/12 is not a binary fraction, so the checker's sums carry a few
    ulp: compared to 4 ulp of the value."""
    defs = [js_indicator(k, 3) for k in range(3)]
    for f in integer_stencils(100):
        got = mdcd_parts(f)["beta"]
        for k in range(3):
            want = exact(defs[k], f)
            assert abs(got[k] - want) <= 4 * np.spacing(max(abs(want), 1.0)), (k, f, got[k], want)


def test_six_cell_indicator_is_the_definition_plus_the_reference_f0_term():
    beta6 = js_indicator(0, 6)
    # the definition's f0^2 coefficient, times 120960: 271779 (the reference writes 271799)
    assert sp.Poly(beta6, *F).coeff_monomial(F[0] ** 2) * 120960 == 271779
    ref = beta6 + sp.Rational(20, 120960) * F[0] ** 2
    for f in integer_stencils(200):
        got = mdcd_parts(f)["beta"][3]
        assert got == exact(ref, f), (f, got, exact(ref, f))
    # constant data: the definition gives 0, the reference c^2 / 6048
    for c in (1.0, 7.0, -3.0):
        assert mdcd_parts(np.full(6, c))["beta"][3] == pytest.approx(c * c / 6048.0, rel=1e-14)


def test_candidates_exact_for_quadratics_on_their_cells():
    rng = np.random.default_rng(5)
    x = np.arange(6) - 2.0  # cell centres; the face is x = 1/2
    for _ in range(50):
        a, b, c = rng.standard_normal(3)
        # point values of a quadratic: the WCNS candidates interpolate nodal values
        f = a + b * x + c * x * x
        q = mdcd_parts(f)["q"]
        face = a + b * 0.5 + c * 0.25
        assert np.allclose(q, face, rtol=0, atol=1e-13 * (1 + abs(face)))
    # and each uses only its own three cells: perturb a cell outside its stencil
    base = rng.standard_normal(6)
    q0 = mdcd_parts(base)["q"]
    for k in range(4):
        for j in range(6):
            if k <= j <= k + 2:
                continue
            g = base.copy()
            g[j] += 1.0
            assert mdcd_parts(g)["q"][k] == q0[k], (k, j)


def test_branch_value_from_tau6_and_weights():
    rng = np.random.default_rng(7)
    for _ in range(100):
        f = np.cumsum(rng.standard_normal(6)) + 10 * (rng.random() < 0.5) * (np.arange(6) >= 3)
        m = mdcd_parts(f)
        b = m["beta"]
        assert m["tau6"] == pytest.approx(abs(b[3] - (b[0] + 4 * b[1] + b[2]) / 6.0), rel=1e-14, abs=1e-14)
        c = m["d"] * (20.0 + m["tau6"] / (b + 1e-40)) ** 2
        assert m["weno"] == pytest.approx(float(np.dot(c, m["q"]) / c.sum()), rel=1e-13, abs=1e-13)
        assert m["value"] == (m["weno"] if m["sai"] <= 0.4 else m["value"])


def test_linear_weights_structure():
    f = np.arange(6.0)
    rng = np.random.default_rng(9)
    for _ in range(20):
        disp, diss = rng.uniform(-0.1, 0.1, 2)
        d = mdcd_parts(f, disp, diss)["d"]
        dm = mdcd_parts(f, disp, -diss)["d"]
        assert d.sum() == pytest.approx(1.0, abs=1e-15)
        assert d[3] == pytest.approx(dm[0], abs=1e-16) and d[2] == pytest.approx(dm[1], abs=1e-16)
        # linear in (disp, diss)
        d1 = mdcd_parts(f, 2 * disp, 2 * diss)["d"]
        d0 = mdcd_parts(f, 0.0, 0.0)["d"]
        assert np.allclose(d1 - d0, 2 * (d - d0), rtol=0, atol=1e-15)
    assert np.array_equal(mdcd_parts(f, 0.0, 0.0)["d"], [0.0, 0.5, 0.5, 0.0])


def test_linear_weights_are_the_published_mdcd_weno_weights():
    """The slopes of d_k in (disp, diss) are fixed by the scheme's definition
    (Sun, Ren, Zhang, Jiang, J. Comput. Phys. 230 (2011) 4616): the MDCD-WENO
    linear weights are those for which the four finite-volume (WENO-JS)
    candidates, combined with them, give the six-point linear MDCD
    reconstruction
      (disp/2 + diss/2) f0 + (-3 disp/2 - 5 diss/2 - 1/12) f1 + (7/12 + disp + 5 diss) f2
      + (7/12 + disp - 5 diss) f3 + (-3 disp/2 + 5 diss/2 - 1/12) f4 + (disp/2 - diss/2) f5
    for every stencil and every (disp, diss).  That identity has the one
    solution slopes (3/2, 9/2) among the consistent, central, mirror
    symmetric linear weights (sympy, below), so checking it on the oracle's
    d_k, exactly in rational arithmetic, pins them.  (The reference then pairs
    them with the WCNS point-value candidates, which is the quirk recorded
    with the smooth-limit test.)"""
    R = sp.Rational
    disp_s, diss_s, a, b = sp.symbols("disp diss a b")
    q_fv = [(2 * F[0] - 7 * F[1] + 11 * F[2]) / 6, (-F[1] + 5 * F[2] + 2 * F[3]) / 6,
            (2 * F[2] + 5 * F[3] - F[4]) / 6, (11 * F[3] - 7 * F[4] + 2 * F[5]) / 6]
    linear = ((disp_s / 2 + diss_s / 2) * F[0] + (-R(3, 2) * disp_s - R(5, 2) * diss_s - R(1, 12)) * F[1]
              + (R(7, 12) + disp_s + 5 * diss_s) * F[2] + (R(7, 12) + disp_s - 5 * diss_s) * F[3]
              + (-R(3, 2) * disp_s + R(5, 2) * diss_s - R(1, 12)) * F[4] + (disp_s / 2 - diss_s / 2) * F[5])
    family = [a * (disp_s + diss_s), R(1, 2) - a * disp_s + b * diss_s, R(1, 2) - a * disp_s - b * diss_s,
              a * (disp_s - diss_s)]
    eqs = sp.Poly(sp.expand(sum(w * q for w, q in zip(family, q_fv)) - linear), *F, disp_s, diss_s).coeffs()
    assert sp.solve(eqs, [a, b], dict=True) == [{a: R(3, 2), b: R(9, 2)}]
    f = np.arange(6.0)
    for disp, diss in ((0.0463783, 0.02), (0.03125, 0.015625), (-0.0625, 0.0078125)):
        d = [sp.nsimplify(float(x), rational=True) for x in mdcd_parts(f, disp, diss)["d"]]
        dv, sv = sp.nsimplify(disp, rational=True), sp.nsimplify(diss, rational=True)
        got = sp.expand(sum(w * q for w, q in zip(d, q_fv)))
        want = sp.expand(linear.subs({disp_s: dv, diss_s: sv}))
        for m in range(6):  # coefficient of each cell, to the weights' rounding
            assert abs(float(got.coeff(F[m]) - want.coeff(F[m]))) < 1e-15, (disp, diss, m)


def test_smooth_limit_order_of_the_branch():
    """sum d_k q_k on monomials x^m (nodes -2..3, face 1/2): exact through
    m = 2; the cubic error is -9 diss / 4 (0 only without dissipation)."""
    for disp, diss in ((0.0463783, 0.02), (0.0463783, 0.0), (0.0, 0.01)):
        x = np.arange(6) - 2.0
        for m in range(4):
            parts = mdcd_parts(x ** m, disp, diss)
            err = float(np.dot(parts["d"], parts["q"])) - 0.5 ** m
            want = -9.0 * diss / 4.0 if m == 3 else 0.0
            assert err == pytest.approx(want, abs=1e-14), (disp, diss, m, err)


# ---------------------------------------------------------------- HLLC and HLLC_p
def conserved(rho, u, p):
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2], p / (GAMMA - 1.0) + 0.5 * rho * float(np.dot(u, u))])


def primitive(U):
    rho = U[0]
    u = U[1:4] / rho
    p = (GAMMA - 1.0) * (U[4] - 0.5 * rho * float(np.dot(u, u)))
    return rho, u, p


def euler_flux(U, dim):
    rho, u, p = primitive(U)
    un = u[dim]
    F = np.array([rho * un, rho * u[0] * un, rho * u[1] * un, rho * u[2] * un, (U[4] + p) * un])
    F[1 + dim] += p
    return F


def hllc_parts(UL, UR, dim, pressure_based):
    out = np.zeros(19)
    _lib().orc_hllc_parts(GAMMA, _dp(np.ascontiguousarray(UL)), _dp(np.ascontiguousarray(UR)), dim,
                          1 if pressure_based else 0, _dp(out))
    return {"SL": out[0], "SR": out[1], "SM": out[2], "den": out[3], "UsL": out[4:9], "UsR": out[9:14],
            "F": out[14:19]}


def random_subsonic_pair(rng):
    while True:
        rhoL, rhoR = rng.uniform(0.3, 3.0, 2)
        pL, pR = rng.uniform(0.3, 5.0, 2)
        uL, uR = rng.uniform(-0.8, 0.8, 3), rng.uniform(-0.8, 0.8, 3)
        UL, UR = conserved(rhoL, uL, pL), conserved(rhoR, uR, pR)
        yield UL, UR


def toro_pvrs_speeds(UL, UR, dim):
    rL, uL, pL = primitive(UL)
    rR, uR, pR = primitive(UR)
    aL, aR = np.sqrt(GAMMA * pL / rL), np.sqrt(GAMMA * pR / rR)
    ps = max(0.0, 0.5 * (pL + pR) - 0.125 * (uR[dim] - uL[dim]) * (rL + rR) * (aL + aR))
    qk = lambda p: 1.0 if ps <= p else np.sqrt(1.0 + (GAMMA + 1.0) / (2.0 * GAMMA) * (ps / p - 1.0))  # noqa: E731
    return uL[dim] - aL * qk(pL), uR[dim] + aR * qk(pR), ps


@pytest.mark.parametrize("pressure_based", [False, True], ids=["HLLC", "HLLC_p"])
def test_hllc_star_states_satisfy_rankine_hugoniot(pressure_based):
    rng = np.random.default_rng(11)
    gen = random_subsonic_pair(rng)
    checked = 0
    while checked < 300:
        UL, UR = next(gen)
        dim = checked % 3
        m = hllc_parts(UL, UR, dim, pressure_based)
        rL, uL, pL = primitive(UL)
        rR, uR, pR = primitive(UR)
        aL, aR = np.sqrt(GAMMA * pL / rL), np.sqrt(GAMMA * pR / rR)
        if pressure_based:
            SL, SR, _ = toro_pvrs_speeds(UL, UR, dim)
        else:  # Davis estimates
            SL, SR = min(uL[dim] - aL, uR[dim] - aR), max(uL[dim] + aL, uR[dim] + aR)
        assert m["SL"] == pytest.approx(SL, rel=1e-14, abs=1e-14)
        assert m["SR"] == pytest.approx(SR, rel=1e-14, abs=1e-14)
        if not (m["SL"] < 0.0 < m["SR"]):
            continue
        SM, UsL, UsR = m["SM"], m["UsL"], m["UsR"]
        FL, FR = euler_flux(UL, dim), euler_flux(UR, dim)
        lhs = (SM - m["SL"]) * UsL + (m["SR"] - SM) * UsR
        rhs = m["SR"] * UR - m["SL"] * UL - (FR - FL)
        scale = np.abs(m["SR"] * UR).max() + np.abs(m["SL"] * UL).max() + np.abs(FR).max() + np.abs(FL).max()
        assert np.max(np.abs(lhs - rhs)) <= 1e-13 * scale, (lhs, rhs)
        for Us, U in ((UsL, UL), (UsR, UR)):
            r, u, p = primitive(Us)
            _, u0, _ = primitive(U)
            assert u[dim] == pytest.approx(SM, rel=1e-13, abs=1e-13)
            t = [c for c in range(3) if c != dim]
            assert np.allclose(u[t], u0[t], rtol=1e-13, atol=1e-13)
        # contact pressure seen from either side (Toro's p*_K = p_K + rho_K (S_K - u_K)(S* - u_K)):
        # equal on both sides exactly when SM balances the two
        psL = pL + rL * (m["SL"] - uL[dim]) * (SM - uL[dim])
        psR = pR + rR * (m["SR"] - uR[dim]) * (SM - uR[dim])
        assert psL == pytest.approx(psR, rel=1e-12)
        want = FL + m["SL"] * (UsL - UL) if SM >= 0 else FR + m["SR"] * (UsR - UR)
        assert np.allclose(m["F"], want, rtol=1e-14, atol=1e-14 * np.abs(want).max())
        checked += 1


def test_hllc_p_pvrs_clips_at_zero_and_uses_acoustic_speeds():
    # two strong rarefactions: the PVRS pressure is negative, clipped to 0,
    # and both speeds are the acoustic ones
    UL = conserved(1.0, np.array([-3.0, 0.1, 0.0]), 1.0)
    UR = conserved(1.0, np.array([3.0, 0.0, -0.2]), 1.0)
    SL, SR, ps = toro_pvrs_speeds(UL, UR, 0)
    assert ps == 0.0
    m = hllc_parts(UL, UR, 0, True)
    a = np.sqrt(GAMMA)
    assert m["SL"] == pytest.approx(-3.0 - a) and m["SR"] == pytest.approx(3.0 + a)


def test_hllc_near_vacuum_falls_back_to_roe_velocity():
    rng = np.random.default_rng(13)
    for _ in range(50):
        rho = rng.uniform(1e-14, 5e-14, 2)
        u = rng.uniform(-0.5, 0.5, (2, 3))
        p = rho * rng.uniform(0.5, 1.0, 2) / GAMMA  # a ~ 1
        UL, UR = conserved(rho[0], u[0], p[0]), conserved(rho[1], u[1], p[1])
        dim = int(rng.integers(3))
        m = hllc_parts(UL, UR, dim, False)
        assert abs(m["den"]) <= 1e-12
        sL, sR = np.sqrt(rho)
        roe = (sL * u[0][dim] + sR * u[1][dim]) / (sL + sR)
        assert m["SM"] == pytest.approx(roe, rel=1e-13, abs=1e-15)
    # and never for physical densities: den = rhoR (SR - unR) - rhoL (SL - unL) > 0
    gen = random_subsonic_pair(np.random.default_rng(14))
    for _ in range(100):
        UL, UR = next(gen)
        assert hllc_parts(UL, UR, 0, False)["den"] > 1e-3


# ---------------------------------------------------------------- Roe with Harten's entropy fix
@functools.lru_cache(maxsize=None)
def euler_jacobian(dim):
    """d F_dim / d U of the 3-D Euler flux for gamma = GAMMA, as a function of U."""
    U = sp.symbols("U0:5")
    rho = U[0]
    u = [U[1] / rho, U[2] / rho, U[3] / rho]
    p = (sp.Rational(7, 5) - 1) * (U[4] - sp.Rational(1, 2) * rho * sum(c * c for c in u))
    un = u[dim]
    Fl = [rho * un, U[1] * un, U[2] * un, U[3] * un, (U[4] + p) * un]
    Fl[1 + dim] += p
    J = sp.Matrix(Fl).jacobian(sp.Matrix(U))
    return sp.lambdify(U, J, "numpy")


def roe_textbook(UL, UR, dim):
    rL, uL, pL = primitive(UL)
    rR, uR, pR = primitive(UR)
    HL, HR = (UL[4] + pL) / rL, (UR[4] + pR) / rR
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    u = (sL * uL + sR * uR) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    rho = sL * sR
    # the Roe matrix of an ideal gas is the Jacobian at the Roe-averaged state
    a2 = (GAMMA - 1.0) * (H - 0.5 * float(np.dot(u, u)))
    p = rho * a2 / GAMMA
    Ut = conserved(rho, u, p)
    A = np.array(euler_jacobian(dim)(*Ut), dtype=np.float64)
    lam = np.sort(np.linalg.eigvals(A).real)
    delta = 0.1 * np.max(np.abs(lam))
    g = lambda x: (x * x + delta * delta) / (2.0 * delta) if abs(x) < delta else abs(x)  # noqa: E731
    # A is diagonalisable with three distinct eigenvalues (u_n - a, u_n triple,
    # u_n + a): g(A) = sum_i g(l_i) prod_{j != i} (A - l_j) / (l_i - l_j)
    distinct = [lam[0], float(np.median(lam)), lam[-1]]
    absA = np.zeros((5, 5))
    for i, li in enumerate(distinct):
        Pi = np.eye(5)
        for j, lj in enumerate(distinct):
            if j != i:
                Pi = Pi @ (A - lj * np.eye(5)) / (li - lj)
        absA += g(li) * Pi
    return 0.5 * (euler_flux(UL, dim) + euler_flux(UR, dim)) - 0.5 * absA @ (UR - UL), lam, delta


@pytest.mark.parametrize("kind", ["sonic", "stagnation", "generic"])
def test_roe_flux_with_harten_fix(kind):
    P = SolverParams(Ma=0.3, Re=100.0).derive()
    P.riemann_solver = 0  # Roe
    rng = np.random.default_rng({"sonic": 1, "stagnation": 2, "generic": 3}[kind])
    fixed = 0
    for it in range(60):
        dim = it % 3
        rho = rng.uniform(0.5, 2.0, 2)
        p = rng.uniform(0.5, 2.0, 2)
        u = rng.uniform(-0.5, 0.5, (2, 3))
        a = np.sqrt(GAMMA * p / rho)
        if kind == "sonic":  # normal velocity close to the sound speed on both sides
            u[:, dim] = a * (1.0 + rng.uniform(-0.03, 0.03, 2))
        elif kind == "stagnation":  # normal velocity close to zero
            u[:, dim] = rng.uniform(-0.02, 0.02, 2)
        UL, UR = conserved(rho[0], u[0], p[0]), conserved(rho[1], u[1], p[1])
        want, lam, delta = roe_textbook(UL, UR, dim)
        fixed += int(np.any(np.abs(lam) < delta))
        got = riemann_batch(P, UL[None, :], UR[None, :], dim)[0]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max()), (kind, got, want)
    if kind != "generic":
        assert fixed == 60  # every case exercised the fix


# ---------------------------------------------------------------- AUSM split functions
def ausm_split(M):
    out = np.zeros(4)
    _lib().orc_ausm_split(float(M), _dp(out))
    return out


def test_ausm_split_functions():
    for M in np.concatenate([np.linspace(-3, 3, 601), [-1.0, 1.0, 0.0]]):
        mp, mm, pp, pm = ausm_split(M)
        assert mp + mm == pytest.approx(M, abs=1e-15)
        assert pp + pm == pytest.approx(1.0, abs=1e-15)
        mp2, mm2, pp2, pm2 = ausm_split(-M)
        assert mp == pytest.approx(-mm2, abs=1e-15) and pp == pytest.approx(pm2, abs=1e-15)
        if abs(M) < 1:
            assert mp == pytest.approx(0.25 * (M + 1) ** 2, abs=1e-15)
            assert pp == pytest.approx(0.25 * (M + 1) ** 2 * (2 - M), abs=1e-15)
        else:
            assert mp == pytest.approx(max(M, 0.0), abs=1e-15)
            assert pp == pytest.approx(1.0 if M > 0 else 0.0, abs=1e-15)
    for M in (1.0, -1.0):  # continuity of value and slope at |M| = 1
        for k in range(4):
            lo, hi = ausm_split(M - 1e-7)[k], ausm_split(M + 1e-7)[k]
            assert abs(hi - lo) < 1e-6, (M, k)
        for k in (0, 1):
            s_in = (ausm_split(M - np.sign(M) * 1e-6)[k] - ausm_split(M - np.sign(M) * 2e-6)[k]) / 1e-6
            s_out = (ausm_split(M + np.sign(M) * 2e-6)[k] - ausm_split(M + np.sign(M) * 1e-6)[k]) / 1e-6
            assert s_in == pytest.approx(s_out, abs=1e-5), (M, k)


# ---------------------------------------------------------------- AUSM+-up flux
def ausm_plus_up(UL, UR, dim, *, minf2=0.5, fa_on_square=True, x_velocity_diffusion=True, a_half="mean"):
    """AUSM+-up (Liou, J. Comput. Phys. 214 (2006) 137-170, eqs. (19), (28)-(30), (73), (75)) with its
    free choices as arguments: the interface sound speed (arithmetic mean or the paper's
    min(a_hat_L, a_hat_R) is out of scope: `mean` only), the cut-off M_inf^2, f_a evaluated on M_o (the
    paper) or on M_o^2 (`fa_on_square`), and the velocity in the pressure diffusion: the face-normal
    one (the paper) or the x velocity (`x_velocity_diffusion`).  Split functions with beta = alpha = 0
    (M(2), P(3)), K_p = 0.25, K_u = 0.75, sigma = 1; the mass flux upwinded at M_1/2 >= 0."""
    assert a_half == "mean"

    def prim(U):
        rho = U[0]
        vel = U[1:4] / rho
        p = max((U[4] - 0.5 * rho * vel @ vel) * (GAMMA - 1.0), 1e-12)
        return rho, vel, p, np.sqrt(GAMMA * p / rho), (U[4] + p) / rho

    rL, vL, pL, aL, HL = prim(UL)
    rR, vR, pR, aR, HR = prim(UR)
    a = 0.5 * (aL + aR)
    ML, MR = vL[dim] / a, vR[dim] / a
    Mbar2 = (vL[dim] ** 2 + vR[dim] ** 2) / (2 * a * a)
    Mo2 = min(1.0, max(Mbar2, minf2))
    Mo = Mo2 if fa_on_square else np.sqrt(Mo2)
    fa = Mo * (2.0 - Mo)

    def m2p(M):
        return 0.5 * (M + abs(M)) if abs(M) >= 1 else 0.25 * (M + 1) ** 2

    def m2m(M):
        return 0.5 * (M - abs(M)) if abs(M) >= 1 else -0.25 * (M - 1) ** 2

    def p3p(M):
        return (1.0 if M > 0 else 0.0) if abs(M) >= 1 else 0.25 * (M + 1) ** 2 * (2 - M)

    def p3m(M):
        return (0.0 if M > 0 else 1.0) if abs(M) >= 1 else 0.25 * (M - 1) ** 2 * (2 + M)

    Mp = -0.25 / fa * max(1.0 - Mbar2, 0.0) * (pR - pL) / (0.5 * (rL + rR) * a * a)
    Mhalf = m2p(ML) + m2m(MR) + Mp
    mdot = a * Mhalf * (rL if Mhalf >= 0 else rR)
    du = (vR[0] - vL[0]) if x_velocity_diffusion else (vR[dim] - vL[dim])
    pu = -0.75 * p3p(ML) * p3m(MR) * (rL + rR) * (fa * a) * du
    phalf = p3p(ML) * pL + p3m(MR) * pR + pu
    psi = np.concatenate([[1.0], vL, [HL]]) if mdot >= 0 else np.concatenate([[1.0], vR, [HR]])
    F = mdot * psi
    F[1 + dim] += phalf
    return F


def test_ausm_flux_is_ausm_plus_up_with_the_reference_choices():
    """The oracle's AUSM (InviscousFlux.cpp:719-846) on transonic and subsonic states with a pressure and a
    velocity jump (both diffusion terms active) equals AUSM+-up restated from Liou's paper with the
    reference's choices: M_inf^2 = 0.5, f_a on M_o^2, the x velocity in the pressure diffusion.  Each of
    those choices changes the flux on these states, so each is pinned: switched back to the paper's
    alternative, the flux differs."""
    P = SolverParams(Ma=0.3, Re=100.0).derive()
    P.riemann_solver = 5  # AUSM
    rng = np.random.default_rng(17)
    worst = 0.0
    differs = {"minf2": 0, "fa": 0, "xvel": 0}
    for it in range(300):
        dim = it % 3
        rho = rng.uniform(0.5, 2.0, 2)
        p = rng.uniform(0.5, 2.0, 2)
        u = rng.uniform(-0.6, 0.6, (2, 3))
        a = np.sqrt(GAMMA * p / rho)
        u[:, dim] = a * rng.uniform(-1.3, 1.3, 2)  # subsonic to supersonic normal Mach numbers
        UL, UR = conserved(rho[0], u[0], p[0]), conserved(rho[1], u[1], p[1])
        got = riemann_batch(P, UL[None, :], UR[None, :], dim)[0]
        want = ausm_plus_up(UL, UR, dim)
        scale = np.abs(want).max()
        worst = max(worst, float(np.abs(got - want).max() / scale))
        for key, alt in (("minf2", dict(minf2=0.0)), ("fa", dict(fa_on_square=False)),
                         ("xvel", dict(x_velocity_diffusion=False))):
            if np.abs(ausm_plus_up(UL, UR, dim, **alt) - got).max() > 1e-8 * scale:
                differs[key] += 1
    assert worst < 1e-12, worst  # rounding of a different operation order
    # every choice is visible on a good share of the states (the x velocity only off the x faces)
    assert differs["minf2"] > 30 and differs["fa"] > 100 and differs["xvel"] > 100, differs
