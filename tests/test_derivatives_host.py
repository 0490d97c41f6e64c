"""The hand-written reverse mode (adjoint) of the device numerics against
their forward-mode tangent (the same numerics on dual numbers), and the
tangent against central finite differences of the FP64 forward, on single
stencils evaluated on the host (no GPU).

- adjoint vs transposed tangent Jacobian: 1e-12 relative (two independent
  derivative implementations of one function for HLLC and Rusanov; for Roe,
  HLL, HLLC_p and AUSM the Riemann reverse is the transposed tangent itself,
  so there the check pins the interpolation reverses and the plumbing), for
  WCNS5, WENO5-JS, MDCD linear and MDCD hybrid (1e-11: the 6-point
  indicator's large coefficients amplify rounding);
- tangent vs central differences (h = 1e-6 of the input scale): 1e-6.
"""
import ctypes

import numpy as np
import pytest

from cases import params
from ns3d_wcns_amd import Interpolation, RiemannSolver
from ns3d_wcns_amd._lib import load_library

ADJ_COMBOS = [(i, r) for i in (Interpolation.WENO5, Interpolation.WENO5_JS, Interpolation.MDCD_LINEAR,
                               Interpolation.MDCD_HYBRID) for r in RiemannSolver]


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stencil(rng, mean_u, gamma=1.4):
    """x-normal 6-cell conserved stencil (var-major), smooth plus noise."""
    x = np.arange(6)
    rho = 1.0 + 0.1 * np.sin(0.7 * x + rng.uniform(0, 6)) + 0.02 * rng.uniform(-1, 1, 6)
    u = mean_u + 0.2 * np.sin(0.5 * x + rng.uniform(0, 6)) + 0.05 * rng.uniform(-1, 1, 6)
    v = 0.3 * rng.uniform(-1, 1, 6)
    w = 0.3 * rng.uniform(-1, 1, 6)
    p = 1.0 / gamma * (1.0 + 0.1 * rng.uniform(-1, 1, 6))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.concatenate([rho, rho * u, rho * v, rho * w, E])


def face_tangent(lib, cp, st, d):
    F, dF = np.zeros(5), np.zeros(5)
    assert lib.ns3d_check_face_tangent(ctypes.byref(cp), _p(st), _p(d), _p(F), _p(dF)) == 0
    return F, dF


@pytest.mark.parametrize("interp,rs", ADJ_COMBOS, ids=lambda x: x.name)
@pytest.mark.parametrize("mean_u", [-2.0, -0.4, 0.0, 0.3, 2.0])
def test_face_adjoint_is_transposed_tangent(interp, rs, mean_u):
    lib = load_library()
    cp = params(0.2, interp, rs).to_c()
    rng = np.random.Generator(np.random.PCG64(int(100 * mean_u) + 1000 + 7 * int(interp) + int(rs)))
    for _ in range(8):
        st = stencil(rng, mean_u)
        J = np.zeros((5, 30))
        for q in range(30):
            e = np.zeros(30)
            e[q] = 1.0
            J[:, q] = face_tangent(lib, cp, st, e)[1]
        Fb = rng.normal(size=5)
        stb = np.zeros(30)
        assert lib.ns3d_check_face_adjoint(ctypes.byref(cp), _p(st), _p(Fb), _p(stb)) == 0
        ref = J.T @ Fb
        # MDCD hybrid: beta3's coefficients (up to 3.6e7 / 120960) amplify
        # the rounding of the two derivative orders
        bar = 1e-11 if interp == Interpolation.MDCD_HYBRID else 1e-12
        assert np.max(np.abs(stb - ref)) <= bar * max(1.0, np.max(np.abs(ref))), (stb, ref)


@pytest.mark.parametrize("interp", [Interpolation.WENO5, Interpolation.WENO5_JS, Interpolation.MDCD_LINEAR,
                                    Interpolation.MDCD_HYBRID], ids=lambda x: x.name)
@pytest.mark.parametrize("rs", list(RiemannSolver), ids=lambda x: x.name)
@pytest.mark.parametrize("char", [False, True])
def test_face_tangent_matches_central_differences(interp, rs, char):
    lib = load_library()
    cp = params(0.2, interp, rs, char=char).to_c()
    rng = np.random.Generator(np.random.PCG64(11 + int(interp) * 10 + int(rs) + 100 * char))
    for _ in range(4):
        st = stencil(rng, 0.3)
        d = rng.normal(size=30)
        d /= np.linalg.norm(d)
        F, dF = face_tangent(lib, cp, st, d)
        h = 1e-6
        Fp, _ = face_tangent(lib, cp, st + h * d, np.zeros(30))
        Fm, _ = face_tangent(lib, cp, st - h * d, np.zeros(30))
        fd = (Fp - Fm) / (2 * h)
        assert np.max(np.abs(fd - dF)) <= 1e-6 * max(1.0, np.max(np.abs(dF))), (fd, dF)


def test_face_adjoint_rejects_unsupported_schemes():
    lib = load_library()
    st = stencil(np.random.Generator(np.random.PCG64(1)), 0.3)
    stb = np.zeros(30)
    for interp, rs, char in [(Interpolation.MDCD_HYBRID, RiemannSolver.HLLC, True),
                             (Interpolation.MDCD_LINEAR, RiemannSolver.Roe, True),
                             (Interpolation.WENO5, RiemannSolver.HLLC, True)]:
        cp = params(0.2, interp, rs, char=char).to_c()
        assert lib.ns3d_check_face_adjoint(ctypes.byref(cp), _p(st), _p(np.ones(5)), _p(stb)) == -2


def visc_case(rng):
    q = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.8, 1.3)])
    g = rng.normal(size=12)
    return q, g


def test_viscous_adjoint_is_transposed_tangent():
    lib = load_library()
    cp = params(0.3).to_c()
    rng = np.random.Generator(np.random.PCG64(5))
    for _ in range(10):
        q, g = visc_case(rng)
        J = np.zeros((9, 16))
        for c in range(16):
            dq, dg = np.zeros(4), np.zeros(12)
            (dq if c < 4 else dg)[c if c < 4 else c - 4] = 1.0
            V, dV = np.zeros(9), np.zeros(9)
            assert lib.ns3d_check_visc_tangent(ctypes.byref(cp), _p(q), _p(g), _p(dq), _p(dg), _p(V), _p(dV)) == 0
            J[:, c] = dV
        Vb = rng.normal(size=9)
        gb, qb = np.zeros(12), np.zeros(4)
        assert lib.ns3d_check_visc_adjoint(ctypes.byref(cp), _p(q), _p(g), _p(Vb), _p(gb), _p(qb)) == 0
        ref = J.T @ Vb
        got = np.concatenate([qb, gb])
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_viscous_tangent_matches_central_differences():
    lib = load_library()
    P = params(0.3)
    cp = P.to_c()
    rng = np.random.Generator(np.random.PCG64(6))

    def V_of(q, g):
        V, dV = np.zeros(9), np.zeros(9)
        lib.ns3d_check_visc_tangent(ctypes.byref(cp), _p(q), _p(g), _p(np.zeros(4)), _p(np.zeros(12)), _p(V), _p(dV))
        return V

    for _ in range(5):
        q, g = visc_case(rng)
        dq, dg = rng.normal(size=4), rng.normal(size=12)
        V, dV = np.zeros(9), np.zeros(9)
        lib.ns3d_check_visc_tangent(ctypes.byref(cp), _p(q), _p(g), _p(dq), _p(dg), _p(V), _p(dV))
        h = 1e-6
        fd = (V_of(q + h * dq, g + h * dg) - V_of(q - h * dq, g - h * dg)) / (2 * h)
        assert np.max(np.abs(fd - dV)) <= 1e-6 * max(1.0, np.max(np.abs(dV)))
        # the forward values are the reference's formulas (ViscousFlux.cpp:151-208)
        mu = P.mu * q[3] ** 1.5 * (1 + P.S_ref) / (q[3] + P.S_ref)
        G = g.reshape(4, 3)
        div = G[0, 0] + G[1, 1] + G[2, 2]
        txx = 2 * mu * G[0, 0] - 2 / 3 * mu * div
        txy = mu * (G[0, 1] + G[1, 0])
        assert abs(V[0] - txx) <= 1e-13 * max(1, abs(txx)) and abs(V[3] - txy) <= 1e-13 * max(1, abs(txy))


def test_c2p_adjoint_matches_central_differences():
    lib = load_library()
    P = params(0.3)
    cp = P.to_c()
    rng = np.random.Generator(np.random.PCG64(8))
    g, R = P.gamma, P.Rgas

    def q_of(U):
        r = U[0]
        u, v, w = U[1] / r, U[2] / r, U[3] / r
        p = (g - 1) * r * (U[4] / r - 0.5 * (u * u + v * v + w * w))
        return np.array([u, v, w, p / (r * R)])

    for _ in range(5):
        U = np.array([rng.uniform(0.8, 1.2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0])
        U[4] = 20.0 + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]
        qb = rng.normal(size=4)
        Ub = np.zeros(5)
        assert lib.ns3d_check_c2p_adjoint(ctypes.byref(cp), _p(U), _p(qb), _p(Ub)) == 0
        h = 1e-6
        ref = np.array([(q_of(U + h * e) - q_of(U - h * e)) @ qb / (2 * h) for e in np.eye(5)])
        assert np.max(np.abs(Ub - ref)) <= 1e-6 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("interp", [Interpolation.WENO5, Interpolation.WENO5_JS], ids=lambda x: x.name)
def test_paired_reconstruction_adjoint(interp):
    """The cell-wise reverse of both reconstructions of a stencil (shared
    smoothness indicators, used by the adjoint kernels) equals the sum of the
    two separate reverses, on smooth and on discontinuous stencils."""
    lib = load_library()
    cp = params(0.2, interp).to_c()
    rng = np.random.Generator(np.random.PCG64(31))
    for trial in range(20):
        f = np.cumsum(rng.normal(size=5)) if trial % 2 else 1.0 + 0.01 * rng.normal(size=5)
        if trial % 5 == 0:
            f[3:] += 5.0  # a jump
        a, b = np.zeros(5), np.zeros(5)
        ybL, ybR = rng.normal(size=2)
        assert lib.ns3d_check_weno_pair_adjoint(ctypes.byref(cp), _p(f), ybL, ybR, _p(a), _p(b)) == 0
        assert np.max(np.abs(a - b)) <= 1e-13 * max(1.0, np.max(np.abs(b))), (a, b)
