"""Tangent-linear and discrete adjoint of the RHS and the RK3 step, on the
device through the C-ABI, and the checkpointed autograd rollout.

- RHS and step dot tests <(dR/dU) dU, lam> = <dU, (dR/dU)^T lam>: the two
  agree to 1e-13 of |(dR/dU) dU| |lam| (the rounding scale of the inner
  product; the tangent runs the forward numerics on dual numbers, the
  adjoint is hand-written reverse mode: independent implementations), on
  boxes of one and of several marching segments along every axis, and x
  lines held whole or in segments (512 cells: two, 517 cells: three, the
  last shorter); the same for Roe, HLL, HLLC_p and AUSM, whose reverse
  transposes their tangent solver by solver (the interpolation reverses,
  the transposed differences and the viscous reverse stay independent of
  the tangent), for MDCD linear and hybrid interpolation with four solvers,
  and the step dot test with Roe, AUSM and MDCD hybrid;
- tangent RHS vs central differences of compute_rhs: 1e-6 (step 1e-6 of
  each field's scale, 1e-7 for MDCD hybrid);
- gradient of J = mean 0.5 rho |u|^2 at the final time with respect to the
  initial conserved fields vs central differences (eps = 1e-7) on 8 seeded
  directions with unit-variance entries: |fd - ad| <= 1e-6 |grad J|_2 (the
  RMS of such directional derivatives) on every direction;
- checkpoint policies give the same gradient to 1e-12, and so do the
  adjoint steps with stored stage states (step_stages / adjoint_step_stages)
  and with recomputed ones;
- time-integrated viscous dissipation J = sum_n dt_n eps(U_{n+1}): the
  checkpointed forward value equals a plain forward's, the gradient matches
  central differences (eps = 1e-7) on 4 seeded directions with
  unit-variance entries to 1e-6 |grad J|_2 on every direction, and a mixed-precision
  forward with the FP64 adjoint gives the FP64 gradient to 1e-4.
"""
import numpy as np
import pytest

from cases import SMALL, box, params, state
from ns3d_wcns_amd import DeviceField, Interpolation, NS3DError, RiemannSolver

pytestmark = pytest.mark.gpu

ADJ = [(Interpolation.WENO5, RiemannSolver.HLLC), (Interpolation.WENO5, RiemannSolver.Rusanov),
       (Interpolation.WENO5_JS, RiemannSolver.HLLC), (Interpolation.WENO5_JS, RiemannSolver.Rusanov)]
# the solvers whose reverse transposes their tangent (Roe, HLL, HLLC_p, AUSM)
ADJ_DUAL = [(i, r) for i in (Interpolation.WENO5, Interpolation.WENO5_JS)
            for r in (RiemannSolver.Roe, RiemannSolver.HLL, RiemannSolver.HLLC_p, RiemannSolver.AUSM)]
# MDCD interpolation (six-cell face stencils; fused march and segmented x kernels)
ADJ_MDCD = [(i, r) for i in (Interpolation.MDCD_LINEAR, Interpolation.MDCD_HYBRID)
            for r in (RiemannSolver.HLLC, RiemannSolver.Rusanov, RiemannSolver.Roe, RiemannSolver.AUSM)]
NG = 3


def field_and_state(n, P, mach_t=0.3, noise=0.02, seed=3):
    import torch
    from ns3d_wcns_amd.adjoint import state_tensor

    prim, _ = state(n, mach_t, noise=noise, seed=seed)
    f = DeviceField(box(n), P)
    f.set_stream(torch.cuda.current_stream().cuda_stream)
    f.initialize_primitive(*prim)
    return f, state_tensor(f)


def rand_interior(shape, seed):
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    X = torch.zeros(shape, dtype=torch.float64)
    X[:, NG:-NG, NG:-NG, NG:-NG] = torch.randn((shape[0],) + tuple(s - 2 * NG for s in shape[1:]), generator=g,
                                               dtype=torch.float64)
    return X.cuda()


def ptrs(t):
    return [t[n].data_ptr() for n in range(5)]


def inner(a, b):
    return float((a[:, NG:-NG, NG:-NG, NG:-NG] * b[:, NG:-NG, NG:-NG, NG:-NG]).sum())


def norm(a):
    return float(a[:, NG:-NG, NG:-NG, NG:-NG].norm())


def dot_agree(tl, lam, dU, ad, what):
    a, b = inner(tl, lam), inner(dU, ad)
    scale = norm(tl) * norm(lam)
    print(f"{what}: <J dU, lam> = {a!r}, <dU, J^T lam> = {b!r}, |diff| / (|J dU| |lam|) = {abs(a - b) / scale:.2e}")
    assert abs(a - b) <= 1e-13 * scale, (a, b, scale)


@pytest.mark.parametrize("n", [SMALL, (40, 24, 20), (512, 6, 8), (517, 6, 8)],
                         ids=["one-segment", "several-segments", "512-cell-x-line", "ragged-x-segments"])
@pytest.mark.parametrize("interp,rs", ADJ, ids=lambda x: x.name)
def test_rhs_dot_product(gpu, interp, rs, n):
    import torch

    P = params(0.17, interp, rs)
    f, U = field_and_state(n, P)
    dU = rand_interior(U.shape, 1) * 1e-3
    lam = rand_interior(U.shape, 2)
    tl, ad = torch.zeros_like(U), torch.zeros_like(U)
    f.tangent_rhs(ptrs(dU), ptrs(tl))
    f.adjoint_rhs(ptrs(lam), ptrs(ad))
    torch.cuda.synchronize()
    dot_agree(tl, lam, dU, ad, f"{interp.name}+{rs.name} {n}")
    f.close()


@pytest.mark.parametrize("n", [SMALL, (40, 24, 20), (517, 6, 8)],
                         ids=["one-segment", "several-segments", "ragged-x-segments"])
@pytest.mark.parametrize("interp,rs", ADJ_DUAL + ADJ_MDCD, ids=lambda x: x.name)
def test_rhs_dot_product_other_schemes(gpu, interp, rs, n):
    import torch

    P = params(0.17, interp, rs)
    f, U = field_and_state(n, P)
    dU = rand_interior(U.shape, 3) * 1e-3
    lam = rand_interior(U.shape, 4)
    tl, ad = torch.zeros_like(U), torch.zeros_like(U)
    f.tangent_rhs(ptrs(dU), ptrs(tl))
    f.adjoint_rhs(ptrs(lam), ptrs(ad))
    torch.cuda.synchronize()
    dot_agree(tl, lam, dU, ad, f"{interp.name}+{rs.name} {n}")
    f.close()


@pytest.mark.parametrize("interp,rs,char", [(Interpolation.WENO5, RiemannSolver.HLLC, False),
                                            (Interpolation.MDCD_LINEAR, RiemannSolver.Roe, False),
                                            (Interpolation.MDCD_HYBRID, RiemannSolver.Rusanov, False),
                                            (Interpolation.WENO5_JS, RiemannSolver.AUSM, False),
                                            (Interpolation.WENO5, RiemannSolver.HLLC_p, True)],
                         ids=lambda x: getattr(x, "name", str(x)))
def test_tangent_rhs_matches_central_differences(gpu, interp, rs, char):
    import torch

    P = params(0.17, interp, rs, char=char)
    f, U = field_and_state(SMALL, P)
    d = rand_interior(U.shape, 4)
    d *= U.abs().amax(dim=(1, 2, 3), keepdim=True)  # per-field scale
    tl = torch.zeros_like(U)
    f.tangent_rhs(ptrs(d), ptrs(tl))
    # MDCD hybrid: the weights 20 + tau6 / beta_k curve strongly where a
    # stencil is smooth, so the difference step is 10 times shorter
    h = 1e-7 if interp == Interpolation.MDCD_HYBRID else 1e-6
    rhs = []
    for sgn in (1.0, -1.0):
        Us = (U + sgn * h * d).contiguous()
        f.set_state(ptrs(Us))
        f.compute_rhs()
        torch.cuda.synchronize()
        rhs.append(np.stack([f.interior(f.download(n)) for n in
                             ("rhs_rho", "rhs_rhou", "rhs_rhov", "rhs_rhow", "rhs_E")]))
    fd = (rhs[0] - rhs[1]) / (2 * h)
    t = tl[:, NG:-NG, NG:-NG, NG:-NG].cpu().numpy()
    for n in range(5):
        err = np.max(np.abs(fd[n] - t[n])) / np.max(np.abs(t[n]))
        assert err < 1e-6, (n, err)
    f.close()


@pytest.mark.parametrize("n,interp,rs", [(16, Interpolation.WENO5, RiemannSolver.HLLC),
                                         (36, Interpolation.WENO5, RiemannSolver.HLLC),
                                         (16, Interpolation.WENO5, RiemannSolver.Roe),
                                         (16, Interpolation.WENO5, RiemannSolver.AUSM),
                                         (16, Interpolation.MDCD_HYBRID, RiemannSolver.HLLC),
                                         (36, Interpolation.MDCD_HYBRID, RiemannSolver.Rusanov)],
                         ids=lambda x: getattr(x, "name", str(x)))
def test_step_dot_product(gpu, n, interp, rs):
    import torch

    P = params(0.17, interp, rs)
    f, U = field_and_state(n, P)
    dt = 0.5 * f.compute_timestep()
    dU = rand_interior(U.shape, 5) * 1e-3
    lam = rand_interior(U.shape, 6)
    tl = dU.clone()
    f.set_state(ptrs(U))
    f.tangent_step(dt, ptrs(tl))
    ad = lam.clone()
    f.set_state(ptrs(U))
    f.adjoint_step(dt, ptrs(ad))
    torch.cuda.synchronize()
    dot_agree(tl, lam, dU, ad, "step")
    f.close()


def rollout_grad(f, U, dts, policy, stats=None):
    from ns3d_wcns_amd.adjoint import mean_kinetic_energy, rollout

    U0 = U.clone().requires_grad_(True)
    J = mean_kinetic_energy(rollout(U0, f, dts, policy, stats))
    J.backward()
    return float(J.detach()), U0.grad


def test_gradient_matches_central_differences(gpu):
    """J = mean 0.5 rho |u|^2 after 10 steps; gradient with respect to the
    initial rho u, rho v, rho w and E, checkpoint every 5 steps, against
    central differences (eps = 1e-7) along 8 seeded random directions d
    with independent unit-variance entries.  For such directions the
    derivative <grad J, d> has RMS |grad J|_2, so every direction is judged
    against that scale: |fd - ad| <= 1e-6 |grad J|_2 (a single direction's
    own derivative can be arbitrarily small, when d happens to be nearly
    orthogonal to the gradient; the relative error is printed).  eps from a
    sweep of 1e-4 .. 1e-8 on this case: above 1e-7 the differences carry
    O(eps^2) truncation and cross branch switches of the flux (min / max of
    wave speeds), up to 1.3e-6 |grad J| at eps = 1e-6; below, the rounding
    floor of the perturbed trajectories (~1e-16 J / eps) takes over; at 1e-7
    every direction agrees to 5e-9 |grad J|."""
    import torch
    from ns3d_wcns_amd.adjoint import central_difference

    P = params(0.3 / np.sqrt(3.0))
    f, U = field_and_state(16, P, noise=0.0, seed=9)
    dts = []
    f.set_state(ptrs(U))
    for _ in range(10):
        dts.append(f.compute_timestep())
        f.step(dts[-1])
    J0, grad = rollout_grad(f, U, dts, "uniform:5")
    eps = 1e-7
    gnorm = float(torch.linalg.vector_norm(grad[1:]))
    for k in range(8):
        d = rand_interior(U.shape, 100 + k)
        d[0] = 0.0  # directions in (rho u, rho v, rho w, E)
        fd = float(central_difference(f, U, d, dts, eps))
        ad = float((grad * d).sum())
        err = abs(fd - ad)
        print(f"direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {err / abs(ad):.2e} over |grad J| {err / gnorm:.2e}")
        assert err <= 1e-6 * gnorm, (k, err / gnorm)
    f.close()


def test_checkpoint_policies_agree(gpu):
    P = params(0.17)
    f, U = field_and_state(12, P, seed=4)
    dts = [0.4 * f.compute_timestep()] * 9
    ref = None
    for policy in ("all", "uniform:4", "uniform:1", "binomial:2", "binomial:0"):
        stats = {}
        _, g = rollout_grad(f, U, dts, policy, stats)
        print(policy, {k: v for k, v in stats.items() if k != "live"})
        assert stats["adjoint_steps"] == len(dts)
        if ref is None:
            ref = g
        else:
            err = float((g - ref).abs().max() / ref.abs().max())
            assert err < 1e-12, (policy, err)
    f.close()


def test_adjoint_with_stored_stages(gpu):
    """step_stages + adjoint_step_stages (stage states kept by the caller)
    against adjoint_step (stages recomputed inside): the stage states are
    the same computation, so the adjoints agree to rounding (1e-13), and
    the state after step_stages is the state after step."""
    import torch

    P = params(0.17)
    f, U = field_and_state(SMALL, P, seed=6)
    dt = 0.5 * f.compute_timestep()
    lam = rand_interior(U.shape, 9)
    U1, U2 = torch.empty_like(U), torch.empty_like(U)
    f.set_state(ptrs(U))
    f.step_stages(dt, ptrs(U1), ptrs(U2))
    after = torch.empty_like(U)
    f.get_state(ptrs(after))
    f.set_state(ptrs(U))
    f.step(dt)
    ref = torch.empty_like(U)
    f.get_state(ptrs(ref))
    assert torch.equal(after, ref)
    a1, a2 = lam.clone(), lam.clone()
    f.set_state(ptrs(U))
    f.adjoint_step(dt, ptrs(a1))
    f.set_state(ptrs(U))
    f.adjoint_step_stages(dt, ptrs(a2), ptrs(U1), ptrs(U2))
    torch.cuda.synchronize()
    err = float((a2 - a1).abs().max() / a1.abs().max())
    print(f"stored-stage adjoint vs recomputed: {err:.2e}")
    assert err < 1e-13, err
    for policy, keep in (("uniform:4", False), ("uniform:4", True)):
        stats = {}
        _, g = rollout_grad_keep(f, U, [dt] * 6, policy, keep, stats)
        if keep:
            assert stats["recomputed_steps"] == 6
            err = float((g - g0).abs().max() / g0.abs().max())
            assert err < 1e-12, err
        else:
            g0 = g
    f.close()


def rollout_grad_keep(f, U, dts, policy, keep, stats):
    from ns3d_wcns_amd.adjoint import mean_kinetic_energy, rollout

    U0 = U.clone().requires_grad_(True)
    mean_kinetic_energy(rollout(U0, f, dts, policy, stats, keep_stages=keep)).backward()
    return None, U0.grad


def test_adjoint_rejects_unsupported_scheme(gpu):
    import torch

    for interp, char in ((Interpolation.MDCD_HYBRID, True), (Interpolation.MDCD_LINEAR, True),
                         (Interpolation.WENO5, True)):
        P = params(0.17, interp, RiemannSolver.HLLC, char=char)
        f, U = field_and_state(SMALL, P)
        lam = torch.zeros_like(U)
        with pytest.raises(NS3DError):
            f.adjoint_rhs(ptrs(lam), ptrs(torch.zeros_like(U)))
        with pytest.raises(NS3DError):
            f.adjoint_step(1e-3, ptrs(lam))
        f.close()


def integral_setup(precision=None, steps=8):
    from ns3d_wcns_amd import PRECISION_FP64
    from ns3d_wcns_amd.adjoint import dissipation_rate, state_tensor

    import torch

    P = params(0.3 / np.sqrt(3.0))
    prim, _ = state(16, 0.3, noise=0.0, seed=11)
    f = DeviceField(box(16), P, precision=PRECISION_FP64 if precision is None else precision)
    f.set_stream(torch.cuda.current_stream().cuda_stream)
    f.initialize_primitive(*prim)
    U = state_tensor(f)
    dts = []
    for _ in range(steps):
        dts.append(f.compute_timestep())
        f.step(dts[-1])
    return f, U, dts, (lambda X: dissipation_rate(X, f.params, f.grid))


def integral_grad(f, U, dts, density, policy):
    from ns3d_wcns_amd.adjoint import time_integral

    U0 = U.clone().requires_grad_(True)
    J = time_integral(U0, f, dts, density, policy)
    J.backward()
    return float(J.detach()), U0.grad


def test_time_integrated_dissipation_gradient(gpu):
    """J = sum_n dt_n eps(U_{n+1}), eps the volume-mean viscous dissipation,
    over 8 CFL steps at 16^3; gradient with respect to the initial rho u,
    rho v, rho w and E."""
    import torch
    from ns3d_wcns_amd.adjoint import time_integral_value

    f, U, dts, density = integral_setup()
    J, grad = integral_grad(f, U, dts, density, "uniform:3")
    J_plain = time_integral_value(f, U, dts, density)
    assert abs(J - J_plain) <= 1e-14 * abs(J_plain), (J, J_plain)
    for policy in ("all", "binomial:2"):
        _, g = integral_grad(f, U, dts, density, policy)
        err = float((g - grad).abs().max() / grad.abs().max())
        assert err < 1e-12, (policy, err)
    # Directions with independent unit-variance entries: their derivative
    # has RMS |grad J|_2, the scale every direction is judged against
    # (|fd - ad| <= 1e-6 |grad J|_2; the relative error is printed).  eps =
    # 1e-7 from a sweep of 1e-4 .. 1e-8: larger steps carry truncation and
    # branch crossings (4.9e-7 |grad J| at 1e-6), smaller ones the rounding of
    # J, a sum of volume means (~1e-16 J / eps); at 1e-7 the worst direction
    # agrees to 5.8e-8 |grad J|.
    eps = 1e-7
    gnorm = float(torch.linalg.vector_norm(grad[1:]))
    for k in range(4):
        d = rand_interior(U.shape, 200 + k)
        d[0] = 0.0
        fd = (time_integral_value(f, U + eps * d, dts, density) - time_integral_value(f, U - eps * d, dts, density)) / (
            2 * eps)
        ad = float((grad * d).sum())
        err = abs(fd - ad)
        print(f"direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {err / abs(ad):.2e} over |grad J| {err / gnorm:.2e}")
        assert err <= 1e-6 * gnorm, (k, err / gnorm)
    f.close()


def test_dissipation_gradient_by_slabs(gpu):
    """The slab-wise gradient of the dissipation rate (z-slabs with periodic
    halo planes, slabs of 5 and 32 on a 16-plane box, the latter wrapping
    one slab around the box) equals torch autograd of the whole-box rate."""
    import torch
    from ns3d_wcns_amd.adjoint import dissipation_rate, dissipation_rate_grad

    f, U, _, _ = integral_setup(steps=2)
    X = U.clone().requires_grad_(True)
    (g_ref,) = torch.autograd.grad(dissipation_rate(X, f.params, f.grid), X)
    for slab in (5, 32):
        g = dissipation_rate_grad(U, f.params, f.grid, slab=slab)
        err = float((g - g_ref).abs().max() / g_ref.abs().max())
        print(f"slab {slab}: {err:.2e}")
        assert err < 1e-13, (slab, err)
    f.close()


def test_time_integral_mixed_forward(gpu):
    """Mixed FP32/FP64 forward, FP64 adjoint about its trajectory: the
    gradient is the FP64 one up to the forward's rounding."""
    from ns3d_wcns_amd import PRECISION_MIXED

    f, U, dts, density = integral_setup()
    _, g64 = integral_grad(f, U, dts, density, "uniform:4")
    f.close()
    fm, Um, dtm, dm = integral_setup(PRECISION_MIXED)
    _, gm = integral_grad(fm, U, dts, dm, "uniform:4")
    err = float((gm - g64).abs().max() / g64.abs().max())
    print(f"mixed-forward gradient vs FP64: {err:.2e}")
    assert err < 1e-4, err
    fm.close()
