"""Device path vs the CPU checker, through the C-ABI, on seeded states.

Tolerances: index maps, conversions and the CFL time step are bit-exact;
per-field RHS, face fluxes and one RK3 step within 1e-12 norm-wise relative
(max|gpu - ref| / max|ref|), the FP64 parity bar for the plain method (a
per-cell relative bound has no meaning where a field's RHS crosses zero).

On a periodic box the half-node and viscous-flux exchanges move no data:
the consuming kernels read periodic images by wrapped index, and
download_face_flux fills the reference layout's ghost faces from them on
the host.  test_face_flux therefore checks the ghost faces of the download;
the device's wrapped reads are checked by every RHS, step and trajectory
test below, whose results depend on them.
"""
import itertools

import numpy as np
import pytest

from cases import ALL_INTERP, ALL_RIEMANN, SMALL, box, params, rel_err, state
from ns3d_wcns_amd import DeviceField, Interpolation, RiemannSolver
from oracle.ns3d_oracle import OracleField

pytestmark = pytest.mark.gpu

CONS = ("rho", "rhou", "rhov", "rhow", "E")
RHS = ("rhs_rho", "rhs_rhou", "rhs_rhov", "rhs_rhow", "rhs_E")
TOL = 1e-12


def faces_periodic(X, g=3) -> bool:
    """The face ghost layers of a ghosted (sz, sy, sx) array hold the periodic
    images of the interior (edges and corners are not part of the map)."""
    I = X[g:-g, g:-g, g:-g]
    return (np.array_equal(X[g:-g, g:-g, :g], I[:, :, -g:]) and np.array_equal(X[g:-g, g:-g, -g:], I[:, :, :g])
            and np.array_equal(X[g:-g, :g, g:-g], I[:, -g:, :]) and np.array_equal(X[g:-g, -g:, g:-g], I[:, :g, :])
            and np.array_equal(X[:g, g:-g, g:-g], I[-g:, :, :]) and np.array_equal(X[-g:, g:-g, g:-g], I[:g, :, :]))


def pair(n, P, prim):
    G = box(n)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    return dev, ref


def test_initial_state_bit_exact(gpu):
    prim, ma = state(SMALL, 0.3, noise=0.05, seed=3)
    dev, ref = pair(SMALL, params(ma), prim)
    for name in CONS + ("u", "v", "w", "p", "T"):
        a, b = dev.download(name), ref.view(name)
        # interior and face ghosts (edges/corners are never written by either)
        assert np.array_equal(dev.interior(a), ref.interior(b)), name
    dev.close()


@pytest.mark.parametrize("seed,n", [(1, SMALL), (2, SMALL), (3, SMALL), (4, (64, 72, 66)), (5, (600, 6, 8)),
                                    (6, (520, 6, 6))],
                         ids=["small-1", "small-2", "small-3", "64x72x66", "x-line-600", "x-line-520"])
def test_timestep_bit_exact(gpu, seed, n):
    """dt of the device's row-wise maximum equals the checker's: rows shorter
    than a block, several blocks of rows, rows longer than a block (a thread
    takes two or three cells of its row)."""
    prim, ma = state(n, 0.6, noise=0.02, seed=seed)
    dev, ref = pair(n, params(ma), prim)
    assert dev.compute_timestep() == ref.compute_timestep()
    dev.close()


def test_conversions_bit_exact(gpu):
    prim, ma = state(SMALL, 0.3, noise=0.05, seed=4)
    dev, ref = pair(SMALL, params(ma), prim)
    # perturb the conserved state so that C2P does real work, identically on both
    rng = np.random.Generator(np.random.PCG64(9))
    for name in CONS:
        g = ref.view(name)
        g *= 1.0 + 1e-3 * rng.uniform(-1, 1, g.shape)
        dev.upload(name, g)
    dev.conservedToPrimitive()
    ref.conserved_to_primitive()
    for name in ("u", "v", "w", "p", "T"):
        assert np.array_equal(dev.download(name), ref.view(name)), name
    dev.primitiveToConserved()
    ref.primitive_to_conserved()
    for name in CONS + ("T",):
        assert np.array_equal(dev.download(name), ref.view(name)), name
    dev.close()


def test_halo_map_bit_exact(gpu):
    """apply_boundary on encoded cell ids: the periodic face-ghost map."""
    prim, ma = state(SMALL, 0.3, seed=5)
    dev, ref = pair(SMALL, params(ma), prim)
    ids = np.arange(ref.sx * ref.sy * ref.sz, dtype=np.float64).reshape(ref.shape)
    for q, name in enumerate(("rho", "u", "v", "w", "p")):
        ref.view(name)[...] = ids + 0.25 * q
        dev.upload(name, ids + 0.25 * q)
    dev.apply_boundary()
    ref.apply_boundary()
    for name in ("rho", "u", "v", "w", "p"):
        assert np.array_equal(dev.download(name), ref.view(name)), name
    dev.close()


@pytest.mark.parametrize("char", [False, True])
def test_face_flux(gpu, char):
    prim, ma = state(SMALL, 0.3, noise=0.03, seed=6)
    dev, ref = pair(SMALL, params(ma, char=char), prim)
    dev.compute_invis_flux()
    dev.apply_boundary_halfnode_flux()
    ref.compute_invis_flux()
    ref.apply_boundary_halfnode_flux()
    for d, axis in enumerate("xyz"):
        for c, comp in enumerate(("mass", "momx", "momy", "momz", "E")):
            a = dev.download_face_flux(d, c)
            b = ref.view(f"flux_f{axis}_{comp}")
            assert rel_err(a, b) < TOL, (axis, comp, rel_err(a, b))
    dev.close()


def test_gradients_and_viscous_flux(gpu):
    prim, ma = state(SMALL, 0.3, noise=0.03, seed=7)
    dev, ref = pair(SMALL, params(ma), prim)
    dev.compute_gradients()
    ref.compute_gradients()
    names = ["du_dx", "du_dy", "du_dz", "dv_dx", "dv_dy", "dv_dz", "dw_dx", "dw_dy", "dw_dz",
             "dT_dx", "dT_dy", "dT_dz"]
    for q, name in enumerate(names):
        assert rel_err(dev.download_gradient(q), ref.view(name)) < TOL, name
    dev.compute_viscous_flux()
    dev.exchange_halos_viscous_flux()
    ref.compute_viscous_flux()
    ref.exchange_halos_viscous_flux()
    for d, pre in enumerate(("Fv", "Hv", "Gv")):
        for c, comp in enumerate(("mass", "momx", "momy", "momz", "E")):
            a = dev.download_viscous_flux(d, c)
            b = ref.view(f"{pre}flux_{comp}")
            a, b = dev.interior(a), ref.interior(b)  # reference leaves the mass component 0
            assert rel_err(a, b) < TOL, (pre, comp, rel_err(a, b))
    dev.close()


@pytest.mark.parametrize("interp,rs,char", list(itertools.product(ALL_INTERP, ALL_RIEMANN, [False, True])))
def test_compute_rhs(gpu, interp, rs, char):
    prim, ma = state(SMALL, 0.35, noise=0.03, seed=11)
    dev, ref = pair(SMALL, params(ma, interp, rs, char), prim)
    dev.compute_rhs()
    ref.compute_rhs()
    for name in RHS:
        a, b = dev.download(name), ref.view(name)
        assert rel_err(a, b) < TOL, (name, rel_err(a, b))
    dev.close()


def test_stepwise_rhs_equals_fused(gpu):
    """The reference's compute_rhs sequence called entry point by entry point."""
    prim, ma = state(SMALL, 0.3, noise=0.03, seed=12)
    dev, ref = pair(SMALL, params(ma), prim)
    for name in RHS:
        dev.upload(name, np.zeros(dev.shape))
    dev.compute_invis_flux()
    dev.apply_boundary_halfnode_flux()
    dev.compute_invis_dflux()
    dev.compute_gradients()
    dev.compute_viscous_flux()
    dev.exchange_halos_viscous_flux()
    dev.compute_vis_flux()
    ref.compute_rhs()
    for name in RHS:
        assert rel_err(dev.download(name), ref.view(name)) < TOL, name
    dev.close()


@pytest.mark.parametrize("rs", [RiemannSolver.HLLC, RiemannSolver.Rusanov])
def test_rk3_step(gpu, rs):
    prim, ma = state(SMALL, 0.3, noise=0.02, seed=13)
    dev, ref = pair(SMALL, params(ma, rs=rs), prim)
    dt = ref.compute_timestep()
    ref.record_conserved_to0()
    ref.runge_kutta_3(dt)
    dev.recordConservedTo0()
    dev.runge_kutta_3(dt)
    for name in CONS:
        a, b = dev.interior(dev.download(name)), ref.interior(ref.view(name))
        b0 = ref.interior(ref.view(name + "0"))
        # the change over the step carries the RHS; compare against its size
        assert np.max(np.abs(a - b)) / np.max(np.abs(b - b0)) < 1e-11, name
        assert rel_err(a, b) < TOL, name
    dev.close()


# x lines of 64 cells (one face per thread, four lines per block) and y / z
# columns long enough for two marching segments each
MULTI = (64, 72, 66)


@pytest.mark.parametrize("interp,char", [(Interpolation.WENO5, False), (Interpolation.WENO5_JS, False),
                                         (Interpolation.MDCD_HYBRID, False), (Interpolation.WENO5, True)],
                         ids=["WENO5", "WENO5_JS", "MDCD_HYBRID", "WENO5-char"])
def test_rk3_step_several_segments(gpu, interp, char):
    prim, ma = state(MULTI, 0.3, noise=0.02, seed=21)
    dev, ref = pair(MULTI, params(ma, interp=interp, char=char), prim)
    dt = ref.compute_timestep()
    ref.record_conserved_to0()
    ref.runge_kutta_3(dt)
    dev.recordConservedTo0()
    dev.runge_kutta_3(dt)
    for name in CONS:
        a, b = dev.interior(dev.download(name)), ref.interior(ref.view(name))
        b0 = ref.interior(ref.view(name + "0"))
        assert np.max(np.abs(a - b)) / np.max(np.abs(b - b0)) < 1e-11, name
        assert rel_err(a, b) < TOL, name
    dev.close()


@pytest.mark.parametrize("n", [(512, 8, 6), (600, 6, 8)], ids=["x-line-512", "x-line-600"])
def test_rk3_step_long_x_lines(gpu, n):
    """x lines of 512 cells (one block of 512 threads per line, paired
    reconstruction) and of 600 (the general x pencil, 256 threads looping over
    the line)."""
    prim, ma = state(n, 0.3, noise=0.02, seed=24)
    dev, ref = pair(n, params(ma), prim)
    dt = ref.compute_timestep()
    ref.record_conserved_to0()
    ref.runge_kutta_3(dt)
    dev.recordConservedTo0()
    dev.runge_kutta_3(dt)
    for name in CONS:
        a, b = dev.interior(dev.download(name)), ref.interior(ref.view(name))
        b0 = ref.interior(ref.view(name + "0"))
        assert np.max(np.abs(a - b)) / np.max(np.abs(b - b0)) < 1e-11, name
        assert rel_err(a, b) < TOL, name
    dev.close()


def test_time_advance_trajectory(gpu):
    """10 steps of time_advance: dt sequence and kinetic energy."""
    prim, ma = state(16, 0.3, seed=21)
    dev, ref = pair(16, params(ma), prim)
    dts_ref = ref.time_advance(10)
    dts_dev = dev.time_advance(10)
    assert len(dts_dev) == len(dts_ref) == 10
    assert rel_err(dts_dev, dts_ref) < 1e-12
    assert abs(dev.total_kinetic_energy() - ref.total_kinetic()) / ref.total_kinetic() < 1e-12
    for name in CONS:
        assert rel_err(dev.interior(dev.download(name)), ref.interior(ref.view(name))) < 1e-11, name
    dev.close()


def test_time_advance_clipping(gpu):
    """dt clipped to the output interval and the total time (maincycle.cpp:33-45)."""
    prim, ma = state(SMALL, 0.3, seed=22)
    dev, ref = pair(SMALL, params(ma), prim)
    dts_ref = ref.time_advance(50, -1.0, 0.05, 0.02)
    dts_dev = dev.time_advance(50, -1.0, 0.05, 0.02)
    assert len(dts_dev) == len(dts_ref)
    assert rel_err(dts_dev, dts_ref) < 1e-12
    assert abs(np.sum(dts_dev) - 0.05) < 1e-12
    dev.close()


def test_advance_steps_matches_time_advance(gpu):
    """The device-only loop (dt kept on the device; every step after the
    first takes it from the CFL maxima that the previous step's stage-3 x
    pencil folded in) takes the same steps as the host-scheduled
    time_advance, whose steps run the separate CFL pass."""
    import torch

    prim, ma = state(16, 0.3, seed=23)
    a, _ = pair(16, params(ma), prim)
    b, _ = pair(16, params(ma), prim)
    dts = a.time_advance(5)
    d_hist = torch.zeros(5, dtype=torch.float64, device="cuda")
    b.advance_steps(5, d_hist.data_ptr())
    b.synchronize()
    assert np.array_equal(d_hist.cpu().numpy(), dts)
    for name in CONS + ("u", "p", "T"):
        assert np.array_equal(a.download(name), b.download(name)), name
        # the loops run their steps without ghost refills and refill once
        # before they return
        assert faces_periodic(a.download(name)) and faces_periodic(b.download(name)), name
    # the device loop records U0 for the interior cells only (ghost cells of
    # U0 are never read)
    for name in ("rho0", "rhou0", "E0"):
        assert np.array_equal(a.interior(a.download(name)), b.interior(b.download(name))), name
    a.close()
    b.close()


@pytest.mark.parametrize("precision,char", [(0, False), (1, False), (0, True)], ids=["fp64", "mixed", "fp64-char"])
def test_folded_cfl_dt_bit_exact(gpu, precision, char):
    """dt from the CFL maxima folded into the stage-3 x pencil equals, bit for
    bit, the separate CFL pass over the same states: x lines of 64 cells
    (the paired x pencil, or the general one characteristic-wise), several
    lines per block, in both precisions."""
    import torch

    n = (64, 40, 36)
    prim, ma = state(n, 0.35, noise=0.02, seed=31)
    P = params(ma, char=char)
    a = DeviceField(box(n), P, precision=precision)
    b = DeviceField(box(n), P, precision=precision)
    a.initialize_primitive(*prim)
    b.initialize_primitive(*prim)
    dts = a.time_advance(6)
    d_hist = torch.zeros(6, dtype=torch.float64, device="cuda")
    b.advance_steps(6, d_hist.data_ptr())
    b.synchronize()
    assert np.array_equal(d_hist.cpu().numpy(), dts)
    for name in CONS:
        assert np.array_equal(a.interior(a.download(name)), b.interior(b.download(name))), name
    a.close()
    b.close()
