"""Shared seeded cases for the parity tests (device path vs the CPU checker)."""
from __future__ import annotations

import numpy as np

from ns3d_wcns_amd import GridDesc, Interpolation, RiemannSolver, SolverParams
from ns3d_wcns_amd.isotropic import isotropic_state

# Non-cubic interior and dx != dy != dz, so an axis mix-up cannot pass.
SMALL = (12, 10, 8)


def box(n, length=(2 * np.pi, 2 * np.pi, 2 * np.pi)) -> GridDesc:
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    return GridDesc(nx, ny, nz, Lx=length[0], Ly=length[1], Lz=length[2]).derive()


def state(n, mach_t=0.3, noise=0.0, seed=1, gamma=1.4):
    """Interior primitives: isotropic field plus optional seeded noise on
    every primitive (rho and p stay positive)."""
    (rho, u, v, w, p), ma = isotropic_state(n, gamma=gamma, mach_t=mach_t, seed=seed)
    if noise > 0:
        rng = np.random.Generator(np.random.PCG64(seed + 1000))
        rho = rho * (1.0 + noise * rng.uniform(-1, 1, rho.shape))
        u = u + noise * rng.uniform(-1, 1, u.shape)
        v = v + noise * rng.uniform(-1, 1, v.shape)
        w = w + noise * rng.uniform(-1, 1, w.shape)
        p = p * (1.0 + noise * rng.uniform(-1, 1, p.shape))
    return (rho, u, v, w, p), ma


def params(ma, interp=Interpolation.WENO5, rs=RiemannSolver.HLLC, char=False, re=100.0, cfl=0.5,
           diss=0.020) -> SolverParams:
    return SolverParams(Ma=ma, Re=re, Pr=0.72, cfl=cfl, interpolation=interp, riemann_solver=rs,
                        char_recon=char, mdcd_diss=diss, mdcd_disp=0.0463783).derive()


def rel_err(a, b) -> float:
    """Norm-wise relative difference max|a - b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    return num / den if den > 0 else num


ALL_INTERP = [Interpolation.WENO5, Interpolation.WENO5_JS, Interpolation.MDCD_LINEAR, Interpolation.MDCD_HYBRID]
ALL_RIEMANN = list(RiemannSolver)
