"""Seeded synthetic decaying isotropic turbulence for a periodic [0, 2pi]^3 box.

Velocity: solenoidal random field with the model spectrum
E(k) = 16 sqrt(2/pi) u0^2 k^4 / k0^5 exp(-2 k^2 / k0^2), uniform random phases
(complex Gaussian modes), projected onto the plane normal to k, amplitude per
mode sqrt(2 E(|k|) / (4 pi |k|^2)), real part of the inverse FFT, then scaled so
that the mean over the three components of <u_i^2> is u0^2.

Thermodynamics: rho = 1, T = 1, p = 1 / (gamma Ma^2) with the reference's
nondimensionalisation (Rgas = 1/(gamma Ma^2), initialization.cpp:196), where
Ma = Ma_t / sqrt(3) so that sqrt(<|u|^2>) / c = Ma_t for u0 = 1.

The same arrays feed the device path and the CPU checker.
"""
from __future__ import annotations

import math

import numpy as np


def model_spectrum(k, k0=4.0, u0=1.0):
    return 16.0 * math.sqrt(2.0 / math.pi) * u0 ** 2 * k ** 4 / k0 ** 5 * np.exp(-2.0 * k ** 2 / k0 ** 2)


def isotropic_velocity(n, k0: float = 4.0, u0: float = 1.0, seed: int = 12345):
    """(u, v, w) arrays of shape (nz, ny, nx), x fastest, for integer
    wavenumbers of a 2pi-periodic box."""
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    rng = np.random.Generator(np.random.PCG64(seed))
    kz = np.fft.fftfreq(nz, 1.0 / nz)[:, None, None]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None]
    kx = np.fft.fftfreq(nx, 1.0 / nx)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    kmag = np.sqrt(k2)
    shape = (nz, ny, nx)
    comps = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    with np.errstate(divide="ignore", invalid="ignore"):
        kdot = (kx * comps[0] + ky * comps[1] + kz * comps[2]) / np.where(k2 > 0, k2, 1.0)
        perp = [comps[0] - kx * kdot, comps[1] - ky * kdot, comps[2] - kz * kdot]
        norm = np.sqrt(sum(np.abs(c) ** 2 for c in perp))
        amp = np.where(k2 > 0, np.sqrt(2.0 * model_spectrum(kmag, k0, u0) / (4.0 * math.pi * np.where(k2 > 0, k2, 1.0))), 0.0)
        scale = np.where(norm > 0, amp / np.where(norm > 0, norm, 1.0), 0.0)
    vel = [np.real(np.fft.ifftn(c * scale)) * (nx * ny * nz) for c in perp]
    ms = sum(float(np.mean(v * v)) for v in vel) / 3.0
    f = u0 / math.sqrt(ms)
    return tuple(np.ascontiguousarray(v * f) for v in vel)


def isotropic_state(n, gamma: float = 1.4, mach_t: float = 0.3, k0: float = 4.0, u0: float = 1.0,
                    seed: int = 12345):
    """Interior primitives (rho, u, v, w, p) and the reference Mach number."""
    u, v, w = isotropic_velocity(n, k0, u0, seed)
    ma = mach_t / math.sqrt(3.0)
    p0 = 1.0 / (gamma * ma * ma)
    rho = np.ones_like(u)
    p = np.full_like(u, p0)
    return (rho, u, v, w, p), ma
