"""Turbulence observables of a periodic box (host side, numpy).

- total kinetic energy sum 0.5 rho |u|^2 dV (compute_total_energy,
  postprocess.cpp:103-128);
- E(k): forward FFT normalised by 1/N, 0.5 (|U|^2 + |V|^2 + |W|^2), summed in
  shells lround(|k|), k = 0..2*NX (compute_energy_spectrum_rank0,
  EnergySpectrum.cpp:63-156);
- TKE, u_rms, nu = <mu>/<rho> and the spectral dissipation 2 nu sum k^2 E(k)
  (compute_turbulence_statistics, ana_isotropic.cpp:238-305).
"""
from __future__ import annotations

import numpy as np


def total_kinetic_energy(rho, u, v, w, dV):
    return float(np.sum(0.5 * rho * (u * u + v * v + w * w) * dV))


def energy_spectrum(u, v, w):
    nz, ny, nx = u.shape
    n = nx * ny * nz
    e = np.zeros(u.shape)
    for c in (u, v, w):
        f = np.fft.fftn(c) / n
        e += 0.5 * (f.real ** 2 + f.imag ** 2)
    k1 = np.fft.fftfreq(nx, 1.0 / nx)
    k2 = np.fft.fftfreq(ny, 1.0 / ny)
    k3 = np.fft.fftfreq(nz, 1.0 / nz)
    # the reference labels i <= N/2 as +i: the Nyquist index counts as +N/2
    k1[nx // 2] = nx // 2
    k2[ny // 2] = ny // 2
    k3[nz // 2] = nz // 2
    kk = np.sqrt(k3[:, None, None] ** 2 + k2[None, :, None] ** 2 + k1[None, None, :] ** 2)
    shell = np.floor(kk + 0.5).astype(np.int64)  # lround for non-negative values
    ek = np.bincount(shell.ravel(), weights=e.ravel(), minlength=2 * nx + 1)
    return ek[: 2 * nx + 1]


def turbulence_statistics(rho, u, v, w, T, params):
    n = u.size
    mu = params.mu * T ** 1.5 * (1.0 + params.S_ref) / (T + params.S_ref)
    mean_mu = float(np.sum(mu)) / n
    mean_rho = float(np.sum(rho)) / n
    nu = mean_mu / mean_rho
    ke = float(np.sum(0.5 * (u * u + v * v + w * w))) / n
    urms = float(np.sqrt(np.sum(u * u + v * v + w * w) / (3 * n)))
    var = [float(np.sum(c * c)) / n - (float(np.sum(c)) / n) ** 2 for c in (u, v, w)]
    tke = 0.5 * sum(var)
    ek = energy_spectrum(u, v, w)
    k = np.arange(ek.size, dtype=np.float64)
    eps = float(np.sum(2.0 * nu * k * k * ek))
    return {"kinetic_energy": ke, "tke": tke, "u_rms": urms, "nu": nu, "dissipation": eps, "spectrum": ek}


def compare_runs(test: dict, ref: dict, params, dV: float) -> dict:
    """Errors of one run against another started from the same state (e.g.
    the mixed method against the plain FP64 one).  test/ref: interior fields
    by name (rho, rhou, rhov, rhow, E, u, v, w, p, T).

    - fields: norm-wise relative error max|a - b| / max|b| per conserved field;
    - kinetic energy (sum 0.5 rho |u|^2 dV), TKE and dissipation: relative error;
    - spectrum: |E_a(k) - E_b(k)| / E_b(k) per shell k = 1..N/2, and the largest
      over shells holding at least 1e-6 of the peak shell energy.
    """
    out = {"fields": {}}
    for name in ("rho", "rhou", "rhov", "rhow", "E"):
        b = ref[name]
        out["fields"][name] = float(np.max(np.abs(test[name] - b)) / np.max(np.abs(b)))
    ka = total_kinetic_energy(test["rho"], test["u"], test["v"], test["w"], dV)
    kb = total_kinetic_energy(ref["rho"], ref["u"], ref["v"], ref["w"], dV)
    out["kinetic_energy"] = abs(ka - kb) / abs(kb)
    sa = turbulence_statistics(test["rho"], test["u"], test["v"], test["w"], test["T"], params)
    sb = turbulence_statistics(ref["rho"], ref["u"], ref["v"], ref["w"], ref["T"], params)
    out["tke"] = abs(sa["tke"] - sb["tke"]) / abs(sb["tke"])
    out["dissipation"] = abs(sa["dissipation"] - sb["dissipation"]) / abs(sb["dissipation"])
    ea, eb = sa["spectrum"], sb["spectrum"]
    kmax = ref["u"].shape[2] // 2
    shell = [float(abs(ea[k] - eb[k]) / eb[k]) if eb[k] > 0 else 0.0 for k in range(1, kmax + 1)]
    out["spectrum_per_shell"] = shell
    big = [e for k, e in enumerate(shell, start=1) if eb[k] >= 1e-6 * eb[1:].max()]
    out["spectrum_max"] = max(big) if big else 0.0
    return out
