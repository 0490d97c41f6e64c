"""N-step parity of the observables: the device path against the CPU checker
on the 64^3 isotropic-turbulence case (Ma_t = 0.3, Re = 100, WCNS5 + HLLC,
CFL 0.5, 100 RK3 steps), and the device kinetic-energy sum against the host.

Tolerances (norm-wise relative): the dt sequence 1e-12; the conserved fields
1e-10; kinetic energy, TKE and the spectral dissipation 1e-11; E(k) 1e-9 for
every shell holding at least 1e-12 of the peak shell energy.  The device
observables (hipFFT spectrum, fixed-order moment sums) against the host
formulas on the same fields: 1e-12.
"""
import numpy as np
import pytest

from cases import box, params
from ns3d_wcns_amd import DeviceField
from ns3d_wcns_amd.isotropic import isotropic_state
from ns3d_wcns_amd.observables import energy_spectrum, turbulence_statistics
from oracle.ns3d_oracle import OracleField, set_threads

pytestmark = pytest.mark.gpu


def test_isotropic_64_cubed_100_steps(gpu):
    n, steps = 64, 100
    prim, ma = isotropic_state(n, mach_t=0.3, seed=12345)
    P = params(ma)
    G = box(n)
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    dts = dev.time_advance(steps)
    set_threads(16)
    ref = OracleField(G, P)
    ref.initialize_primitive(*prim)
    rdts = ref.time_advance(steps)
    print(f"{steps} steps, t = {dts.sum():.6f}", flush=True)
    assert len(dts) == len(rdts) == steps
    assert np.max(np.abs(dts - rdts) / rdts) < 1e-12
    snap = {name: dev.interior(dev.download(name)) for name in ("rho", "rhou", "rhov", "rhow", "E", "u", "v", "w", "T")}
    for name in ("rho", "rhou", "rhov", "rhow", "E"):
        b = ref.interior(ref.view(name))
        err = np.max(np.abs(snap[name] - b)) / np.max(np.abs(b))
        print(name, f"{err:.2e}")
        assert err < 1e-10, (name, err)
    ke_dev = dev.total_kinetic_energy()
    ke_ref = ref.total_kinetic()
    assert abs(ke_dev - ke_ref) <= 1e-11 * ke_ref
    rv = {name: ref.interior(ref.view(name)) for name in ("rho", "u", "v", "w", "T")}
    sd = turbulence_statistics(snap["rho"], snap["u"], snap["v"], snap["w"], snap["T"], P)
    sr = turbulence_statistics(rv["rho"], rv["u"], rv["v"], rv["w"], rv["T"], P)
    for k in ("kinetic_energy", "tke", "dissipation"):
        assert abs(sd[k] - sr[k]) <= 1e-11 * abs(sr[k]), (k, sd[k], sr[k])
    ed, er = sd["spectrum"], sr["spectrum"]
    big = er >= 1e-12 * er.max()
    shell_err = np.abs(ed[big] - er[big]) / er[big]
    print(f"E(k): max shell error {shell_err.max():.2e} over {big.sum()} shells")
    assert shell_err.max() < 1e-9
    dev.close()
    ref.close()


def test_device_kinetic_energy_matches_host_sum(gpu):
    n = 32
    prim, ma = isotropic_state(n, mach_t=0.5, seed=3)
    P = params(ma)
    G = box(n)
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    dev.advance_steps(3)
    rho, u, v, w = (dev.interior(dev.download(q)) for q in ("rho", "u", "v", "w"))
    host = float(np.sum(0.5 * rho * (u * u + v * v + w * w))) * G.dx * G.dy * G.dz
    assert abs(dev.total_kinetic_energy() - host) <= 1e-12 * host
    # spectrum normalisation: sum_k E(k) = 0.5 <|u|^2> (Parseval, 1/N forward transform)
    ek = energy_spectrum(u, v, w)
    assert abs(ek.sum() - 0.5 * np.mean(u * u + v * v + w * w)) <= 1e-12 * ek.sum()
    dev.close()


@pytest.mark.parametrize("n", [32, (24, 20, 16)])
def test_device_spectrum_and_statistics_match_host(gpu, n):
    prim, ma = state_for(n)
    P = params(ma)
    G = box(n)
    dev = DeviceField(G, P)
    dev.initialize_primitive(*prim)
    dev.advance_steps(2)
    f = {q: dev.interior(dev.download(q)) for q in ("rho", "u", "v", "w", "T")}
    host = turbulence_statistics(f["rho"], f["u"], f["v"], f["w"], f["T"], P)
    ek = dev.energy_spectrum()
    eh = host["spectrum"]
    big = eh >= 1e-12 * eh.max()
    assert np.max(np.abs(ek[big] - eh[big]) / eh[big]) < 1e-12
    assert np.all(np.abs(ek[~big]) <= 1e-12 * eh.max())
    st = dev.turbulence_statistics()
    for k in ("kinetic_energy", "tke", "u_rms", "nu", "dissipation"):
        assert abs(st[k] - host[k]) <= 1e-12 * abs(host[k]), (k, st[k], host[k])
    dev.close()


def state_for(n):
    from cases import state

    return state(n, 0.5, noise=0.01, seed=17)
