"""The drop-in's scalar surface (ns3d_func.h:73-131) and monitoring functions
on the CPU, and a time-advance loop written against the reference's headers
compiling unchanged.

- ns3d_scalar_interpolate / _riemann / _face_flux (the kernels' numerics
  compiled for the host) against the oracle's interpolate_select,
  X_Riemann_solver and WCNS_Riemann_InviscidFlux on seeded random stencils:
  every interpolation with both flags, every solver along every axis, and
  every interpolation x solver x {component-, characteristic-wise} x axis,
  smooth stencils and stencils with a jump; 1e-12 relative to the magnitude
  of the case (the host build forms the
  WCNS weights over one common denominator, the oracle divides each, so the
  bits differ at rounding level);
- the C++ functions with the reference's signatures
  (tests/dropin_scalar_probe.cpp, built here with g++) against the C-ABI,
  bit for bit, including weno5_interpolate / mdcd_*_interpolate /
  weno5_reconstruction / zero_interpolate against interpolate_select of the
  same scheme, the std::runtime_error of a wrong stencil size, and the
  refusal of a normal that is not a coordinate axis;
- compute_diagnostics and write_residuals_tecplot against the definition of
  postprocess.cpp:7-101 and output.cpp:208-242 (res = U - U0 on every cell,
  sqrt(sum over the interior of res^2 dx^3) / max|U|, the table's header at
  step 1 and appended lines after);
- nm -DC: every reference signature of the surface is exported;
- maincycle.cpp of the reference compiles unchanged against
  include/compat/ and links against libns3d_dropin (skipped where the
  reference's sources are absent).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from ns3d_wcns_amd import Interpolation, RiemannSolver, SolverParams, _lib
from oracle.ns3d_oracle import face_flux_batch, interpolate_batch, riemann_batch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ns3d_wcns_amd", "lib")
DROPIN = os.path.join(LIB, "libns3d_dropin.so")
REF_MAINCYCLE = "/root/reference/maincycle.cpp"
GAMMA = 1.4
INTERPS = [Interpolation.WENO5, Interpolation.MDCD_LINEAR, Interpolation.MDCD_HYBRID, Interpolation.WENO5_JS]
SOLVERS = list(RiemannSolver)
DP = ctypes.POINTER(ctypes.c_double)
E_ARG, E_UNSUPPORTED = -1, -2  # NS3D_E_ARG, NS3D_E_UNSUPPORTED (include/ns3d_hip.h)


def dp(a):
    return a.ctypes.data_as(DP)


def P(interp=Interpolation.WENO5, rs=RiemannSolver.HLLC, char=False, diss=0.020, disp=0.0463783):
    return SolverParams(Ma=0.5, Re=100.0, Pr=0.72, interpolation=interp, riemann_solver=rs, char_recon=char,
                        mdcd_diss=diss, mdcd_disp=disp).derive()


def conserved(rho, u, v, w, p, g=GAMMA):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def random_prims(rng, shape, jump=False):
    """rho, u, v, w, p of random states; with jump, a contact and a pressure
    jump at a random cell of the last axis (kept mild enough that every
    interpolated state has a real sound speed)."""
    rho = rng.uniform(0.5, 2.0, shape)
    u, v, w = (rng.uniform(-1.0, 1.0, shape) for _ in range(3))
    p = rng.uniform(0.5, 2.0, shape)
    if jump:
        u, v, w = 0.5 * u, 0.5 * v, 0.5 * w
        p = p + 1.0
        at = rng.integers(1, shape[-1], size=shape[:-1])[..., None]
        right = np.arange(shape[-1]) >= at
        rho = np.where(right, rho * 2.0, rho)
        p = np.where(right, p * 3.0, p)
    return rho, u, v, w, p


def case_scale(*arrays):
    return max(1.0, max(float(np.max(np.abs(a))) for a in arrays))


@pytest.fixture(scope="module")
def lib():
    return _lib.load_library()


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists(DROPIN):
        pytest.skip("g++ or the drop-in library is missing")
    out = str(tmp_path_factory.mktemp("probe") / "probe")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "dropin_scalar_probe.cpp"), "-L", LIB, "-lns3d_dropin",
                    "-lns3d_hip", f"-Wl,-rpath,{LIB}", "-o", out], check=True)
    return out


def run_probe(probe, lines, cwd=None):
    r = subprocess.run([probe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True,
                       cwd=cwd)
    return r.stdout.splitlines(), r.stderr


def fmt(xs):
    return " ".join(repr(float(x)) for x in np.ravel(xs))


# ------------------------------------------------------------------ C-ABI vs oracle
@pytest.mark.parametrize("interp", INTERPS, ids=lambda i: i.name)
def test_scalar_interpolate_matches_oracle(lib, interp):
    rng = np.random.Generator(np.random.PCG64(11 + int(interp)))
    st = np.concatenate([rng.uniform(-1, 1, (200, 6)),                       # rough
                         np.cumsum(rng.uniform(-0.1, 0.1, (200, 6)), axis=1),  # smooth-ish
                         np.where(np.arange(6) >= 3, 5.0, 1.0)[None, :] + 0 * rng.uniform(size=(1, 6))])
    pr = P(interp)
    cp = pr.to_c()
    for flag in (1.0, -1.0):
        ref = interpolate_batch(pr, st, flag)
        got = np.empty(len(st))
        out = ctypes.c_double()
        for t in range(len(st)):
            row = np.ascontiguousarray(st[t])
            assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(row), 6, flag, ctypes.byref(out)) == 0
            got[t] = out.value
        scale = np.maximum(1.0, np.max(np.abs(st), axis=1))
        err = float(np.max(np.abs(got - ref) / scale))
        assert err < 1e-12, (interp.name, flag, err)


def test_scalar_interpolate_zero_and_sizes(lib):
    cp = P(Interpolation.ZERO).to_c()
    out = ctypes.c_double()
    two = np.array([3.0, 7.0])
    assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(two), 2, 1.0, ctypes.byref(out)) == 0
    assert out.value == 3.0
    assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(two), 2, -1.0, ctypes.byref(out)) == 0
    assert out.value == 7.0
    six = np.arange(6.0)
    assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(six), 6, 1.0, ctypes.byref(out)) == E_ARG
    assert b"'ZERO requires stencil size 2, got 6" in lib.ns3d_last_error()
    cp = P(Interpolation.WENO5).to_c()
    assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(six), 5, 1.0, ctypes.byref(out)) == E_ARG
    assert b"'WENO5 requires stencil size 6, got 5" in lib.ns3d_last_error()


@pytest.mark.parametrize("rs", SOLVERS, ids=lambda r: r.name)
def test_scalar_riemann_matches_oracle(lib, rs):
    rng = np.random.Generator(np.random.PCG64(31 + int(rs)))
    n = 300
    UL = conserved(*random_prims(rng, n)).T.copy()
    UR = conserved(*random_prims(rng, n)).T.copy()
    UR[:50] = UL[:50]  # F(U, U) cases
    pr = P(rs=rs)
    cp = pr.to_c()
    for dim in range(3):
        ref = riemann_batch(pr, UL, UR, dim)
        got = np.empty_like(ref)
        for t in range(n):
            a, b, f = UL[t].copy(), UR[t].copy(), np.empty(5)
            assert lib.ns3d_scalar_riemann(ctypes.byref(cp), dp(a), dp(b), dim, dp(f)) == 0
            got[t] = f
        err = float(np.max(np.abs(got - ref) / np.maximum(1.0, np.max(np.abs(ref), axis=1))[:, None]))
        assert err < 1e-12, (rs.name, dim, err)


@pytest.mark.parametrize("char", [False, True], ids=["component", "characteristic"])
@pytest.mark.parametrize("interp", INTERPS, ids=lambda i: i.name)
def test_scalar_face_flux_matches_oracle(lib, interp, char):
    rng = np.random.Generator(np.random.PCG64(71 + 2 * int(interp) + int(char)))
    n = 60
    prim = random_prims(rng, (n, 6), jump=False)
    smooth = tuple(np.cumsum(q * 0.05, axis=1) + q[:, :1] for q in prim)  # gentle variation along the stencil
    rough = random_prims(rng, (n, 6), jump=True)
    prims = tuple(np.concatenate([a, b]) for a, b in zip(smooth, rough))
    prims = (np.abs(prims[0]) + 0.3, prims[1], prims[2], prims[3], np.abs(prims[4]) + 0.3)
    Ut = np.moveaxis(conserved(*prims), 0, 1).copy()  # (n, 5, 6)
    ut = np.moveaxis(np.stack(prims), 0, 1).copy()
    for rs in SOLVERS:
        pr = P(interp, rs, char)
        cp = pr.to_c()
        for dim in range(3):
            ref = face_flux_batch(pr, Ut, ut, dim)
            got = np.empty_like(ref)
            for t in range(len(Ut)):
                a, b, f = Ut[t].copy(), ut[t].copy(), np.empty(5)
                assert lib.ns3d_scalar_face_flux(ctypes.byref(cp), dp(a), dp(b), dim, dp(f)) == 0
                got[t] = f
            assert np.isfinite(ref).all()
            err = float(np.max(np.abs(got - ref) / case_scale(ref)))
            assert err < 1e-12, (interp.name, rs.name, char, dim, err)


def test_scalar_refusals(lib):
    cp = P(Interpolation.WENO5).to_c()
    a, f = np.ones(30), np.empty(5)
    assert lib.ns3d_scalar_face_flux(ctypes.byref(cp), dp(a), dp(a), 3, dp(f)) == E_ARG
    assert lib.ns3d_scalar_riemann(ctypes.byref(cp), dp(a), dp(a), -1, dp(f)) == E_ARG
    cz = P(Interpolation.ZERO).to_c()
    assert lib.ns3d_scalar_face_flux(ctypes.byref(cz), dp(a), dp(a), 0, dp(f)) == E_UNSUPPORTED
    bad = P().to_c()
    bad.riemann_solver = 9
    assert lib.ns3d_scalar_riemann(ctypes.byref(bad), dp(a), dp(a), 0, dp(f)) == E_ARG


# ------------------------------------------------------------------ reference signatures
def capi_interp(lib, pr, st, flag):
    cp = pr.to_c()
    out = ctypes.c_double()
    row = np.ascontiguousarray(st, dtype=np.float64)
    assert lib.ns3d_scalar_interpolate(ctypes.byref(cp), dp(row), len(row), flag, ctypes.byref(out)) == 0
    return out.value


def test_reference_signatures_equal_capi(lib, probe):
    rng = np.random.Generator(np.random.PCG64(5))
    diss, disp = 0.02, 0.0463783
    lines, expect = [], []
    for interp in INTERPS:
        for flag in (1.0, -1.0):
            st = rng.uniform(-1, 1, 6)
            lines.append(f"interp {int(interp)} {flag!r} {diss!r} {disp!r} 6 {fmt(st)}")
            expect.append(capi_interp(lib, P(interp, diss=diss, disp=disp), st, flag))
    # the named interpolations equal interpolate_select of their scheme (+1)
    named = [(0, Interpolation.WENO5), (1, Interpolation.MDCD_LINEAR), (2, Interpolation.MDCD_HYBRID),
             (3, Interpolation.WENO5_JS)]
    for which, interp in named:
        st = rng.uniform(-1, 1, 6)
        lines.append(f"named {which} {diss!r} {disp!r} {fmt(st)}")
        # weno5_interpolate and weno5_reconstruction take the struct defaults
        pr = P(interp, diss=diss, disp=disp) if which in (1, 2) else P(interp)
        expect.append(capi_interp(lib, pr, st, 1.0))
    lines.append(f"named 4 0 0 {fmt([3.5, 9.0, 1, 1, 1, 1])}")
    expect.append(3.5)
    for rs in SOLVERS:
        for dim in range(3):
            L = conserved(*random_prims(rng, ())).ravel()
            R = conserved(*random_prims(rng, ())).ravel()
            nrm = [0.0, 0.0, 0.0]
            nrm[dim] = 1.0
            lines.append(f"riemann {int(rs)} {fmt(nrm)} {GAMMA!r} {fmt(L)} {fmt(R)}")
            cp = P(rs=rs).to_c()
            f = np.empty(5)
            assert lib.ns3d_scalar_riemann(ctypes.byref(cp), dp(L.copy()), dp(R.copy()), dim, dp(f)) == 0
            expect.extend(f)
    for interp in INTERPS:
        for char in (False, True):
            prims = random_prims(rng, (6,), jump=True)
            Ut, ut = conserved(*prims), np.stack(prims)
            rs = SOLVERS[int(interp) % len(SOLVERS)]
            dim = int(interp) % 3
            lines.append(f"face {int(interp)} {int(rs)} {int(char)} {dim} {GAMMA!r} {fmt(Ut)} {fmt(ut)}")
            cp = P(interp, rs, char, diss=0.01).to_c()  # the struct default the probe keeps
            f = np.empty(5)
            assert lib.ns3d_scalar_face_flux(ctypes.byref(cp), dp(Ut.copy()), dp(ut.copy()), dim, dp(f)) == 0
            expect.extend(f)
    out, _ = run_probe(probe, lines)
    got = [float(x) for x in out]
    assert len(got) == len(expect)
    diff = np.nonzero(np.array(got) != np.array(expect))[0]
    assert len(diff) == 0, [(int(i), lines[min(i, len(lines) - 1)][:40], got[i], expect[i]) for i in diff[:5]]


def test_reference_signature_errors(probe):
    out, err = run_probe(probe, ["interp 0 1.0 0.02 0.04 5 1 2 3 4 5",
                                 "interp 1 1.0 0.02 0.04 6 1 2 3 4 5 6",
                                 "riemann 2 0.6 0.8 0 1.4 1 0 0 0 2.5 1 0 0 0 2.5"])
    assert out[0] == "throw interpolate_select: interpolation 'WENO5 requires stencil size 6, got 5"
    assert out[1] == "throw interpolate_select: interpolation 'ZERO requires stencil size 2, got 6"
    assert [float(x) for x in out[2:7]] == [0.0] * 5  # Fface untouched
    assert "is not (1,0,0), (0,1,0) or (0,0,1)" in err


def cell_averages(coef, cells, lo):
    """Averages of the polynomial sum coef[m] x^m over the unit cells
    [lo + j, lo + j + 1], j = 0 .. cells - 1."""
    out = []
    for j in range(cells):
        a, b = lo + j, lo + j + 1
        out.append(sum(c * (b ** (m + 1) - a ** (m + 1)) / (m + 1) for m, c in enumerate(coef)))
    return out


def test_reconstruction_functions(probe):
    """reconstruction.cpp:13-124 (not on the reference's path): c6th / c4th
    reproduce the interface value p(0) of polynomials of degree 5 / 3 from
    their cell averages (and miss degree 6 / 4, so the orders are pinned);
    mdcd_reconstruction keeps the reference's integer-quotient constants, so
    a constant stencil gives 0 and its coefficients are (0, -disp - 2 diss,
    disp + 5 diss, disp - 5 diss, -disp + 2 diss, 0); reconstruct_select
    dispatches on P.recon with the reference's reversal, size errors and
    fallback mean."""
    rng = np.random.Generator(np.random.PCG64(77))
    lines, want = [], []
    for deg, exact in ((5, True), (6, False)):
        coef = rng.uniform(-1, 1, deg + 1)
        lines.append(f"recon 3 0 1 0 0 6 {fmt(cell_averages(coef, 6, -3))}")
        want.append((coef[0], exact))
    for deg, exact in ((3, True), (4, False)):
        coef = rng.uniform(-1, 1, deg + 1)
        lines.append(f"recon 4 0 1 0 0 4 {fmt(cell_averages(coef, 4, -2))}")
        want.append((coef[0], exact))
    out, _ = run_probe(probe, lines)
    for got, (p0, exact) in zip(out, want):
        assert (abs(float(got) - p0) < 1e-13) == exact, (got, p0, exact)
    diss, disp = 0.02, 0.0463783
    expect_c = [0.0, -disp - 2 * diss, disp + 5 * diss, disp - 5 * diss, -disp + 2 * diss, 0.0]
    lines = [f"recon 2 0 1 {diss!r} {disp!r} 6 {fmt(np.eye(6)[m])}" for m in range(6)]
    lines.append(f"recon 2 0 1 {diss!r} {disp!r} 6 {fmt(np.full(6, 3.7))}")
    out, _ = run_probe(probe, lines)
    np.testing.assert_allclose([float(x) for x in out[:6]], expect_c, rtol=0, atol=1e-17)
    assert abs(float(out[6])) < 1e-15
    st = rng.uniform(-1, 1, 6)
    WENO5, WCNS, LINEAR, MDCD = 0, 1, 2, 3  # SolverParams::Reconstruction
    out, _ = run_probe(probe, [f"recon 0 {WENO5} -1 0.01 0.0463783 6 {fmt(st)}",
                               f"named 3 0 0 {fmt(st[::-1])}",
                               f"recon 0 {MDCD} 1 {diss!r} {disp!r} 6 {fmt(st)}",
                               f"recon 2 0 1 {diss!r} {disp!r} 6 {fmt(st)}",
                               f"recon 0 {LINEAR} -1 0 0 2 {fmt([4.0, 9.0])}",
                               f"recon 0 {WCNS} 1 0 0 6 {fmt(st)}",
                               f"recon 0 {WCNS} -1 0 0 6 {fmt(st)}",
                               f"recon 0 {WENO5} 1 0 0 5 {fmt(st[:5])}",
                               f"recon 0 {LINEAR} 1 0 0 6 {fmt(st)}"])
    assert out[0] == out[1]  # flag -1 reverses, then weno5_reconstruction
    assert out[2] == out[3]
    assert float(out[4]) == 9.0
    assert float(out[5]) == 0.5 * (st[3] + st[2]) and float(out[6]) == 0.5 * (st[3] + st[4])
    assert out[7] == "throw reconstruct_select: reconstruction 'WENO5 requires stencil size 6, got 5"
    assert out[8] == "throw reconstruct_select: reconstruction 'LINEAR requires stencil size 2, got 6"


# ------------------------------------------------------------------ monitoring
def lcg_fields(n, seed):
    """The probe's deterministic fields: U0 = base + r - 1/2, U = U0 + 1e-3 (r' - 1/2)."""
    mask = (1 << 64) - 1
    s = seed
    U, U0 = np.empty((5, n)), np.empty((5, n))
    for q in range(5):
        base = 1.0 if q in (0, 4) else 0.0
        for i in range(n):
            s = (s * 6364136223846793005 + 1442695040888963407) & mask
            U0[q, i] = base + (s >> 11) / 9007199254740992.0 - 0.5
            s = (s * 6364136223846793005 + 1442695040888963407) & mask
            U[q, i] = U0[q, i] + 1e-3 * ((s >> 11) / 9007199254740992.0 - 0.5)
    return U, U0


def test_compute_diagnostics_and_residual_table(probe, tmp_path):
    nx, ny, nz, ng, dx, seed = 5, 4, 3, 3, 0.25, 9
    out, _ = run_probe(probe, [f"diag {nx} {ny} {nz} {ng} {dx!r} 1 {seed}", f"diag {nx} {ny} {nz} {ng} {dx!r} 2 {seed}"],
                       cwd=str(tmp_path))
    vals = [float(x) for x in out]
    sx, sy, sz = nx + 2 * ng, ny + 2 * ng, nz + 2 * ng
    U, U0 = lcg_fields(sx * sy * sz, seed)
    shape = (5, sz, sy, sx)
    Ui = U.reshape(shape)[:, ng:ng + nz, ng:ng + ny, ng:ng + nx]
    Ri = (U - U0).reshape(shape)[:, ng:ng + nz, ng:ng + ny, ng:ng + nx]
    want = np.sqrt(np.sum(Ri ** 2 * dx ** 3, axis=(1, 2, 3))) / np.max(np.abs(Ui), axis=(1, 2, 3))
    np.testing.assert_allclose(vals[:5], want, rtol=1e-13)
    assert vals[5] == sx * sy * sz  # res_* cover every cell
    assert vals[6] == 0.0  # res = U - U0 exactly
    table = (tmp_path / "output" / "residuals_probe.dat").read_text().splitlines()
    assert table[0] == 'TITLE = "Residuals vs Step"'
    assert table[1] == 'VARIABLES = "Step" "Res_rho" "Res_rhou" "Res_rhov" "Res_rhow" "Res_E" "Etot"'
    assert len(table) == 4 and table[2].startswith("1 ") and table[3].startswith("2 ")
    cols = [float(x) for x in table[3].split()[1:]]
    np.testing.assert_allclose(cols[:5], want, rtol=1e-8)
    assert cols[5] == 0.25


# ------------------------------------------------------------------ exports and compile
def test_dropin_exports_scalar_and_monitor_signatures():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", DROPIN], capture_output=True, text=True, check=True).stdout
    vec = "std::vector<double, std::allocator<double> >"
    vvec = f"std::vector<{vec}, std::allocator<{vec} > >"
    string = "std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> >"
    sigs = [f"WCNS_Riemann_InviscidFlux({vec}&, {vvec} const&, {vvec} const&, SolverParams const&, int)",
            f"interpolate_select({vec} const&, double, SolverParams)",
            "weno5_interpolate(std::array<double, 6ul> const&)",
            "zero_interpolate(std::array<double, 2ul> const&)",
            "mdcd_linear_interpolate(std::array<double, 6ul> const&, SolverParams)",
            "mdcd_hybrid_interpolate(std::array<double, 6ul> const&, SolverParams)",
            "weno5_reconstruction(std::array<double, 6ul> const&)",
            f"reconstruct_select({vec} const&, double, SolverParams)",
            "linear_reconstruction(std::array<double, 2ul> const&)",
            "mdcd_reconstruction(std::array<double, 6ul> const&, SolverParams)",
            "c6th_reconstruction(std::array<double, 6ul> const&)",
            "c4th_reconstruction(std::array<double, 4ul> const&)",
            "compute_diagnostics(Field3D&, SolverParams const&, GridDesc const&)",
            "compute_total_energy(Field3D&, GridDesc const&, CartDecomp const&, SolverParams const&)",
            f"write_residuals_tecplot(Field3D const&, int, {string} const&)",
            f"compute_energy_spectrum(Field3D const&, GridDesc const&, CartDecomp const&, {string} const&)",
            "compute_turbulence_statistics(Field3D&, GridDesc const&, SolverParams const&, CartDecomp const&, "
            "double)",
            "isotropic_post_process(Field3D&, GridDesc const&, CartDecomp const&, SolverParams const&, double)",
            "Field3D::updateResiduals()"]
    for name in ("Roe", "Rusanov", "HLLC", "HLLC_p", "HLL", "AUSM"):
        sigs.append(f"{name}_Riemann_solver({vec}&, {vec} const&, {vec} const&, double, double, double, double)")
    for sig in sigs:
        assert sig in out, sig


def test_reference_maincycle_compiles_unchanged(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists(REF_MAINCYCLE):
        pytest.skip("g++ or the reference's maincycle.cpp is absent")
    # compiled from a copy, so that "ns3d_func.h" resolves to include/compat/
    # and not to the reference's own header beside it
    src = tmp_path / "maincycle.cpp"
    shutil.copyfile(REF_MAINCYCLE, src)
    main = tmp_path / "main.cpp"
    main.write_text('#include "ns3d_func.h"\nint main() { return 0; }\n')
    compat = os.path.join(ROOT, "include", "compat")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", compat, str(src), str(main), "-L", LIB,
                    "-lns3d_dropin", "-lns3d_hip", f"-Wl,-rpath,{LIB}", "-o", str(tmp_path / "loop")], check=True)


# ------------------------------------------------------------------ post-processing (device)
def post_field(n, p0):
    """The probe's analytic field on the cell centres of a periodic n^3 box
    of side 2 pi, arrays indexed [z, y, x]."""
    x = (np.arange(n) + 0.5) * (2 * np.pi / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rho = 1.0 + 0.1 * np.sin(X) * np.cos(Y)
    u = np.sin(X) * np.cos(Y) * np.cos(Z) + 0.1 * np.sin(2.0 * Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = 0.05 * np.cos(3.0 * X)
    p = p0 + 0.1 * np.cos(X + Y)
    return rho, u, v, w, p


@pytest.mark.gpu
def test_isotropic_post_process_files(gpu, probe, tmp_path):
    """isotropic_post_process through the drop-in (faithful mode) on an
    analytic 16^3 field: the spectrum file against numpy's FFT and shell sums
    (EnergySpectrum.cpp:63-134), 1e-12 of its largest value; every column of
    output/turbulence_stats.dat against its definition
    (ana_isotropic.cpp:196-340) to the file's 8 significant digits."""
    n, t = 16, 0.125
    print(f"isotropic_post_process {n}^3", flush=True)
    out, err = run_probe(probe, [f"post {n} {t!r}"], cwd=str(tmp_path))
    assert "ns3d:" not in err, err
    g, ma, re, s_ref = 1.4, 0.4, 500.0, 110.4 / 288.15  # SolverParams defaults
    rgas = 1.0 / (ma * ma * g)
    rho, u, v, w, p = post_field(n, 1.0 / (g * ma * ma))
    T = p / (rho * rgas)
    assert out[0].startswith("[Post] nu = ")  # the reference's console line (ana_isotropic.cpp:296-301)
    assert abs(float(out[-1]) - T[0, 0, 0]) <= 1e-15 * T[0, 0, 0]
    # spectrum
    N = n ** 3
    E = sum(0.5 * np.abs(np.fft.fftn(q) / N) ** 2 for q in (u, v, w))
    k = np.fft.fftfreq(n, 1.0 / n)
    K3, K2, K1 = np.meshgrid(k, k, k, indexing="ij")
    shell = np.floor(np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2) + 0.5).astype(int)
    Ek = np.bincount(shell.ravel(), weights=E.ravel(), minlength=2 * n + 1)
    rows = np.loadtxt(tmp_path / "output" / f"output_time_{t:.5f}_spectrum.dat")
    assert rows.shape == (2 * n + 1, 2)
    assert np.array_equal(rows[:, 0], np.arange(2 * n + 1))
    assert float(np.max(np.abs(rows[:, 1] - Ek))) <= 1e-12 * float(np.max(Ek))
    # statistics
    mu = (1.0 / re) * T ** 1.5 * (1.0 + s_ref) / (T + s_ref)
    nu = mu.mean() / rho.mean()
    ke = 0.5 * np.mean(u * u + v * v + w * w)
    u_rms = np.sqrt(np.mean(u * u + v * v + w * w) / 3.0)
    tke = 0.5 * sum(np.mean(q * q) - np.mean(q) ** 2 for q in (u, v, w))
    eps = float(np.sum(2.0 * nu * np.arange(2 * n + 1) ** 2 * Ek))
    h = 2 * np.pi / n

    def sh(a, m):
        return np.roll(a, -m, axis=2)

    dudx = (sh(u, 3) - 9 * sh(u, 2) + 45 * sh(u, 1) - 45 * sh(u, -1) + 9 * sh(u, -2) - sh(u, -3)) / (60 * h)
    taylor_li = u_rms / np.sqrt(np.mean(dudx ** 2))
    mach_t = np.sqrt(2 * ke) / np.mean(np.sqrt(g * rgas * T))
    want = [t, ke, tke, u_rms, eps, np.sqrt(15 * nu * u_rms ** 2 / eps), (nu ** 3 / eps) ** 0.25,
            (nu * eps) ** 0.25, np.sqrt(nu / eps), taylor_li, u_rms * taylor_li / nu, mach_t]
    lines = (tmp_path / "output" / "turbulence_stats.dat").read_text().splitlines()
    assert lines[0].split() == ["current_time", "KineticEnergy", "TKE", "u_rms", "Dissipation", "Taylor",
                                "Kol_scale", "Kol_vel", "Kol_time", "Taylor_Li", "Re_lambda", "Mach_t"]
    got = [float(x) for x in lines[1].split()]
    np.testing.assert_allclose(got, want, rtol=2e-8)
