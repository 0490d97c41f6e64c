// ns3d_flops.hip -- algorithmic floating-point operation counts of the RHS +
// RK3 path, for the roofline report.
//
// The face flux is counted exactly: the device numerics (ns3d_device.h) are
// instantiated on a scalar type that counts every +, -, *, /, sqrt, fabs,
// fmin and fmax (one each; negation and comparisons are free), on a smooth
// subsonic stencil so that the Riemann solvers take their common branch.
// The cell-wise parts (half-node differences, viscous terms, stage update,
// CFL) are straight-line formulas whose counts are written out below next to
// the expressions they count.
#include <cmath>

#include "../../include/ns3d_hip.h"
#include "ns3d_device.h"

namespace {

thread_local long g_ops = 0;

struct CF {
    double v;
    CF() : v(0) {}
    CF(double x) : v(x) {}
    explicit operator double() const { return v; }
};
inline CF operator+(CF a, CF b) { ++g_ops; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { ++g_ops; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { ++g_ops; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { ++g_ops; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline CF &operator+=(CF &a, CF b) { a = a + b; return a; }
inline CF &operator*=(CF &a, CF b) { a = a * b; return a; }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline CF sqrt(CF a) { ++g_ops; return CF(std::sqrt(a.v)); }
inline CF fabs(CF a) { ++g_ops; return CF(std::fabs(a.v)); }
inline CF fmin(CF a, CF b) { ++g_ops; return CF(std::fmin(a.v, b.v)); }
inline CF fmax(CF a, CF b) { ++g_ops; return CF(std::fmax(a.v, b.v)); }

template <int INTERP, int RS, bool CHAR>
long count_face(const ns3d::SchemeCoef<CF> &c) {
    // smooth subsonic 6-point stencil, x normal (rho, rho u, rho v, rho w, E)
    CF s[5][6], pc[2][5];
    const double g = c.gamma.v;
    for (int m = 0; m < 6; ++m) {
        const double rho = 1.0 + 0.01 * m, u = 0.3 + 0.02 * m, v = -0.2 + 0.01 * m, w = 0.1, p = 6.0 + 0.05 * m;
        s[0][m] = rho;
        s[1][m] = rho * u;
        s[2][m] = rho * v;
        s[3][m] = rho * w;
        s[4][m] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    }
    for (int side = 0; side < 2; ++side) {
        const int m = 2 + side;
        const double rho = s[0][m].v, u = s[1][m].v / rho, v = s[2][m].v / rho, w = s[3][m].v / rho;
        pc[side][0] = rho;
        pc[side][1] = u;
        pc[side][2] = v;
        pc[side][3] = w;
        pc[side][4] = (g - 1.0) * (s[4][m].v - 0.5 * rho * (u * u + v * v + w * w));
    }
    CF F[5];
    g_ops = 0;
    if constexpr (CHAR) {
        // the pencils derive the cell primitives from U (ns3d_pencil_impl.h prim_of)
        for (int side = 0; side < 2; ++side) {
            CF u[5] = {s[0][2 + side], s[1][2 + side], s[2][2 + side], s[3][2 + side], s[4][2 + side]};
            const CF ir = CF(1) / u[0];
            pc[side][1] = u[1] * ir;
            pc[side][2] = u[2] * ir;
            pc[side][3] = u[3] * ir;
            pc[side][4] = (c.gamma - CF(1)) *
                          (u[4] - CF(0.5) * u[0] * (pc[side][1] * pc[side][1] + pc[side][2] * pc[side][2] +
                                                    pc[side][3] * pc[side][3]));
        }
    }
    ns3d::face_flux<INTERP, RS, CHAR, 1>(c, s, pc, F);
    return g_ops;
}

template <int INTERP, bool CHAR>
long count_face_rs(int rs, const ns3d::SchemeCoef<CF> &c) {
    switch (rs) {
        case ns3d::R_ROE: return count_face<INTERP, ns3d::R_ROE, CHAR>(c);
        case ns3d::R_RUSANOV: return count_face<INTERP, ns3d::R_RUSANOV, CHAR>(c);
        case ns3d::R_HLLC: return count_face<INTERP, ns3d::R_HLLC, CHAR>(c);
        case ns3d::R_HLL: return count_face<INTERP, ns3d::R_HLL, CHAR>(c);
        case ns3d::R_HLLC_P: return count_face<INTERP, ns3d::R_HLLC_P, CHAR>(c);
        default: return count_face<INTERP, ns3d::R_AUSM, CHAR>(c);
    }
}

template <bool CHAR>
long count_face_all(int interp, int rs, const ns3d::SchemeCoef<CF> &c) {
    switch (interp) {
        case ns3d::I_WENO5: return count_face_rs<ns3d::I_WENO5, CHAR>(rs, c);
        case ns3d::I_MDCD_LINEAR: return count_face_rs<ns3d::I_MDCD_LINEAR, CHAR>(rs, c);
        case ns3d::I_MDCD_HYBRID: return count_face_rs<ns3d::I_MDCD_HYBRID, CHAR>(rs, c);
        default: return count_face_rs<ns3d::I_WENO5_JS, CHAR>(rs, c);
    }
}

}  // namespace

extern "C" int ns3d_algorithmic_flops(const ns3d_params *p, ns3d_flops *out) {
    if (!p || !out) return NS3D_E_ARG;
    if (p->interpolation == NS3D_INTERP_ZERO || p->interpolation < 0 || p->interpolation > 4 ||
        p->riemann_solver < 0 || p->riemann_solver > 5)
        return NS3D_E_ARG;
    ns3d::SchemeCoef<CF> c;
    c.gamma = p->gamma;
    c.p_ref = 0.0;
    c.H_ref = 0.0;
    c.E_ref = 0.0;
    const double diss = p->mdcd_diss, disp = p->mdcd_disp;
    const double lin[6] = {(3.0 * disp + 3.0 * diss) / 8.0, (-18.0 * disp - 30 * diss - 1.0) / 16.0,
                           (12.0 * disp + 60.0 * diss + 9.0) / 16.0, (12.0 * disp - 60.0 * diss + 9.0) / 16.0,
                           (-18.0 * disp + 30.0 * diss - 1.0) / 16.0, (3.0 * disp - 3.0 * diss) / 8.0};
    for (int q = 0; q < 6; ++q) c.mdcd_lin[q] = lin[q];
    c.mdcd_d[0] = 1.5 * (disp + diss);
    c.mdcd_d[1] = 0.5 - 1.5 * (disp - 3 * diss);
    c.mdcd_d[2] = 0.5 - 1.5 * (disp + 3 * diss);
    c.mdcd_d[3] = 1.5 * (disp - diss);
    out->face = (double)(p->char_recon ? count_face_all<true>(p->interpolation, p->riemann_solver, c)
                                       : count_face_all<false>(p->interpolation, p->riemann_solver, c));
    // half-node difference per variable and direction: 3 differences, 3
    // products, 2 sums, 1 scaling = 9; x 5 variables x 3 directions, plus 2
    // sums per variable to combine the directions
    out->diff = 9.0 * 5 * 3 + 2.0 * 5;
    // viscous terms per cell:
    //   12 gradients, 6th-order central: 4 products, 5 sums, 1 scaling = 120
    //   Sutherland mu: sqrt, 3 products, 2 sums, 1 division = 7
    //   divergence 2; normal stresses 3 x 5; shear stresses 3 x 2; mu Cp/Pr 1
    //   energy fluxes 3 x (4 products + 3 sums) = 21
    //   divergence of 12 flux components: 12 x 10, plus 8 sums = 128
    out->visc_flux = 120 + 7 + 2 + 15 + 6 + 1 + 21;
    out->visc_div = 128;
    // stage update: 5 x (u0 + dt r: 2 | a u0 + b (u + dt r): 5), averaged over
    // the three stages = 5 x 4; C2P/P2C: 1 division, 3 products (velocity),
    // 6 (kinetic), 4 (pressure), 2 + 1 division (temperature), 3 products
    // (momentum), 4 (energy)
    out->update = 5.0 * 4 + (1 + 3 + 6 + 4 + 3 + 3 + 4);
    // CFL: u (1), p (5), a (3), |u| + a (2), max (1)
    out->cfl = 12;
    out->per_stage = 3.0 * out->face + out->diff + out->visc_flux + out->visc_div + out->update;
    out->per_step = 3.0 * out->per_stage + out->cfl;
    return NS3D_OK;
}
