// ns3d_device.h -- gfx950 device numerics for the WCNS face flux.
//
// Every face is evaluated in a rotated frame in which the face normal is the
// first velocity component: conserved vectors are (rho, rho*un, rho*ut1,
// rho*ut2, E) with (un, ut1, ut2) = (u, v, w) for x faces, (v, w, u) for y
// faces and (w, u, v) for z faces.  With the reference's axis-aligned normals
// (n = e_dim, InviscousFlux.cpp:160-162) its formulas reduce to these x-normal
// forms; the cyclic permutation maps the reference's tangential
// characteristic rows onto the x-normal ones with identical scaling
// (build_eigen_matrices, InviscousFlux.cpp:888-987), so the interpolated
// characteristic variables are the same quantities.
//
// Every function is templated on the arithmetic type R (double for the plain
// method, float for the mixed one).
#pragma once
#include <hip/hip_runtime.h>

// The numerics are host-callable too, so that the library can count their
// floating-point operations on a counting scalar type (ns3d_flops.cpp).
#define NS3D_HD __host__ __device__ __forceinline__

namespace ns3d {

enum { I_WENO5 = 0, I_ZERO = 1, I_MDCD_LINEAR = 2, I_MDCD_HYBRID = 3, I_WENO5_JS = 4 };
enum { R_ROE = 0, R_RUSANOV = 1, R_HLLC = 2, R_HLL = 3, R_HLLC_P = 4, R_AUSM = 5 };

// Scheme constants handed to the face kernels.
template <class R>
struct SchemeCoef {
    R gamma;
    R mdcd_lin[6];  // linear MDCD weights (interpolation.cpp:128-132)
    R mdcd_d[4];    // MDCD-WENO linear weights d0..d3 (interpolation.cpp:182-185)
};

// ------------------------------------------------------------ interpolation
// Value at i+1/2 from f0..f5 = cells i-2..i+3 (left-biased; the right state
// passes the mirrored stencil).

// WCNS5 nonlinear weights 1/16, 5/8, 5/16 with eps = 1e-6
// (weno5_interpolate, interpolation.cpp:86-115).  alpha_k = d_k / q_k with
// q_k = (eps + beta_k)^2; the weighted sum is formed over the common
// denominator q0*q1*q2, one division instead of six.
template <class R>
NS3D_HD R wcns5(R f0, R f1, R f2, R f3, R f4) {
    const R eps = R(1e-6);
    const R t0 = f0 - R(2) * f1 + f2, s0 = f0 - R(4) * f1 + R(3) * f2;
    const R t1 = f1 - R(2) * f2 + f3, s1 = f1 - f3;
    const R t2 = f2 - R(2) * f3 + f4, s2 = R(3) * f2 - R(4) * f3 + f4;
    const R b0 = R(13.0 / 12.0) * t0 * t0 + R(0.25) * s0 * s0;
    const R b1 = R(13.0 / 12.0) * t1 * t1 + R(0.25) * s1 * s1;
    const R b2 = R(13.0 / 12.0) * t2 * t2 + R(0.25) * s2 * s2;
    const R q0 = (eps + b0) * (eps + b0);
    const R q1 = (eps + b1) * (eps + b1);
    const R q2 = (eps + b2) * (eps + b2);
    const R e0 = R(1.0 / 16.0) * (q1 * q2);
    const R e1 = R(5.0 / 8.0) * (q0 * q2);
    const R e2 = R(5.0 / 16.0) * (q0 * q1);
    const R p0 = R(3) * f0 - R(10) * f1 + R(15) * f2;  // 8 x candidate 0
    const R p1 = -f1 + R(6) * f2 + R(3) * f3;
    const R p2 = R(3) * f2 + R(6) * f3 - f4;
    return (e0 * p0 + e1 * p1 + e2 * p2) / (R(8) * (e0 + e1 + e2));
}

// Jiang-Shu WENO5 weights 0.1, 0.6, 0.3 (weno5_reconstruction,
// reconstruction.cpp:69-100), same common-denominator form.
template <class R>
NS3D_HD R weno5js(R f0, R f1, R f2, R f3, R f4) {
    const R eps = R(1e-6);
    const R t0 = f0 - R(2) * f1 + f2, s0 = f0 - R(4) * f1 + R(3) * f2;
    const R t1 = f1 - R(2) * f2 + f3, s1 = f1 - f3;
    const R t2 = f2 - R(2) * f3 + f4, s2 = R(3) * f2 - R(4) * f3 + f4;
    const R b0 = R(13.0 / 12.0) * t0 * t0 + R(0.25) * s0 * s0;
    const R b1 = R(13.0 / 12.0) * t1 * t1 + R(0.25) * s1 * s1;
    const R b2 = R(13.0 / 12.0) * t2 * t2 + R(0.25) * s2 * s2;
    const R q0 = (eps + b0) * (eps + b0);
    const R q1 = (eps + b1) * (eps + b1);
    const R q2 = (eps + b2) * (eps + b2);
    const R e0 = R(0.1) * (q1 * q2);
    const R e1 = R(0.6) * (q0 * q2);
    const R e2 = R(0.3) * (q0 * q1);
    const R p0 = R(2) * f0 - R(7) * f1 + R(11) * f2;  // 6 x candidate 0
    const R p1 = -f1 + R(5) * f2 + R(2) * f3;
    const R p2 = R(2) * f2 + R(5) * f3 - f4;
    return (e0 * p0 + e1 * p1 + e2 * p2) / (R(6) * (e0 + e1 + e2));
}

template <class R>
NS3D_HD R mdcd_linear(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5) {
    return c.mdcd_lin[0] * f0 + c.mdcd_lin[1] * f1 + c.mdcd_lin[2] * f2 + c.mdcd_lin[3] * f3 +
           c.mdcd_lin[4] * f4 + c.mdcd_lin[5] * f5;
}

// MDCD hybrid (mdcd_hybrid_interpolate, interpolation.cpp:135-209): the
// smoothness sensor picks the linear MDCD scheme when sai > 0.4, otherwise a
// four-stencil WENO-Z-type combination with the 6-point indicator beta3.
template <class R>
NS3D_HD R mdcd_hybrid(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5) {
    const R eps = R(0.9 * 0.4 / (1.0 - 0.9 * 0.4) * 1e-4);
    const R a1 = fabs(f2 - f1) + fabs(f2 - R(2) * f1 + f0);
    const R b1 = fabs(f2 - f3) + fabs(f2 - R(2) * f3 + f4);
    const R a2 = fabs(f3 - f2) + fabs(f3 - R(2) * f2 + f1);
    const R b2 = fabs(f3 - f4) + fabs(f3 - R(2) * f4 + f5);
    const R sai1 = (R(2) * a1 * b1 + eps) / (a1 * a1 + b1 * b1 + eps);
    const R sai2 = (R(2) * a2 * b2 + eps) / (a2 * a2 + b2 * b2 + eps);
    // std::min(sai1, sai2) > 0.4
    const R sai = sai2 < sai1 ? sai2 : sai1;
    if (sai > R(0.4)) return mdcd_linear(c, f0, f1, f2, f3, f4, f5);
    const R eps_small = R(1e-40);
    const R q0 = (R(3) * f0 - R(10) * f1 + R(15) * f2) * R(0.125);
    const R q1 = (-f1 + R(6) * f2 + R(3) * f3) * R(0.125);
    const R q2 = (R(3) * f2 + R(6) * f3 - f4) * R(0.125);
    const R q3 = (R(15) * f3 - R(10) * f4 + R(3) * f5) * R(0.125);
    const R t0 = f0 - R(2) * f1 + f2, s0 = f0 - R(4) * f1 + R(3) * f2;
    const R t1 = f1 - R(2) * f2 + f3, s1 = f1 - f3;
    const R t2 = f2 - R(2) * f3 + f4, s2 = R(3) * f2 - R(4) * f3 + f4;
    const R beta0 = R(13.0 / 12.0) * (t0 * t0) + R(0.25) * (s0 * s0);
    const R beta1 = R(13.0 / 12.0) * (t1 * t1) + R(0.25) * (s1 * s1);
    const R beta2 = R(13.0 / 12.0) * (t2 * t2) + R(0.25) * (s2 * s2);
    const R beta3 = (R(271799) * (f0 * f0) +
                     f0 * (R(-2380800) * f1 + R(4086352) * f2 - R(3462252) * f3 + R(1458762) * f4 - R(245620) * f5) +
                     f1 * (R(5653317) * f1 - R(20427884) * f2 + R(17905032) * f3 - R(7727988) * f4 + R(1325006) * f5) +
                     f2 * (R(19510972) * f2 - R(35817664) * f3 + R(15929912) * f4 - R(2792660) * f5) +
                     f3 * (R(17195652) * f3 - R(15880404) * f4 + R(2863984) * f5) +
                     f4 * (R(3824847) * f4 - R(1429976) * f5) + R(139633) * (f5 * f5)) *
                    R(1.0 / 120960.0);
    const R tau6 = fabs(beta3 - (beta0 + R(4) * beta1 + beta2) * R(1.0 / 6.0));
    R z0 = R(20) + tau6 / (beta0 + eps_small);
    R z1 = R(20) + tau6 / (beta1 + eps_small);
    R z2 = R(20) + tau6 / (beta2 + eps_small);
    R z3 = R(20) + tau6 / (beta3 + eps_small);
    const R w0 = c.mdcd_d[0] * (z0 * z0), w1 = c.mdcd_d[1] * (z1 * z1);
    const R w2 = c.mdcd_d[2] * (z2 * z2), w3 = c.mdcd_d[3] * (z3 * z3);
    return (w0 * q0 + w1 * q1 + w2 * q2 + w3 * q3) / (w0 + w1 + w2 + w3);
}

template <int INTERP, class R>
NS3D_HD R interp6(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5) {
    if constexpr (INTERP == I_WENO5) return wcns5(f0, f1, f2, f3, f4);
    else if constexpr (INTERP == I_WENO5_JS) return weno5js(f0, f1, f2, f3, f4);
    else if constexpr (INTERP == I_MDCD_LINEAR) return mdcd_linear(c, f0, f1, f2, f3, f4, f5);
    else return mdcd_hybrid(c, f0, f1, f2, f3, f4, f5);
}

// Left (+1) and right (-1) states of one 6-point line s[0..5] = cells i-2..i+3
// (interpolate_select, interpolation.cpp:13-77: the right state reverses the
// stencil).
template <int INTERP, class R>
NS3D_HD void interp_lr(const SchemeCoef<R> &c, const R *s, R &left, R &right) {
    left = interp6<INTERP>(c, s[0], s[1], s[2], s[3], s[4], s[5]);
    right = interp6<INTERP>(c, s[5], s[4], s[3], s[2], s[1], s[0]);
}

// ------------------------------------------------------------ Riemann solvers
// Primitive unpacking of a rotated conserved state.
template <class R>
struct Prim {
    R rho, un, u1, u2, E, p, a, ir;
};

template <class R>
NS3D_HD Prim<R> unpack(const R *U, R g) {
    Prim<R> s;
    s.rho = U[0];
    s.ir = R(1) / s.rho;
    s.un = U[1] * s.ir;
    s.u1 = U[2] * s.ir;
    s.u2 = U[3] * s.ir;
    s.E = U[4];
    s.p = (U[4] - R(0.5) * s.rho * (s.un * s.un + s.u1 * s.u1 + s.u2 * s.u2)) * (g - R(1));
    s.a = sqrt(g * s.p * s.ir);
    return s;
}

// Physical x-normal flux of a rotated state.
template <class R>
NS3D_HD void phys_flux(const Prim<R> &s, R *F) {
    const R m = s.rho * s.un;
    F[0] = m;
    F[1] = m * s.un + s.p;
    F[2] = m * s.u1;
    F[3] = m * s.u2;
    F[4] = (s.E + s.p) * s.un;
}

// Rusanov (InviscousFlux.cpp:270-311)
template <class R>
NS3D_HD void rusanov(const R *UL, const R *UR, R g, R *F) {
    const Prim<R> L = unpack(UL, g), Rr = unpack(UR, g);
    R FL[5], FR[5];
    phys_flux(L, FL);
    phys_flux(Rr, FR);
    const R sl = fabs(L.un) + L.a, sr = fabs(Rr.un) + Rr.a;
    const R smax = sl > sr ? sl : sr;
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = R(0.5) * (FL[n] + FR[n]) - R(0.5) * smax * (UR[n] - UL[n]);
}

// HLL (InviscousFlux.cpp:384-447), Davis wave speeds.
template <class R>
NS3D_HD void hll(const R *UL, const R *UR, R g, R *F) {
    const Prim<R> L = unpack(UL, g), Rr = unpack(UR, g);
    R FL[5], FR[5];
    phys_flux(L, FL);
    phys_flux(Rr, FR);
    const R SL = fmin(L.un - L.a, Rr.un - Rr.a);
    const R SR = fmax(L.un + L.a, Rr.un + Rr.a);
    if (SL >= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FL[n];
    } else if (SR <= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FR[n];
    } else {
        const R id = R(1) / (SR - SL);
        const R f1 = SR * id, f2 = SL * id, f3 = SL * SR * id;
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = f1 * FL[n] - f2 * FR[n] + f3 * (UR[n] - UL[n]);
    }
}

// HLLC star-region flux, shared by HLLC and HLLC_p (InviscousFlux.cpp:538-591).
// Tangential velocities are unchanged across the contact; the star energy is
// E/rho + (SM - un)(SM + p / (rho (S - un))).
template <class R>
NS3D_HD void hllc_star(const Prim<R> &L, const Prim<R> &Rr, const R *UL, const R *UR,
                                          const R *FL, const R *FR, R SL, R SR, R SM, R *F) {
    const bool left = SM >= R(0);
    const Prim<R> &s = left ? L : Rr;
    const R S = left ? SL : SR;
    const R *U = left ? UL : UR;
    const R *Fs = left ? FL : FR;
    const R dS = S - s.un;
    const R rs = s.rho * dS / (S - SM);
    const R Es = s.E * s.ir + (SM - s.un) * (SM + s.p / (s.rho * dS));
    const R Us[5] = {rs, rs * SM, rs * s.u1, rs * s.u2, rs * Es};
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = Fs[n] + S * (Us[n] - U[n]);
}

// HLLC (InviscousFlux.cpp:449-592): Davis SL/SR, supersonic early-outs,
// pressure-based contact speed with the Roe-averaged normal velocity as the
// fallback when |den| <= 1e-12.
template <class R>
NS3D_HD void hllc(const R *UL, const R *UR, R g, R *F) {
    const Prim<R> L = unpack(UL, g), Rr = unpack(UR, g);
    R FL[5], FR[5];
    phys_flux(L, FL);
    phys_flux(Rr, FR);
    const R SL = fmin(L.un - L.a, Rr.un - Rr.a);
    const R SR = fmax(L.un + L.a, Rr.un + Rr.a);
    if (SL >= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FL[n];
        return;
    }
    if (SR <= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FR[n];
        return;
    }
    const R num = Rr.rho * Rr.un * (SR - Rr.un) - L.rho * L.un * (SL - L.un) + L.p - Rr.p;
    const R den = Rr.rho * (SR - Rr.un) - L.rho * (SL - L.un);
    R SM;
    if (fabs(den) > R(1e-12)) {
        SM = num / den;
    } else {
        const R sqL = sqrt(L.rho), sqR = sqrt(Rr.rho);
        SM = (sqL * L.un + sqR * Rr.un) / (sqL + sqR);
    }
    hllc_star(L, Rr, UL, UR, FL, FR, SL, SR, SM, F);
}

// HLLC with pressure-based wave speeds (HLLC_p_Riemann_solver,
// InviscousFlux.cpp:594-717).
template <class R>
NS3D_HD void hllc_p(const R *UL, const R *UR, R g, R *F) {
    const Prim<R> L = unpack(UL, g), Rr = unpack(UR, g);
    R FL[5], FR[5];
    phys_flux(L, FL);
    phys_flux(Rr, FR);
    R ps = R(0.5) * (L.p + Rr.p) - R(0.125) * (Rr.un - L.un) * (L.rho + Rr.rho) * (L.a + Rr.a);
    ps = fmax(R(0), ps);
    const R c1 = (g + R(1)) / (R(2) * g), c2 = (g - R(1)) / (R(2) * g);
    const R SL = ps <= L.p ? L.un - L.a : L.un - L.a * sqrt(c1 * ps / L.p + c2);
    const R SR = ps <= Rr.p ? Rr.un + Rr.a : Rr.un + Rr.a * sqrt(c1 * ps / Rr.p + c2);
    if (SL >= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FL[n];
        return;
    }
    if (SR <= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FR[n];
        return;
    }
    const R SM = (Rr.p - L.p + L.rho * L.un * (SL - L.un) - Rr.rho * Rr.un * (SR - Rr.un)) /
                 (L.rho * (SL - L.un) - Rr.rho * (SR - Rr.un));
    hllc_star(L, Rr, UL, UR, FL, FR, SL, SR, SM, F);
}

// Characteristic matrices for the x normal at a Roe-type state (rho, u, v, w,
// H, c) (build_eigen_matrices with n = e_x, InviscousFlux.cpp:888-987).  Rows
// of L: acoustic -, entropy, tangential u1, tangential u2, acoustic +.
template <class R>
struct Eig {
    R L[5][5];
    R Rm[5][5];
};

template <class R>
NS3D_HD void eigen_x(R rho, R u, R v, R w, R c, R g, Eig<R> &e) {
    const R a1 = g - R(1);
    const R phi = R(0.5) * a1 * (u * u + v * v + w * w);
    const R ic = R(1) / c;
    const R ic2 = ic * ic;
    const R irho = R(1) / rho;
    const R a2 = R(0.70710678118654752440) * irho * ic;  // 1 / (sqrt(2) rho c)
    const R a3 = R(0.70710678118654752440) * rho * ic;   // rho / (sqrt(2) c)
    const R a4 = (phi + c * c) / a1;
    const R a5 = R(1) - phi * ic2;
    const R a6 = phi / a1;
    e.L[0][0] = a2 * (phi + c * u);
    e.L[0][1] = -a2 * (a1 * u + c);
    e.L[0][2] = -a2 * (a1 * v);
    e.L[0][3] = -a2 * (a1 * w);
    e.L[0][4] = a1 * a2;
    e.L[1][0] = a5;
    e.L[1][1] = a1 * u * ic2;
    e.L[1][2] = a1 * v * ic2;
    e.L[1][3] = a1 * w * ic2;
    e.L[1][4] = -a1 * ic2;
    e.L[2][0] = v * irho;
    e.L[2][1] = R(0);
    e.L[2][2] = -irho;
    e.L[2][3] = R(0);
    e.L[2][4] = R(0);
    e.L[3][0] = -w * irho;
    e.L[3][1] = R(0);
    e.L[3][2] = R(0);
    e.L[3][3] = irho;
    e.L[3][4] = R(0);
    e.L[4][0] = a2 * (phi - c * u);
    e.L[4][1] = -a2 * (a1 * u - c);
    e.L[4][2] = -a2 * (a1 * v);
    e.L[4][3] = -a2 * (a1 * w);
    e.L[4][4] = a1 * a2;
    // columns of R
    e.Rm[0][0] = a3;            e.Rm[0][1] = R(1);  e.Rm[0][2] = R(0);     e.Rm[0][3] = R(0);    e.Rm[0][4] = a3;
    e.Rm[1][0] = a3 * (u - c);  e.Rm[1][1] = u;     e.Rm[1][2] = R(0);     e.Rm[1][3] = R(0);    e.Rm[1][4] = a3 * (u + c);
    e.Rm[2][0] = a3 * v;        e.Rm[2][1] = v;     e.Rm[2][2] = -rho;     e.Rm[2][3] = R(0);    e.Rm[2][4] = a3 * v;
    e.Rm[3][0] = a3 * w;        e.Rm[3][1] = w;     e.Rm[3][2] = R(0);     e.Rm[3][3] = rho;     e.Rm[3][4] = a3 * w;
    e.Rm[4][0] = a3 * (a4 - c * u); e.Rm[4][1] = a6; e.Rm[4][2] = -rho * v; e.Rm[4][3] = rho * w; e.Rm[4][4] = a3 * (a4 + c * u);
}

// L * x for the sparse x-normal rows.
template <class R>
NS3D_HD void proj_L(const Eig<R> &e, const R *x, R *y) {
    y[0] = e.L[0][0] * x[0] + e.L[0][1] * x[1] + e.L[0][2] * x[2] + e.L[0][3] * x[3] + e.L[0][4] * x[4];
    y[1] = e.L[1][0] * x[0] + e.L[1][1] * x[1] + e.L[1][2] * x[2] + e.L[1][3] * x[3] + e.L[1][4] * x[4];
    y[2] = e.L[2][0] * x[0] + e.L[2][2] * x[2];
    y[3] = e.L[3][0] * x[0] + e.L[3][3] * x[3];
    y[4] = e.L[4][0] * x[0] + e.L[4][1] * x[1] + e.L[4][2] * x[2] + e.L[4][3] * x[3] + e.L[4][4] * x[4];
}

// R * q for the sparse x-normal columns.
template <class R>
NS3D_HD void proj_R(const Eig<R> &e, const R *q, R *y) {
    y[0] = e.Rm[0][0] * q[0] + q[1] + e.Rm[0][4] * q[4];
    y[1] = e.Rm[1][0] * q[0] + e.Rm[1][1] * q[1] + e.Rm[1][4] * q[4];
    y[2] = e.Rm[2][0] * q[0] + e.Rm[2][1] * q[1] + e.Rm[2][2] * q[2] + e.Rm[2][4] * q[4];
    y[3] = e.Rm[3][0] * q[0] + e.Rm[3][1] * q[1] + e.Rm[3][3] * q[3] + e.Rm[3][4] * q[4];
    y[4] = e.Rm[4][0] * q[0] + e.Rm[4][1] * q[1] + e.Rm[4][2] * q[2] + e.Rm[4][3] * q[3] + e.Rm[4][4] * q[4];
}

// Roe flux difference splitting with the entropy fix delta = 0.1 max|lambda|
// (Roe_Riemann_solver, InviscousFlux.cpp:313-382; Roe average 849-884).
template <class R>
NS3D_HD void roe(const R *UL, const R *UR, R g, R *F) {
    const Prim<R> L = unpack(UL, g), Rr = unpack(UR, g);
    R FL[5], FR[5];
    phys_flux(L, FL);
    phys_flux(Rr, FR);
    const R sL = sqrt(L.rho), sR = sqrt(Rr.rho);
    const R rb = sL * sR;
    const R dn = R(1) / (sL + sR);
    const R ub = (sL * L.un + sR * Rr.un) * dn;
    const R vb = (sL * L.u1 + sR * Rr.u1) * dn;
    const R wb = (sL * L.u2 + sR * Rr.u2) * dn;
    const R HL = (L.E + L.p) * L.ir, HR = (Rr.E + Rr.p) * Rr.ir;
    const R hb = (sL * HL + sR * HR) * dn;
    const R cb = sqrt((g - R(1)) * (hb - R(0.5) * (ub * ub + vb * vb + wb * wb)));
    Eig<R> e;
    eigen_x(rb, ub, vb, wb, cb, g, e);
    R dU[5], al[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) dU[n] = UR[n] - UL[n];
    proj_L(e, dU, al);
    R lam[5] = {ub - cb, ub, ub, ub, ub + cb};
    R mx = R(0);
#pragma unroll
    for (int m = 0; m < 5; ++m) mx = fmax(mx, fabs(lam[m]));
    const R delta = R(0.1) * mx;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        R al_m = fabs(lam[m]);
        if (al_m < delta) al_m = fabs((lam[m] * lam[m] + delta * delta) / (R(2) * delta));
        al[m] *= al_m;
    }
    R D[5];
    proj_R(e, al, D);
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = R(0.5) * (FL[n] + FR[n]) - R(0.5) * D[n];
}

// AUSM-type flux (AUSM_Riemann_solver, InviscousFlux.cpp:719-846).  XS is the
// rotated slot of the x velocity: the reference's pressure-diffusion term uses
// (u_R - u_L) of the x velocity for every face direction (line 823).
template <int XS, class R>
NS3D_HD void ausm(const R *UL, const R *UR, R g, R *F) {
    const R Kp = R(0.25), Ku = R(0.75);
    const R rL = UL[0], irL = R(1) / rL;
    const R vL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
    R pL = (UL[4] - R(0.5) * rL * (vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2])) * (g - R(1));
    pL = fmax(pL, R(1e-12));
    const R aL = sqrt(g * pL * irL);
    const R HL = (UL[4] + pL) * irL;
    const R rR = UR[0], irR = R(1) / rR;
    const R vR[3] = {UR[1] * irR, UR[2] * irR, UR[3] * irR};
    R pR = (UR[4] - R(0.5) * rR * (vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2])) * (g - R(1));
    pR = fmax(pR, R(1e-12));
    const R aR = sqrt(g * pR * irR);
    const R HR = (UR[4] + pR) * irR;
    const R ab = R(0.5) * (aL + aR);
    const R rhob = R(0.5) * (rL + rR);
    const R iab = R(1) / ab;
    const R ML = vL[0] * iab, MR = vR[0] * iab;
    const R Mb2 = (vL[0] * vL[0] + vR[0] * vR[0]) / (R(2) * ab * ab);
    const R Mo2 = fmin(R(1), fmax(Mb2, R(0.5)));
    const R fa = Mo2 * (R(2) - Mo2);
    const R Mp = fabs(ML) >= R(1) ? R(0.5) * (ML + fabs(ML)) : R(0.25) * (ML + R(1)) * (ML + R(1));
    const R Mm = fabs(MR) >= R(1) ? R(0.5) * (MR - fabs(MR)) : R(-0.25) * (MR - R(1)) * (MR - R(1));
    const R Mh = Mp + Mm - Kp / fa * fmax(R(1) - Mb2, R(0)) * ((pR - pL) / (rhob * ab * ab));
    const R mdot = ab * Mh * (Mh >= R(0) ? rL : rR);
    const R Pp = fabs(ML) >= R(1) ? R(0.5) * (R(1) + (ML > R(0) ? R(1) : R(-1)))
                                  : R(0.25) * (ML + R(1)) * (ML + R(1)) * (R(2) - ML);
    const R Pm = fabs(MR) >= R(1) ? R(0.5) * (R(1) - (MR > R(0) ? R(1) : R(-1)))
                                  : R(0.25) * (MR - R(1)) * (MR - R(1)) * (R(2) + MR);
    const R dux = vR[XS - 1] - vL[XS - 1];
    const R ph = Pp * pL + Pm * pR - Ku * Pm * Pp * (rL + rR) * fa * dux * ab;
    const bool up = mdot >= R(0);
    F[0] = mdot;
    F[1] = mdot * (up ? vL[0] : vR[0]) + ph;
    F[2] = mdot * (up ? vL[1] : vR[1]);
    F[3] = mdot * (up ? vL[2] : vR[2]);
    F[4] = mdot * (up ? HL : HR);
}

template <int RS, int XS, class R>
NS3D_HD void riemann(const R *UL, const R *UR, R g, R *F) {
    if constexpr (RS == R_ROE) roe(UL, UR, g, F);
    else if constexpr (RS == R_RUSANOV) rusanov(UL, UR, g, F);
    else if constexpr (RS == R_HLLC) hllc(UL, UR, g, F);
    else if constexpr (RS == R_HLL) hll(UL, UR, g, F);
    else if constexpr (RS == R_HLLC_P) hllc_p(UL, UR, g, F);
    else ausm<XS>(UL, UR, g, F);
}

// One WCNS face flux from the rotated 6-point conserved stencil st[var][m]
// (cells i-2..i+3) and, for characteristic-wise interpolation, the rotated
// primitives (rho, un, u1, u2, p) of cells i and i+1 (WCNS_Riemann_InviscidFlux,
// InviscousFlux.cpp:146-268).
template <int INTERP, int RS, bool CHAR, int XS, class R>
NS3D_HD void face_flux(const SchemeCoef<R> &c, const R (&st)[5][6], const R (&pc)[2][5],
                                          R *F) {
    R UL[5], UR[5];
    if constexpr (!CHAR) {
#pragma unroll
        for (int n = 0; n < 5; ++n) interp_lr<INTERP>(c, st[n], UL[n], UR[n]);
    } else {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            // eigenvectors of the single cell state i (left) or i+1 (right)
            const R rho = pc[side][0], u = pc[side][1], v = pc[side][2], w = pc[side][3], p = pc[side][4];
            const R cs = sqrt(c.gamma * p / rho);
            Eig<R> e;
            eigen_x(rho, u, v, w, cs, c.gamma, e);
            R W[5][6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                R x[5] = {st[0][m], st[1][m], st[2][m], st[3][m], st[4][m]};
                R y[5];
                proj_L(e, x, y);
#pragma unroll
                for (int n = 0; n < 5; ++n) W[n][m] = y[n];
            }
            R Q[5];
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                if (side == 0) Q[n] = interp6<INTERP>(c, W[n][0], W[n][1], W[n][2], W[n][3], W[n][4], W[n][5]);
                else Q[n] = interp6<INTERP>(c, W[n][5], W[n][4], W[n][3], W[n][2], W[n][1], W[n][0]);
            }
            proj_R(e, Q, side == 0 ? UL : UR);
        }
    }
    riemann<RS, XS>(UL, UR, c.gamma, F);
}

}  // namespace ns3d
