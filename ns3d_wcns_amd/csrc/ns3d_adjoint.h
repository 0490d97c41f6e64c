// ns3d_adjoint.h -- hand-written reverse mode (vector-Jacobian products) of
// the FP64 face flux and the pointwise viscous terms, for the discrete
// adjoint of the RHS.  Each function recomputes its forward values, then
// propagates the output adjoint back to the inputs, accumulating (+=).
//
// The nonlinear WCNS5 / WENO5-JS weights are differentiated exactly.  The
// non-smooth parts take the fixed subgradient of the tangent-linear model
// (ns3d_dual.h): min / max follow the operand they return (the first on
// ties), d|x| = sign(x) with sign(0) = +1, and every branch condition (the
// HLLC wave pattern) is frozen at the forward values.
//
// Supported: component-wise WCNS5, WENO5-JS and MDCD (linear, hybrid)
// interpolation with every Riemann solver: HLLC and Rusanov by hand, Roe,
// HLL, HLLC_p and AUSM by transposing their forward-mode tangent
// (riemann_vjp_dual).
#pragma once
#include "ns3d_common.h"
#include "ns3d_dual.h"
#include "ns3d_visc.h"

namespace ns3d {

// ------------------------------------------------------------ interpolation
// Reverse mode of wcns5 (JS = false) / weno5js (JS = true) on f[0..4]: fb +=
// yb dy/df; returns y.
template <bool JS>
NS3D_HD double weno_vjp(const double *f, double yb, double *fb) {
    const double eps = 1e-6;
    const double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    const double t[3] = {f0 - 2.0 * f1 + f2, f1 - 2.0 * f2 + f3, f2 - 2.0 * f3 + f4};
    const double s[3] = {f0 - 4.0 * f1 + 3.0 * f2, f1 - f3, 3.0 * f2 - 4.0 * f3 + f4};
    double b[3], q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b[k] = (13.0 / 12.0) * t[k] * t[k] + 0.25 * s[k] * s[k];
        q[k] = (eps + b[k]) * (eps + b[k]);
    }
    const double d0 = JS ? 0.1 : 1.0 / 16.0, d1 = JS ? 0.6 : 5.0 / 8.0, d2 = JS ? 0.3 : 5.0 / 16.0;
    const double K = JS ? 6.0 : 8.0;
    const double e[3] = {d0 * q[1] * q[2], d1 * q[0] * q[2], d2 * q[0] * q[1]};
    // K x candidate values
    const double pc[3][3] = {{JS ? 2.0 : 3.0, JS ? -7.0 : -10.0, JS ? 11.0 : 15.0},
                             {-1.0, JS ? 5.0 : 6.0, JS ? 2.0 : 3.0},
                             {JS ? 2.0 : 3.0, JS ? 5.0 : 6.0, -1.0}};
    const double p[3] = {pc[0][0] * f0 + pc[0][1] * f1 + pc[0][2] * f2, pc[1][0] * f1 + pc[1][1] * f2 + pc[1][2] * f3,
                         pc[2][0] * f2 + pc[2][1] * f3 + pc[2][2] * f4};
    const double N = e[0] * p[0] + e[1] * p[1] + e[2] * p[2];
    const double D = K * (e[0] + e[1] + e[2]);
    // y = N / D: Nb = yb / D and Db = -Nb y from one reciprocal
    const double iD = frcp(D);
    const double y = N * iD;
    const double Nb = yb * iD, Db = -Nb * y;
    double eb[3], pb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        eb[k] = Nb * p[k] + K * Db;
        pb[k] = Nb * e[k];
    }
    const double qb[3] = {eb[1] * d1 * q[2] + eb[2] * d2 * q[1], eb[0] * d0 * q[2] + eb[2] * d2 * q[0],
                          eb[0] * d0 * q[1] + eb[1] * d1 * q[0]};
    double tb[3], sb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double bb = qb[k] * 2.0 * (eps + b[k]);
        tb[k] = bb * (13.0 / 6.0) * t[k];
        sb[k] = bb * 0.5 * s[k];
    }
    // t, s and p are linear in f
    fb[0] += tb[0] + sb[0] + pb[0] * pc[0][0];
    fb[1] += -2.0 * tb[0] + tb[1] - 4.0 * sb[0] + sb[1] + pb[0] * pc[0][1] + pb[1] * pc[1][0];
    fb[2] += tb[0] - 2.0 * tb[1] + tb[2] + 3.0 * sb[0] + 3.0 * sb[2] + pb[0] * pc[0][2] + pb[1] * pc[1][1] +
             pb[2] * pc[2][0];
    fb[3] += tb[1] - 2.0 * tb[2] - sb[1] - 4.0 * sb[2] + pb[1] * pc[1][2] + pb[2] * pc[2][1];
    fb[4] += tb[2] + sb[2] + pb[2] * pc[2][2];
    return y;
}

// Both reconstructions of one cell (recon_pair, ns3d_device.h) in FP64.
template <bool JS>
NS3D_HD void weno_pair(const double *f, double &yL, double &yR) {
    recon_pair<JS ? I_WENO5_JS : I_WENO5>(f, yL, yR);
}

// Reverse mode of both reconstructions of one cell from its 5-cell stencil
// f[0..4] = cells c-2..c+2: yL = weno(f0..f4) (the value at the right face,
// the left state of face c+1/2) and yR = weno(f4..f0) (the value at the left
// face, the right state of face c-1/2).  The two share their smoothness
// indicators (the mirrored stencil swaps beta_0 and beta_2), which are
// formed once.  fb += ybL dyL/df + ybR dyR/df.
template <bool JS>
NS3D_HD void weno_pair_vjp(const double *f, double ybL, double ybR, double *fb) {
    const double eps = 1e-6;
    const double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    const double t[3] = {f0 - 2.0 * f1 + f2, f1 - 2.0 * f2 + f3, f2 - 2.0 * f3 + f4};
    const double s[3] = {f0 - 4.0 * f1 + 3.0 * f2, f1 - f3, 3.0 * f2 - 4.0 * f3 + f4};
    double b[3], q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b[k] = (13.0 / 12.0) * t[k] * t[k] + 0.25 * s[k] * s[k];
        q[k] = (eps + b[k]) * (eps + b[k]);
    }
    const double d0 = JS ? 0.1 : 1.0 / 16.0, d1 = JS ? 0.6 : 5.0 / 8.0, d2 = JS ? 0.3 : 5.0 / 16.0;
    const double K = JS ? 6.0 : 8.0;
    const double c0 = JS ? 2.0 : 3.0, c1 = JS ? -7.0 : -10.0, c2 = JS ? 11.0 : 15.0;  // K x candidate 0
    const double m1 = JS ? 5.0 : 6.0, m2 = JS ? 2.0 : 3.0;                            // candidate 1: -1, m1, m2
    // left (right-face) reconstruction: weights on (q0, q1, q2), stencils f0..f4
    const double eL[3] = {d0 * q[1] * q[2], d1 * q[0] * q[2], d2 * q[0] * q[1]};
    const double pL[3] = {c0 * f0 + c1 * f1 + c2 * f2, -f1 + m1 * f2 + m2 * f3, m2 * f2 + m1 * f3 - f4};
    // mirrored: beta'_0 = beta_2, beta'_2 = beta_0
    const double eR[3] = {d0 * q[1] * q[0], d1 * q[2] * q[0], d2 * q[2] * q[1]};
    const double pR[3] = {c0 * f4 + c1 * f3 + c2 * f2, -f3 + m1 * f2 + m2 * f1, m2 * f2 + m1 * f1 - f0};
    // y = N / D, so Nb = yb / D and Db = -yb y / D = -Nb y: one reciprocal
    // per side serves the value and both adjoints
    const double DL = K * (eL[0] + eL[1] + eL[2]), DR = K * (eR[0] + eR[1] + eR[2]);
    const double iDL = frcp(DL), iDR = frcp(DR);
    const double yL = (eL[0] * pL[0] + eL[1] * pL[1] + eL[2] * pL[2]) * iDL;
    const double yR = (eR[0] * pR[0] + eR[1] * pR[1] + eR[2] * pR[2]) * iDR;
    const double NbL = ybL * iDL, DbL = -NbL * yL;
    const double NbR = ybR * iDR, DbR = -NbR * yR;
    double ebL[3], ebR[3], pbL[3], pbR[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ebL[k] = NbL * pL[k] + K * DbL;
        pbL[k] = NbL * eL[k];
        ebR[k] = NbR * pR[k] + K * DbR;
        pbR[k] = NbR * eR[k];
    }
    // q adjoints: left e = (d0 q1 q2, d1 q0 q2, d2 q0 q1); mirrored e = (d0 q1 q0, d1 q2 q0, d2 q2 q1)
    const double qb0 = ebL[1] * d1 * q[2] + ebL[2] * d2 * q[1] + ebR[0] * d0 * q[1] + ebR[1] * d1 * q[2];
    const double qb1 = ebL[0] * d0 * q[2] + ebL[2] * d2 * q[0] + ebR[0] * d0 * q[0] + ebR[2] * d2 * q[2];
    const double qb2 = ebL[0] * d0 * q[1] + ebL[1] * d1 * q[0] + ebR[1] * d1 * q[0] + ebR[2] * d2 * q[1];
    const double qb[3] = {qb0, qb1, qb2};
    double tb[3], sb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double bb = qb[k] * 2.0 * (eps + b[k]);
        tb[k] = bb * (13.0 / 6.0) * t[k];
        sb[k] = bb * 0.5 * s[k];
    }
    fb[0] += tb[0] + sb[0] + pbL[0] * c0 - pbR[2];
    fb[1] += -2.0 * tb[0] + tb[1] - 4.0 * sb[0] + sb[1] + pbL[0] * c1 - pbL[1] + pbR[1] * m2 + pbR[2] * m1;
    fb[2] += tb[0] - 2.0 * tb[1] + tb[2] + 3.0 * sb[0] + 3.0 * sb[2] + (pbL[0] + pbR[0]) * c2 +
             (pbL[1] + pbR[1]) * m1 + (pbL[2] + pbR[2]) * m2;
    fb[3] += tb[1] - 2.0 * tb[2] - sb[1] - 4.0 * sb[2] + pbL[1] * m2 + pbL[2] * m1 + pbR[0] * c1 - pbR[1];
    fb[4] += tb[2] + sb[2] - pbL[2] + pbR[0] * c0;
}

// ------------------------------------------------------------ Riemann flux
// Adjoints of the Prim fields of one side.
struct PrimBar {
    double rho, un, u1, u2, E, p, a, ir;
};

// The symmetric matrix of beta3_of's quadratic form times 120960 (the
// diagonal doubled): d beta3 / d f_i = sum_j kB3[i][j] f_j / 120960.
__device__ constexpr double kB3[6][6] = {
    {2.0 * 271799, -2380800, 4086352, -3462252, 1458762, -245620},
    {-2380800, 2.0 * 5653317, -20427884, 17905032, -7727988, 1325006},
    {4086352, -20427884, 2.0 * 19510972, -35817664, 15929912, -2792660},
    {-3462252, 17905032, -35817664, 2.0 * 17195652, -15880404, 2863984},
    {1458762, -7727988, 15929912, -15880404, 2.0 * 3824847, -1429976},
    {-245620, 1325006, -2792660, 2863984, -1429976, 2.0 * 139633}};

// Reverse mode of the FP64 MDCD interpolations (interp6 of MDCD_LINEAR /
// MDCD_HYBRID, off = 0) on f[0..5]: fb += yb dy/df; returns y.  The
// hybrid's sensor branch is frozen at the forward values, and |.| takes
// sign(0) = +1, as in the tangent.
template <int INTERP>
NS3D_HD double mdcd_vjp(const SchemeCoef<double> &c, const double *f, double yb, double *fb) {
    static_assert(INTERP == I_MDCD_LINEAR || INTERP == I_MDCD_HYBRID, "MDCD interpolation");
    const double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4], f5 = f[5];
    if constexpr (INTERP == I_MDCD_HYBRID) {
        // the forward's sensor and branch (ns3d_device.h mdcd_hybrid)
        const double eps = 0.9 * 0.4 / (1.0 - 0.9 * 0.4) * 1e-4;
        const double a1 = ::fabs(f2 - f1) + ::fabs(f2 - 2.0 * f1 + f0);
        const double b1 = ::fabs(f2 - f3) + ::fabs(f2 - 2.0 * f3 + f4);
        const double a2 = ::fabs(f3 - f2) + ::fabs(f3 - 2.0 * f2 + f1);
        const double b2 = ::fabs(f3 - f4) + ::fabs(f3 - 2.0 * f4 + f5);
        const double sai1 = fdiv(2.0 * a1 * b1 + eps, a1 * a1 + b1 * b1 + eps);
        const double sai2 = fdiv(2.0 * a2 * b2 + eps, a2 * a2 + b2 * b2 + eps);
        const double sai = sai2 < sai1 ? sai2 : sai1;
        if (!(sai > 0.4)) {
            const double eps_small = 1e-40;
            const double q[4] = {(3.0 * f0 - 10.0 * f1 + 15.0 * f2) * 0.125, (-f1 + 6.0 * f2 + 3.0 * f3) * 0.125,
                                 (3.0 * f2 + 6.0 * f3 - f4) * 0.125, (15.0 * f3 - 10.0 * f4 + 3.0 * f5) * 0.125};
            const double t0 = f0 - 2.0 * f1 + f2, s0 = f0 - 4.0 * f1 + 3.0 * f2;
            const double t1 = f1 - 2.0 * f2 + f3, s1 = f1 - f3;
            const double t2 = f2 - 2.0 * f3 + f4, s2 = 3.0 * f2 - 4.0 * f3 + f4;
            double beta[4];
            beta[0] = (13.0 / 12.0) * (t0 * t0) + 0.25 * (s0 * s0);
            beta[1] = (13.0 / 12.0) * (t1 * t1) + 0.25 * (s1 * s1);
            beta[2] = (13.0 / 12.0) * (t2 * t2) + 0.25 * (s2 * s2);
            beta[3] = beta3_of<double>(f0, f1, f2, f3, f4, f5);
            const double X = beta[3] - (beta[0] + 4.0 * beta[1] + beta[2]) * (1.0 / 6.0);
            const double tau6 = ::fabs(X);
            double r[4], z[4], w[4], N = 0.0, D = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                r[k] = frcp(beta[k] + eps_small);
                z[k] = 20.0 + tau6 * r[k];
                w[k] = c.mdcd_d[k] * (z[k] * z[k]);
            }
            N = w[0] * q[0] + w[1] * q[1] + w[2] * q[2] + w[3] * q[3];
            D = w[0] + w[1] + w[2] + w[3];
            const double iD = frcp(D), y = N * iD;
            // y = N / D
            const double Nb = yb * iD, Db = -yb * y * iD;
            double qb[4], betab[4], tau6b = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                qb[k] = Nb * w[k];
                const double zb = (Nb * q[k] + Db) * c.mdcd_d[k] * 2.0 * z[k];
                tau6b += zb * r[k];
                betab[k] = -zb * tau6 * r[k] * r[k];
            }
            const double Xb = X >= 0.0 ? tau6b : -tau6b;
            betab[3] += Xb;
            betab[0] -= Xb * (1.0 / 6.0);
            betab[1] -= Xb * (4.0 / 6.0);
            betab[2] -= Xb * (1.0 / 6.0);
            const double tb0 = betab[0] * (13.0 / 6.0) * t0, sb0 = betab[0] * 0.5 * s0;
            const double tb1 = betab[1] * (13.0 / 6.0) * t1, sb1 = betab[1] * 0.5 * s1;
            const double tb2 = betab[2] * (13.0 / 6.0) * t2, sb2 = betab[2] * 0.5 * s2;
            fb[0] += tb0 + sb0 + 0.375 * qb[0];
            fb[1] += -2.0 * tb0 - 4.0 * sb0 + tb1 + sb1 - 1.25 * qb[0] - 0.125 * qb[1];
            fb[2] += tb0 + 3.0 * sb0 - 2.0 * tb1 + tb2 + 3.0 * sb2 + 1.875 * qb[0] + 0.75 * qb[1] + 0.375 * qb[2];
            fb[3] += tb1 - sb1 - 2.0 * tb2 - 4.0 * sb2 + 0.375 * qb[1] + 0.75 * qb[2] + 1.875 * qb[3];
            fb[4] += tb2 + sb2 - 0.125 * qb[2] - 1.25 * qb[3];
            fb[5] += 0.375 * qb[3];
            const double b3 = betab[3] * (1.0 / 120960.0);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double g = 0.0;
#pragma unroll
                for (int jj = 0; jj < 6; ++jj) g += kB3[i][jj] * f[jj];
                fb[i] += b3 * g;
            }
            return y;
        }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) fb[m] += c.mdcd_lin[m] * yb;
    return mdcd_linear(c, f0, f1, f2, f3, f4, f5);
}

// Reverse mode of phys_flux: sb += (dF/ds)^T Fb.
NS3D_HD void phys_flux_vjp(const Prim<double> &s, const double *Fb, PrimBar &sb) {
    const double m = s.rho * s.un;
    const double mb = Fb[0] + Fb[1] * s.un + Fb[2] * s.u1 + Fb[3] * s.u2;
    sb.un += Fb[1] * m + Fb[4] * (s.E + s.p) + mb * s.rho;
    sb.p += Fb[1] + Fb[4] * s.un;
    sb.u1 += Fb[2] * m;
    sb.u2 += Fb[3] * m;
    sb.E += Fb[4] * s.un;
    sb.rho += mb * s.un;
}

// Reverse mode of unpack: Ub += (ds/dU)^T sb.
NS3D_HD void unpack_vjp(double g, const double *U, const Prim<double> &s, PrimBar sb, double *Ub) {
    // a = sqrt(g p ir)
    const double i2a = frcp(2.0 * s.a);
    sb.p += sb.a * g * s.ir * i2a;
    sb.ir += sb.a * g * s.p * i2a;
    // p = (E - rho k / 2)(g - 1), k = un^2 + u1^2 + u2^2
    const double k = s.un * s.un + s.u1 * s.u1 + s.u2 * s.u2;
    sb.E += sb.p * (g - 1.0);
    sb.rho -= sb.p * (g - 1.0) * 0.5 * k;
    const double kb = -sb.p * (g - 1.0) * 0.5 * s.rho;
    sb.un += 2.0 * kb * s.un;
    sb.u1 += 2.0 * kb * s.u1;
    sb.u2 += 2.0 * kb * s.u2;
    // velocities = momenta * ir, ir = 1 / rho
    Ub[1] += sb.un * s.ir;
    Ub[2] += sb.u1 * s.ir;
    Ub[3] += sb.u2 * s.ir;
    sb.ir += sb.un * U[1] + sb.u1 * U[2] + sb.u2 * U[3];
    sb.rho -= sb.ir * s.ir * s.ir;
    Ub[0] += sb.rho;
    Ub[4] += sb.E;
}

// Reverse mode of hllc (ns3d_device.h): ULb += (dF/dUL)^T Fb, URb likewise.
NS3D_HD void hllc_vjp(const SchemeCoef<double> &c, const double *UL, const double *UR, const double *Fb, double *ULb,
                      double *URb) {
    const Prim<double> L = unpack(c, UL), Rr = unpack(c, UR);
    PrimBar Lb{}, Rb{};
    const double SLa = L.un - L.a, SLr = Rr.un - Rr.a;
    const double SRa = L.un + L.a, SRr = Rr.un + Rr.a;
    const bool selL = !(SLr < SLa), selR = !(SRr > SRa);  // fmin / fmax pick the first on ties
    const double SL = selL ? SLa : SLr, SR = selR ? SRa : SRr;
    if (SL >= 0.0) {
        phys_flux_vjp(L, Fb, Lb);
    } else if (SR <= 0.0) {
        phys_flux_vjp(Rr, Fb, Rb);
    } else {
        double SLb = 0.0, SRb = 0.0;
        const double num = Rr.rho * Rr.un * (SR - Rr.un) - L.rho * L.un * (SL - L.un) + L.p - Rr.p;
        const double den = Rr.rho * (SR - Rr.un) - L.rho * (SL - L.un);
        const bool roeavg = !(::fabs(den) > 1e-12);
        double SM, iden = 0.0;
        double sqL = 0.0, sqR = 0.0;
        if (!roeavg) {
            iden = frcp(den);  // SM = num / den; its adjoints below reuse 1 / den
            SM = num * iden;
        } else {
            sqL = ::sqrt(L.rho);
            sqR = ::sqrt(Rr.rho);
            SM = (sqL * L.un + sqR * Rr.un) / (sqL + sqR);
        }
        const bool left = SM >= 0.0;
        const Prim<double> &s = left ? L : Rr;
        PrimBar &sb = left ? Lb : Rb;
        const double S = left ? SL : SR;
        const double *U = left ? UL : UR;
        double *Ubd = left ? ULb : URb;
        const double dS = S - s.un;
        const double iq = frcp(S - SM);
        const double idS = frcp(dS), irdS = s.ir * idS;  // 1 / dS, 1 / (rho dS)
        const double rs = s.rho * dS * iq;
        const double Y = s.p * irdS;
        const double Z = SM + Y;
        const double Es = s.E * s.ir + (SM - s.un) * Z;
        // F = Fs + S (Us - U)
        phys_flux_vjp(s, Fb, sb);
        const double Us[5] = {rs, rs * SM, rs * s.u1, rs * s.u2, rs * Es};
        double Sb = 0.0;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            Sb += Fb[n] * (Us[n] - U[n]);
            Ubd[n] -= S * Fb[n];
        }
        const double rsb = S * (Fb[0] + Fb[1] * SM + Fb[2] * s.u1 + Fb[3] * s.u2 + Fb[4] * Es);
        double SMb = S * Fb[1] * rs;
        sb.u1 += S * Fb[2] * rs;
        sb.u2 += S * Fb[3] * rs;
        const double Esb = S * Fb[4] * rs;
        // Es = E ir + (SM - un) Z, Z = SM + p / (rho dS)
        sb.E += Esb * s.ir;
        sb.ir += Esb * s.E;
        SMb += Esb * Z;
        sb.un -= Esb * Z;
        const double Zb = Esb * (SM - s.un);
        SMb += Zb;
        sb.p += Zb * irdS;
        sb.rho -= Zb * Y * s.ir;
        double dSb = -Zb * Y * idS;
        // rs = rho dS iq, iq = 1 / (S - SM), dS = S - un
        sb.rho += rsb * dS * iq;
        dSb += rsb * s.rho * iq;
        const double iqb = rsb * s.rho * dS;
        const double tq = -iqb * iq * iq;
        Sb += tq;
        SMb -= tq;
        Sb += dSb;
        sb.un -= dSb;
        if (left) SLb += Sb;
        else SRb += Sb;
        // contact speed
        if (!roeavg) {
            const double numb = SMb * iden, denb = -numb * SM;
            Rb.rho += numb * Rr.un * (SR - Rr.un) + denb * (SR - Rr.un);
            Rb.un += numb * Rr.rho * (SR - 2.0 * Rr.un) - denb * Rr.rho;
            SRb += numb * Rr.rho * Rr.un + denb * Rr.rho;
            Lb.rho -= numb * L.un * (SL - L.un) + denb * (SL - L.un);
            Lb.un -= numb * L.rho * (SL - 2.0 * L.un) - denb * L.rho;
            SLb -= numb * L.rho * L.un + denb * L.rho;
            Lb.p += numb;
            Rb.p -= numb;
        } else {
            const double Dn = sqL + sqR;
            const double Nmb = SMb / Dn, Dnb = -SMb * SM / Dn;
            const double sqLb = Nmb * L.un + Dnb, sqRb = Nmb * Rr.un + Dnb;
            Lb.un += Nmb * sqL;
            Rb.un += Nmb * sqR;
            Lb.rho += sqLb / (2.0 * sqL);
            Rb.rho += sqRb / (2.0 * sqR);
        }
        // wave speeds
        if (selL) {
            Lb.un += SLb;
            Lb.a -= SLb;
        } else {
            Rb.un += SLb;
            Rb.a -= SLb;
        }
        if (selR) {
            Lb.un += SRb;
            Lb.a += SRb;
        } else {
            Rb.un += SRb;
            Rb.a += SRb;
        }
    }
    unpack_vjp(c.gamma, UL, L, Lb, ULb);
    unpack_vjp(c.gamma, UR, Rr, Rb, URb);
}

// Reverse mode of rusanov.
NS3D_HD void rusanov_vjp(const SchemeCoef<double> &c, const double *UL, const double *UR, const double *Fb,
                         double *ULb, double *URb) {
    const Prim<double> L = unpack(c, UL), Rr = unpack(c, UR);
    PrimBar Lb{}, Rb{};
    const double sl = ::fabs(L.un) + L.a, sr = ::fabs(Rr.un) + Rr.a;
    const bool pickL = sl > sr;  // smax = sl > sr ? sl : sr
    const double smax = pickL ? sl : sr;
    double hb[5], smb = 0.0;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        hb[n] = 0.5 * Fb[n];
        ULb[n] += 0.5 * smax * Fb[n];
        URb[n] -= 0.5 * smax * Fb[n];
        smb -= 0.5 * Fb[n] * (UR[n] - UL[n]);
    }
    phys_flux_vjp(L, hb, Lb);
    phys_flux_vjp(Rr, hb, Rb);
    if (pickL) {
        Lb.un += L.un >= 0.0 ? smb : -smb;
        Lb.a += smb;
    } else {
        Rb.un += Rr.un >= 0.0 ? smb : -smb;
        Rb.a += smb;
    }
    unpack_vjp(c.gamma, UL, L, Lb, ULb);
    unpack_vjp(c.gamma, UR, Rr, Rb, URb);
}

// The scheme coefficients on Dual (constants: zero tangents).
NS3D_HD SchemeCoef<Dual> dual_coef(const SchemeCoef<double> &c) {
    SchemeCoef<Dual> d;
    d.gamma = Dual(c.gamma);
    d.p_ref = Dual(c.p_ref);
    d.H_ref = Dual(c.H_ref);
    d.E_ref = Dual(c.E_ref);
    for (int q = 0; q < 6; ++q) d.mdcd_lin[q] = Dual(c.mdcd_lin[q]);
    for (int q = 0; q < 4; ++q) d.mdcd_d[q] = Dual(c.mdcd_d[q]);
    return d;
}

// Reverse mode of riemann<RS, XS> by transposing its tangent: one
// forward-mode evaluation per input state component (ten), each contracted
// with Fb.  The same code and the same subgradients as the tangent-linear
// model, so the two stay exact transposes; it costs about ten Dual
// evaluations of the solver, for the solvers without a hand-written
// reverse (Roe with its entropy fix, HLL, HLLC_p, AUSM).
template <int RS, int XS>
NS3D_HD void riemann_vjp_dual(const SchemeCoef<double> &c, const double *UL, const double *UR, const double *Fb,
                              double *ULb, double *URb) {
    const SchemeCoef<Dual> cd = dual_coef(c);
#pragma unroll 1
    for (int i = 0; i < 10; ++i) {
        Dual L[5], R[5], F[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            L[n] = Dual(UL[n], i == n ? 1.0 : 0.0);
            R[n] = Dual(UR[n], i == 5 + n ? 1.0 : 0.0);
        }
        riemann<RS, XS>(cd, L, R, F);
        double sum = 0.0;
#pragma unroll
        for (int n = 0; n < 5; ++n) sum += F[n].d * Fb[n];
        if (i < 5) ULb[i] += sum;
        else URb[i - 5] += sum;
    }
}

// Reverse mode of riemann<RS, XS>: ULb += (dF/dUL)^T Fb, URb likewise.
template <int RS, int XS>
NS3D_HD void riemann_vjp(const SchemeCoef<double> &c, const double *UL, const double *UR, const double *Fb,
                         double *ULb, double *URb) {
    if constexpr (RS == R_HLLC) hllc_vjp(c, UL, UR, Fb, ULb, URb);
    else if constexpr (RS == R_RUSANOV) rusanov_vjp(c, UL, UR, Fb, ULb, URb);
    else riemann_vjp_dual<RS, XS>(c, UL, UR, Fb, ULb, URb);
}

// Reverse mode of the component-wise face flux face_flux<INTERP, RS, false>
// on the rotated stencil st[var][m] (cells i-2..i+3): stb += (dF/dst)^T Fb.
// XS: slot of the x velocity in the rotated frame (AUSM).
template <int INTERP, int RS, int XS = 1>
NS3D_HD void face_vjp(const SchemeCoef<double> &c, const double (&st)[5][6], const double *Fb, double (&stb)[5][6]) {
    double UL[5], UR[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) interp_lr<INTERP>(c, st[n], UL[n], UR[n]);
    double ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
    riemann_vjp<RS, XS>(c, UL, UR, Fb, ULb, URb);
    if constexpr (INTERP == I_MDCD_LINEAR || INTERP == I_MDCD_HYBRID) {
        // both states of the face from the same 6 cells, the right one mirrored
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            double fb[6] = {0, 0, 0, 0, 0, 0}, gb[6] = {0, 0, 0, 0, 0, 0};
            const double fr[6] = {st[n][5], st[n][4], st[n][3], st[n][2], st[n][1], st[n][0]};
            mdcd_vjp<INTERP>(c, st[n], ULb[n], fb);
            mdcd_vjp<INTERP>(c, fr, URb[n], gb);
#pragma unroll
            for (int m = 0; m < 6; ++m) stb[n][m] += fb[m] + gb[5 - m];
        }
        return;
    } else {
        static_assert(INTERP == I_WENO5 || INTERP == I_WENO5_JS, "adjoint: WCNS5, WENO5-JS or MDCD");
    }
    constexpr bool JS = INTERP == I_WENO5_JS;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        double fb[5] = {0, 0, 0, 0, 0};
        const double fl[5] = {st[n][0], st[n][1], st[n][2], st[n][3], st[n][4]};
        weno_vjp<JS>(fl, ULb[n], fb);
#pragma unroll
        for (int m = 0; m < 5; ++m) stb[n][m] += fb[m];
        double gb[5] = {0, 0, 0, 0, 0};
        const double fr[5] = {st[n][5], st[n][4], st[n][3], st[n][2], st[n][1]};
        weno_vjp<JS>(fr, URb[n], gb);
#pragma unroll
        for (int m = 0; m < 5; ++m) stb[n][5 - m] += gb[m];
    }
}

// ------------------------------------------------------------ viscous terms
// Reverse mode of visc_local: from the adjoints Vb of the nine flux
// components, accumulate gb (twelve gradients) and qb (u, v, w, T).
NS3D_HD void visc_local_vjp(double mu_ref, double S_ref, double cp_pr, double u, double v, double w, double T,
                            const double (&g)[4][3], const double (&Vb)[NVISC], double (&gb)[4][3],
                            double (&qb)[4]) {
    const double mu = sutherland(mu_ref, S_ref, T);
    const double divU = g[0][0] + g[1][1] + g[2][2];
    const double txx = 2.0 * mu * g[0][0] - (2.0 / 3.0) * mu * divU;
    const double tyy = 2.0 * mu * g[1][1] - (2.0 / 3.0) * mu * divU;
    const double tzz = 2.0 * mu * g[2][2] - (2.0 / 3.0) * mu * divU;
    const double txy = mu * (g[0][1] + g[1][0]);
    const double txz = mu * (g[0][2] + g[2][0]);
    const double tyz = mu * (g[1][2] + g[2][1]);
    const double kq = mu * cp_pr;
    const double ex = Vb[VEX], ey = Vb[VEY], ez = Vb[VEZ];
    // energy fluxes u . tau + kq grad T
    double txxb = Vb[VXX] + ex * u, tyyb = Vb[VYY] + ey * v, tzzb = Vb[VZZ] + ez * w;
    double txyb = Vb[VXY] + ex * v + ey * u, txzb = Vb[VXZ] + ex * w + ez * u, tyzb = Vb[VYZ] + ey * w + ez * v;
    qb[0] += ex * txx + ey * txy + ez * txz;
    qb[1] += ex * txy + ey * tyy + ez * tyz;
    qb[2] += ex * txz + ey * tyz + ez * tzz;
    const double kqb = ex * g[3][0] + ey * g[3][1] + ez * g[3][2];
    gb[3][0] += ex * kq;
    gb[3][1] += ey * kq;
    gb[3][2] += ez * kq;
    // stresses
    double mub = kqb * cp_pr;
    mub += txxb * (2.0 * g[0][0] - (2.0 / 3.0) * divU) + tyyb * (2.0 * g[1][1] - (2.0 / 3.0) * divU) +
           tzzb * (2.0 * g[2][2] - (2.0 / 3.0) * divU) + txyb * (g[0][1] + g[1][0]) + txzb * (g[0][2] + g[2][0]) +
           tyzb * (g[1][2] + g[2][1]);
    const double divb = -(2.0 / 3.0) * mu * (txxb + tyyb + tzzb);
    gb[0][0] += 2.0 * mu * txxb + divb;
    gb[1][1] += 2.0 * mu * tyyb + divb;
    gb[2][2] += 2.0 * mu * tzzb + divb;
    gb[0][1] += mu * txyb;
    gb[1][0] += mu * txyb;
    gb[0][2] += mu * txzb;
    gb[2][0] += mu * txzb;
    gb[1][2] += mu * tyzb;
    gb[2][1] += mu * tyzb;
    // mu = mu_ref T^1.5 (1 + S) / (T + S): dmu/dT = mu (1.5 / T - 1 / (T + S))
    qb[3] += mub * mu * (1.5 * frcp(T) - frcp(T + S_ref));
}

// Reverse mode of the cell primitives (u, v, w, T) of a conserved state U
// (conservedToPrimitive, field_structures.h:444-466, with T = p / (rho Rgas)):
// Ub += (dq/dU)^T qb.
NS3D_HD void c2p_vjp(double gamma, double Rgas, const double *U, const double *qb, double *Ub) {
    const double ir = frcp(U[0]);
    const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
    const double cT = fdiv(gamma - 1.0, Rgas) * ir;  // dT/dE
    // T = (gamma - 1) / Rgas (E / rho - |u|^2 / 2)
    Ub[0] += -(qb[0] * u + qb[1] * v + qb[2] * w) * ir + qb[3] * cT * (u * u + v * v + w * w - U[4] * ir);
    Ub[1] += qb[0] * ir - qb[3] * cT * u;
    Ub[2] += qb[1] * ir - qb[3] * cT * v;
    Ub[3] += qb[2] * ir - qb[3] * cT * w;
    Ub[4] += qb[3] * cT;
}

}  // namespace ns3d
