// ns3d_hostcheck.hip -- host-side evaluation of the device derivative
// numerics on single stencils, for verification without a GPU: the tangent
// (the forward numerics on Dual) and the hand-written reverse mode of one
// x-normal face flux, of the pointwise viscous terms and of the
// conserved -> primitive map.  The adjoint kernels call the same functions.
#include "../../include/ns3d_hip.h"
#include "ns3d_adjoint.h"
#include "ns3d_dual.h"
#include "ns3d_kernels.h"

using namespace ns3d;

namespace {

template <class R>
SchemeCoef<R> coef_of(const ns3d_params *p) {
    SchemeCoef<R> c;
    c.gamma = R(p->gamma);
    c.p_ref = c.H_ref = c.E_ref = R(0.0);
    const double diss = p->mdcd_diss, disp = p->mdcd_disp;
    const double lin[6] = {(3.0 * disp + 3.0 * diss) / 8.0, (-18.0 * disp - 30 * diss - 1.0) / 16.0,
                           (12.0 * disp + 60.0 * diss + 9.0) / 16.0, (12.0 * disp - 60.0 * diss + 9.0) / 16.0,
                           (-18.0 * disp + 30.0 * diss - 1.0) / 16.0, (3.0 * disp - 3.0 * diss) / 8.0};
    for (int q = 0; q < 6; ++q) c.mdcd_lin[q] = R(lin[q]);
    c.mdcd_d[0] = R(1.5 * (disp + diss));
    c.mdcd_d[1] = R(0.5 - 1.5 * (disp - 3 * diss));
    c.mdcd_d[2] = R(0.5 - 1.5 * (disp + 3 * diss));
    c.mdcd_d[3] = R(1.5 * (disp - diss));
    return c;
}

template <int INTERP, int RS, bool CHAR>
void face_tangent(const SchemeCoef<Dual> &c, const double *st, const double *dst, double *F, double *dF) {
    Dual s[5][6], pc[2][5], out[5];
    for (int n = 0; n < 5; ++n)
        for (int m = 0; m < 6; ++m) s[n][m] = Dual(st[6 * n + m], dst[6 * n + m]);
    if constexpr (CHAR) {
        for (int side = 0; side < 2; ++side) {
            const Dual u[5] = {s[0][2 + side], s[1][2 + side], s[2][2 + side], s[3][2 + side], s[4][2 + side]};
            const Dual ir = Dual(1.0) / u[0];
            pc[side][0] = u[0];
            pc[side][1] = u[1] * ir;
            pc[side][2] = u[2] * ir;
            pc[side][3] = u[3] * ir;
            pc[side][4] = (c.gamma - Dual(1.0)) *
                          (u[4] - Dual(0.5) * u[0] *
                                      (pc[side][1] * pc[side][1] + pc[side][2] * pc[side][2] + pc[side][3] * pc[side][3]));
        }
    }
    face_flux<INTERP, RS, CHAR, 1>(c, s, pc, out);
    for (int n = 0; n < 5; ++n) {
        F[n] = out[n].v;
        dF[n] = out[n].d;
    }
}

template <int INTERP, bool CHAR>
int face_tangent_rs(int rs, const SchemeCoef<Dual> &c, const double *st, const double *dst, double *F, double *dF) {
    switch (rs) {
        case R_ROE: face_tangent<INTERP, R_ROE, CHAR>(c, st, dst, F, dF); break;
        case R_RUSANOV: face_tangent<INTERP, R_RUSANOV, CHAR>(c, st, dst, F, dF); break;
        case R_HLLC: face_tangent<INTERP, R_HLLC, CHAR>(c, st, dst, F, dF); break;
        case R_HLL: face_tangent<INTERP, R_HLL, CHAR>(c, st, dst, F, dF); break;
        case R_HLLC_P: face_tangent<INTERP, R_HLLC_P, CHAR>(c, st, dst, F, dF); break;
        case R_AUSM: face_tangent<INTERP, R_AUSM, CHAR>(c, st, dst, F, dF); break;
        default: return NS3D_E_ARG;
    }
    return NS3D_OK;
}

template <bool CHAR>
int face_tangent_all(const ns3d_params *p, const double *st, const double *dst, double *F, double *dF) {
    const SchemeCoef<Dual> c = coef_of<Dual>(p);
    switch (p->interpolation) {
        case I_WENO5: return face_tangent_rs<I_WENO5, CHAR>(p->riemann_solver, c, st, dst, F, dF);
        case I_MDCD_LINEAR: return face_tangent_rs<I_MDCD_LINEAR, CHAR>(p->riemann_solver, c, st, dst, F, dF);
        case I_MDCD_HYBRID: return face_tangent_rs<I_MDCD_HYBRID, CHAR>(p->riemann_solver, c, st, dst, F, dF);
        case I_WENO5_JS: return face_tangent_rs<I_WENO5_JS, CHAR>(p->riemann_solver, c, st, dst, F, dF);
        default: return NS3D_E_ARG;
    }
}

template <int INTERP>
void face_vjp_rs(int rs, const SchemeCoef<double> &c, const double (&s)[5][6], const double *Fb,
                 double (&sb)[5][6]) {
    switch (rs) {
        case R_ROE: face_vjp<INTERP, R_ROE>(c, s, Fb, sb); break;
        case R_RUSANOV: face_vjp<INTERP, R_RUSANOV>(c, s, Fb, sb); break;
        case R_HLLC: face_vjp<INTERP, R_HLLC>(c, s, Fb, sb); break;
        case R_HLL: face_vjp<INTERP, R_HLL>(c, s, Fb, sb); break;
        case R_HLLC_P: face_vjp<INTERP, R_HLLC_P>(c, s, Fb, sb); break;
        default: face_vjp<INTERP, R_AUSM>(c, s, Fb, sb); break;
    }
}

}  // namespace

extern "C" int ns3d_check_face_tangent(const ns3d_params *p, const double *st, const double *dst, double *F,
                                       double *dF) {
    if (!p || !st || !dst || !F || !dF) return NS3D_E_ARG;
    return p->char_recon ? face_tangent_all<true>(p, st, dst, F, dF) : face_tangent_all<false>(p, st, dst, F, dF);
}

extern "C" int ns3d_check_face_adjoint(const ns3d_params *p, const double *st, const double *Fb, double *stb) {
    if (!p || !st || !Fb || !stb) return NS3D_E_ARG;
    if (!adjoint_supported(p->interpolation, p->riemann_solver, p->char_recon != 0)) return NS3D_E_UNSUPPORTED;
    const SchemeCoef<double> c = coef_of<double>(p);
    double s[5][6], sb[5][6] = {};
    for (int n = 0; n < 5; ++n)
        for (int m = 0; m < 6; ++m) s[n][m] = st[6 * n + m];
    switch (p->interpolation) {
        case I_WENO5_JS: face_vjp_rs<I_WENO5_JS>(p->riemann_solver, c, s, Fb, sb); break;
        case I_MDCD_LINEAR: face_vjp_rs<I_MDCD_LINEAR>(p->riemann_solver, c, s, Fb, sb); break;
        case I_MDCD_HYBRID: face_vjp_rs<I_MDCD_HYBRID>(p->riemann_solver, c, s, Fb, sb); break;
        default: face_vjp_rs<I_WENO5>(p->riemann_solver, c, s, Fb, sb); break;
    }
    for (int n = 0; n < 5; ++n)
        for (int m = 0; m < 6; ++m) stb[6 * n + m] = sb[n][m];
    return NS3D_OK;
}

extern "C" int ns3d_check_visc_tangent(const ns3d_params *p, const double *q, const double *g, const double *dq,
                                       const double *dg, double *V, double *dV) {
    if (!p || !q || !g || !dq || !dg || !V || !dV) return NS3D_E_ARG;
    Dual gd[4][3], out[NVISC];
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) gd[a][d] = Dual(g[3 * a + d], dg[3 * a + d]);
    visc_local(Dual(p->mu), Dual(p->S_ref), Dual(p->Cp / p->Pr), Dual(q[0], dq[0]), Dual(q[1], dq[1]),
               Dual(q[2], dq[2]), Dual(q[3], dq[3]), gd, out);
    for (int n = 0; n < NVISC; ++n) {
        V[n] = out[n].v;
        dV[n] = out[n].d;
    }
    return NS3D_OK;
}

extern "C" int ns3d_check_visc_adjoint(const ns3d_params *p, const double *q, const double *g, const double *Vb,
                                       double *gb, double *qb) {
    if (!p || !q || !g || !Vb || !gb || !qb) return NS3D_E_ARG;
    double gg[4][3], gbb[4][3] = {}, vb[NVISC], qq[4] = {};
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) gg[a][d] = g[3 * a + d];
    for (int n = 0; n < NVISC; ++n) vb[n] = Vb[n];
    visc_local_vjp(p->mu, p->S_ref, p->Cp / p->Pr, q[0], q[1], q[2], q[3], gg, vb, gbb, qq);
    for (int a = 0; a < 4; ++a) {
        for (int d = 0; d < 3; ++d) gb[3 * a + d] = gbb[a][d];
        qb[a] = qq[a];
    }
    return NS3D_OK;
}

extern "C" int ns3d_check_c2p_adjoint(const ns3d_params *p, const double *U, const double *qb, double *Ub) {
    if (!p || !U || !qb || !Ub) return NS3D_E_ARG;
    for (int n = 0; n < 5; ++n) Ub[n] = 0.0;
    c2p_vjp(p->gamma, p->Rgas, U, qb, Ub);
    return NS3D_OK;
}

extern "C" int ns3d_check_weno_pair_adjoint(const ns3d_params *p, const double *f, double ybL, double ybR,
                                            double *fb_pair, double *fb_single) {
    if (!p || !f || !fb_pair || !fb_single) return NS3D_E_ARG;
    if (p->interpolation != I_WENO5 && p->interpolation != I_WENO5_JS) return NS3D_E_UNSUPPORTED;
    const bool js = p->interpolation == I_WENO5_JS;
    for (int m = 0; m < 5; ++m) fb_pair[m] = fb_single[m] = 0.0;
    if (js) weno_pair_vjp<true>(f, ybL, ybR, fb_pair);
    else weno_pair_vjp<false>(f, ybL, ybR, fb_pair);
    const double g[5] = {f[4], f[3], f[2], f[1], f[0]};
    double gb[5] = {0, 0, 0, 0, 0};
    if (js) {
        weno_vjp<true>(f, ybL, fb_single);
        weno_vjp<true>(g, ybR, gb);
    } else {
        weno_vjp<false>(f, ybL, fb_single);
        weno_vjp<false>(g, ybR, gb);
    }
    for (int m = 0; m < 5; ++m) fb_single[m] += gb[4 - m];
    return NS3D_OK;
}
