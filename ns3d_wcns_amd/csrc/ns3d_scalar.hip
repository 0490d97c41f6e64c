// ns3d_scalar.hip -- single stencils on the host: the scalar functions of
// ns3d_func.h:73-131 (interpolate_select, the six X_Riemann_solver,
// WCNS_Riemann_InviscidFlux) evaluated with the numerics the kernels run
// (ns3d_device.h, compiled here for the host; IEEE division and square
// root).  The face normal is a coordinate axis, as in every call the
// reference's flux loops make: the state is permuted so that the normal
// velocity comes first (the rotated frame of the kernels) and the flux is
// permuted back.
#include <cstdarg>
#include <cstdio>

#include "ns3d_internal.h"

using namespace ns3d;

namespace {

// conserved slots of the normal and the two tangential velocities for a
// normal along axis DIR, and the rotated slot of the x velocity
template <int DIR>
struct Axes {
    static constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    static constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
};

template <int RS, int DIR>
void riemann_axis(const SchemeCoef<double> &c, const double *l, const double *r, double *F) {
    using A = Axes<DIR>;
    const double a[5] = {l[0], l[A::P1], l[A::P2], l[A::P3], l[4]};
    const double b[5] = {r[0], r[A::P1], r[A::P2], r[A::P3], r[4]};
    double f[5];
    riemann<RS, A::XS>(c, a, b, f);
    F[0] = f[0];
    F[A::P1] = f[1];
    F[A::P2] = f[2];
    F[A::P3] = f[3];
    F[4] = f[4];
}

template <int RS>
void riemann_dim(int dim, const SchemeCoef<double> &c, const double *l, const double *r, double *F) {
    if (dim == 0) riemann_axis<RS, 0>(c, l, r, F);
    else if (dim == 1) riemann_axis<RS, 1>(c, l, r, F);
    else riemann_axis<RS, 2>(c, l, r, F);
}

void riemann_any(int rs, int dim, const SchemeCoef<double> &c, const double *l, const double *r, double *F) {
    switch (rs) {
        case R_ROE: riemann_dim<R_ROE>(dim, c, l, r, F); break;
        case R_RUSANOV: riemann_dim<R_RUSANOV>(dim, c, l, r, F); break;
        case R_HLLC: riemann_dim<R_HLLC>(dim, c, l, r, F); break;
        case R_HLL: riemann_dim<R_HLL>(dim, c, l, r, F); break;
        case R_HLLC_P: riemann_dim<R_HLLC_P>(dim, c, l, r, F); break;
        default: riemann_dim<R_AUSM>(dim, c, l, r, F); break;
    }
}

template <int INTERP, int RS, bool CHAR, int DIR>
void face_axis(const SchemeCoef<double> &c, const double *Ut, const double *ut, double *F) {
    using A = Axes<DIR>;
    const int slot[5] = {0, A::P1, A::P2, A::P3, 4};
    double s[5][6], pc[2][5] = {};
    for (int n = 0; n < 5; ++n)
        for (int m = 0; m < 6; ++m) s[n][m] = Ut[6 * slot[n] + m];
    if constexpr (CHAR)
        for (int side = 0; side < 2; ++side)
            for (int n = 0; n < 5; ++n) pc[side][n] = ut[6 * slot[n] + 2 + side];
    double f[5];
    face_flux<INTERP, RS, CHAR, A::XS>(c, s, pc, f);
    for (int n = 0; n < 5; ++n) F[slot[n]] = f[n];
}

template <int INTERP, int RS, bool CHAR>
void face_dim(int dim, const SchemeCoef<double> &c, const double *Ut, const double *ut, double *F) {
    if (dim == 0) face_axis<INTERP, RS, CHAR, 0>(c, Ut, ut, F);
    else if (dim == 1) face_axis<INTERP, RS, CHAR, 1>(c, Ut, ut, F);
    else face_axis<INTERP, RS, CHAR, 2>(c, Ut, ut, F);
}

template <int INTERP, bool CHAR>
void face_rs(int rs, int dim, const SchemeCoef<double> &c, const double *Ut, const double *ut, double *F) {
    switch (rs) {
        case R_ROE: face_dim<INTERP, R_ROE, CHAR>(dim, c, Ut, ut, F); break;
        case R_RUSANOV: face_dim<INTERP, R_RUSANOV, CHAR>(dim, c, Ut, ut, F); break;
        case R_HLLC: face_dim<INTERP, R_HLLC, CHAR>(dim, c, Ut, ut, F); break;
        case R_HLL: face_dim<INTERP, R_HLL, CHAR>(dim, c, Ut, ut, F); break;
        case R_HLLC_P: face_dim<INTERP, R_HLLC_P, CHAR>(dim, c, Ut, ut, F); break;
        default: face_dim<INTERP, R_AUSM, CHAR>(dim, c, Ut, ut, F); break;
    }
}

template <bool CHAR>
void face_any(int interp, int rs, int dim, const SchemeCoef<double> &c, const double *Ut, const double *ut,
              double *F) {
    switch (interp) {
        case I_WENO5: face_rs<I_WENO5, CHAR>(rs, dim, c, Ut, ut, F); break;
        case I_MDCD_LINEAR: face_rs<I_MDCD_LINEAR, CHAR>(rs, dim, c, Ut, ut, F); break;
        case I_MDCD_HYBRID: face_rs<I_MDCD_HYBRID, CHAR>(rs, dim, c, Ut, ut, F); break;
        default: face_rs<I_WENO5_JS, CHAR>(rs, dim, c, Ut, ut, F); break;
    }
}

int error(int code, const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return set_error(code, buf);
}

const char *interp_name(int i) {
    switch (i) {
        case I_WENO5: return "WENO5";
        case I_ZERO: return "ZERO";
        case I_MDCD_LINEAR: return "MDCD_LINEAR";
        case I_MDCD_HYBRID: return "MDCD_HYBRID";
        default: return "WENO5_JS";
    }
}

// interpolation, Riemann solver and gamma of a parameter set (boundary
// types play no part here)
int check_scheme(const ns3d_params *p, bool need_interp) {
    if (!p) return error(NS3D_E_ARG, "null params");
    if (need_interp && (p->interpolation < NS3D_INTERP_WENO5 || p->interpolation > NS3D_INTERP_WENO5_JS))
        return error(NS3D_E_ARG, "unknown interpolation %d", p->interpolation);
    if (p->riemann_solver < NS3D_RIEMANN_ROE || p->riemann_solver > NS3D_RIEMANN_AUSM)
        return error(NS3D_E_ARG, "Unknown Riemann solver %d", p->riemann_solver);
    if (!(p->gamma > 1.0)) return error(NS3D_E_ARG, "gamma must exceed 1 (got %g)", p->gamma);
    return NS3D_OK;
}

}  // namespace

extern "C" {

int ns3d_scalar_interpolate(const ns3d_params *p, const double *st, int n, double flag, double *out) {
    if (!p || !out || (n > 0 && !st)) return error(NS3D_E_ARG, "scalar_interpolate: null argument");
    if (p->interpolation < NS3D_INTERP_WENO5 || p->interpolation > NS3D_INTERP_WENO5_JS)
        return error(NS3D_E_ARG, "unknown interpolation %d", p->interpolation);
    const int want = p->interpolation == I_ZERO ? 2 : 6;
    if (n != want) {
        return error(NS3D_E_ARG, "interpolate_select: interpolation '%s requires stencil size %d, got %d",
                     interp_name(p->interpolation), want, n);
    }
    if (p->interpolation == I_ZERO) {
        *out = flag < 0.0 ? st[1] : st[0];
        return NS3D_OK;
    }
    const SchemeCoef<double> c = scheme_coef64(*p);
    double l = 0.0, r = 0.0;
    switch (p->interpolation) {
        case I_WENO5: interp_lr<I_WENO5>(c, st, l, r); break;
        case I_MDCD_LINEAR: interp_lr<I_MDCD_LINEAR>(c, st, l, r); break;
        case I_MDCD_HYBRID: interp_lr<I_MDCD_HYBRID>(c, st, l, r); break;
        default: interp_lr<I_WENO5_JS>(c, st, l, r); break;
    }
    *out = flag < 0.0 ? r : l;
    return NS3D_OK;
}

int ns3d_scalar_riemann(const ns3d_params *p, const double *UL, const double *UR, int dim, double *F) {
    int rc = check_scheme(p, false);
    if (rc) return rc;
    if (!UL || !UR || !F) return error(NS3D_E_ARG, "scalar_riemann: null argument");
    if (dim < 0 || dim > 2) return error(NS3D_E_ARG, "scalar_riemann: dim %d is not 0, 1 or 2", dim);
    SchemeCoef<double> c{};
    c.gamma = p->gamma;
    riemann_any(p->riemann_solver, dim, c, UL, UR, F);
    return NS3D_OK;
}

int ns3d_scalar_face_flux(const ns3d_params *p, const double *Ut, const double *ut, int dim, double *F) {
    int rc = check_scheme(p, true);
    if (rc) return rc;
    if (p->interpolation == I_ZERO)
        return error(NS3D_E_UNSUPPORTED,
                     "scalar_face_flux: interpolation ZERO is not supported (2-point stencils; the face flux "
                     "takes 6)");
    if (!Ut || !F || (p->char_recon && !ut)) return error(NS3D_E_ARG, "scalar_face_flux: null argument");
    if (dim < 0 || dim > 2) return error(NS3D_E_ARG, "scalar_face_flux: dim %d is not 0, 1 or 2", dim);
    const SchemeCoef<double> c = scheme_coef64(*p);
    if (p->char_recon) face_any<true>(p->interpolation, p->riemann_solver, dim, c, Ut, ut, F);
    else face_any<false>(p->interpolation, p->riemann_solver, dim, c, Ut, ut, F);
    return NS3D_OK;
}

}  // extern "C"
