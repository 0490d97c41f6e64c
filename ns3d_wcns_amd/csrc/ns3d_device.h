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

#include <type_traits>

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
    R p_ref, H_ref;  // reference pressure and enthalpy of the perturbation form (0: plain)
    R E_ref;         // reference total energy, H_ref - p_ref
    R mdcd_lin[6];  // linear MDCD weights (interpolation.cpp:128-132)
    R mdcd_d[4];    // MDCD-WENO linear weights d0..d3 (interpolation.cpp:182-185)
};

// The FP32 arithmetic (mixed method) works in the perturbation form below;
// the FP64 arithmetic is plain, with the reference's operation sequence
// (its divisions and square roots by the iterations below).
template <class R>
constexpr bool kPert = sizeof(R) == 4;

// Division and reciprocal of the face, Riemann and viscous numerics (the
// state conversions, the stage update and the CFL formula divide with IEEE
// operations, as the reference).  FP32 (the mixed method): the hardware
// reciprocal (v_rcp_f32, about 1 ulp).  FP64: Newton-Raphson refinement of
// the hardware reciprocal estimate (v_rcp_f64), r <- r + r (1 - b r), twice,
// which roughly doubles the correct bits each time, and a / b = a r: within
// about 2 ulp, in six instructions instead of the eleven of the correctly
// rounded expansion (which rescales, refines and then corrects the
// quotient).  Operands are normal numbers here: densities, wave-speed
// differences and sums of positive WENO weights, all far from the
// subnormal and overflow ranges.  Host builds divide with IEEE operations.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double nr_rcp64(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}
#endif
template <class R>
NS3D_HD R frcp(R a) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (kPert<R>) return __builtin_amdgcn_rcpf(a);
    if constexpr (std::is_same<R, double>::value) return nr_rcp64(a);
#endif
    return R(1) / a;
}
template <class R>
NS3D_HD R fdiv(R a, R b) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (kPert<R>) return a * __builtin_amdgcn_rcpf(b);
    if constexpr (std::is_same<R, double>::value) return a * nr_rcp64(b);
#endif
    return a / b;
}
// Square root: FP32, the hardware v_sqrt_f32.  FP64: Goldschmidt's
// iteration from the hardware estimate y of 1/sqrt(a) (v_rsq_f64): g = a y,
// h = y / 2, r = 1/2 - g h, g += g r, h += h r, then two Newton corrections
// g += (a - g g) h; ten instructions instead of about seventeen with the
// operand scaling and special-case fix-ups of the correctly rounded
// expansion.  Operands are positive normal numbers (gamma p / rho, the
// Sutherland temperature).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double gs_sqrt64(double a) {
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, a), h, g);
    return __builtin_fma(__builtin_fma(-g, g, a), h, g);
}
#endif
template <class R>
NS3D_HD R fsqrt(R a) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (kPert<R>) return __builtin_amdgcn_sqrtf(a);
    if constexpr (std::is_same<R, double>::value) return gs_sqrt64(a);
#endif
    return sqrt(a);
}

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
    return fdiv(e0 * p0 + e1 * p1 + e2 * p2, R(8) * (e0 + e1 + e2));
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
    return fdiv(e0 * p0 + e1 * p1 + e2 * p2, R(6) * (e0 + e1 + e2));
}

template <class R>
NS3D_HD R mdcd_linear(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5) {
    return c.mdcd_lin[0] * f0 + c.mdcd_lin[1] * f1 + c.mdcd_lin[2] * f2 + c.mdcd_lin[3] * f3 +
           c.mdcd_lin[4] * f4 + c.mdcd_lin[5] * f5;
}

// MDCD hybrid (mdcd_hybrid_interpolate, interpolation.cpp:135-209): the
// smoothness sensor picks the linear MDCD scheme when sai > 0.4, otherwise a
// four-stencil WENO-Z-type combination with the 6-point indicator beta3.
// The 6-point smoothness indicator beta3 of the MDCD hybrid scheme
// (interpolation.cpp:193-200), in the reference's evaluation order and with
// its f0^2 coefficient 271799 (the indicator's definition gives 271779, so
// this beta3 exceeds it by f0^2 / 6048 and is c^2 / 6048 on constant data).
template <class T>
NS3D_HD T beta3_of(T f0, T f1, T f2, T f3, T f4, T f5) {
    return (T(271799) * (f0 * f0) +
            f0 * (T(-2380800) * f1 + T(4086352) * f2 - T(3462252) * f3 + T(1458762) * f4 - T(245620) * f5) +
            f1 * (T(5653317) * f1 - T(20427884) * f2 + T(17905032) * f3 - T(7727988) * f4 + T(1325006) * f5) +
            f2 * (T(19510972) * f2 - T(35817664) * f3 + T(15929912) * f4 - T(2792660) * f5) +
            f3 * (T(17195652) * f3 - T(15880404) * f4 + T(2863984) * f5) + f4 * (T(3824847) * f4 - T(1429976) * f5) +
            T(139633) * (f5 * f5)) *
           T(1.0 / 120960.0);
}

// off: the constant the FP32 arithmetic has removed from f (E_ref of the
// energy, or its characteristic image); beta3 is not shift-invariant (its
// integer coefficients do not sum to zero) and its large coefficients cancel
// to O(h^2): the FP32 arithmetic evaluates it in FP64 on f + off.
template <class R>
NS3D_HD R mdcd_hybrid(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5, R off) {
    const R eps = R(0.9 * 0.4 / (1.0 - 0.9 * 0.4) * 1e-4);
    const R a1 = fabs(f2 - f1) + fabs(f2 - R(2) * f1 + f0);
    const R b1 = fabs(f2 - f3) + fabs(f2 - R(2) * f3 + f4);
    const R a2 = fabs(f3 - f2) + fabs(f3 - R(2) * f2 + f1);
    const R b2 = fabs(f3 - f4) + fabs(f3 - R(2) * f4 + f5);
    const R sai1 = fdiv(R(2) * a1 * b1 + eps, a1 * a1 + b1 * b1 + eps);
    const R sai2 = fdiv(R(2) * a2 * b2 + eps, a2 * a2 + b2 * b2 + eps);
    // std::min(sai1, sai2) > 0.4
    const R sai = sai2 < sai1 ? sai2 : sai1;
    if (sai > R(0.4)) return mdcd_linear(c, f0, f1, f2, f3, f4, f5);
    // 1e-40 is below the FP32 normal range: the FP32 arithmetic uses 1e-30
    // and caps z at 1e15 (z^2 within range), which only rescales stencils
    // whose weight ratio to the others already exceeds 1e30
    const R eps_small = kPert<R> ? R(1e-30) : R(1e-40);
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
    R beta3;
    if constexpr (kPert<R>) {
        const double o = off;
        beta3 = (R)beta3_of<double>((double)f0 + o, (double)f1 + o, (double)f2 + o, (double)f3 + o, (double)f4 + o,
                                    (double)f5 + o);
    } else {
        (void)off;
        beta3 = beta3_of<R>(f0, f1, f2, f3, f4, f5);
    }
    const R tau6 = fabs(beta3 - (beta0 + R(4) * beta1 + beta2) * R(1.0 / 6.0));
    R z0 = R(20) + fdiv(tau6, beta0 + eps_small);
    R z1 = R(20) + fdiv(tau6, beta1 + eps_small);
    R z2 = R(20) + fdiv(tau6, beta2 + eps_small);
    R z3 = R(20) + fdiv(tau6, beta3 + eps_small);
    if constexpr (kPert<R>) {
        z0 = fmin(z0, R(1e15));
        z1 = fmin(z1, R(1e15));
        z2 = fmin(z2, R(1e15));
        z3 = fmin(z3, R(1e15));
    }
    const R w0 = c.mdcd_d[0] * (z0 * z0), w1 = c.mdcd_d[1] * (z1 * z1);
    const R w2 = c.mdcd_d[2] * (z2 * z2), w3 = c.mdcd_d[3] * (z3 * z3);
    return fdiv(w0 * q0 + w1 * q1 + w2 * q2 + w3 * q3, w0 + w1 + w2 + w3);
}

// off: see mdcd_hybrid (0 for the plain method).
template <int INTERP, class R>
NS3D_HD R interp6(const SchemeCoef<R> &c, R f0, R f1, R f2, R f3, R f4, R f5, R off = R(0)) {
    if constexpr (INTERP == I_WENO5) return wcns5(f0, f1, f2, f3, f4);
    else if constexpr (INTERP == I_WENO5_JS) return weno5js(f0, f1, f2, f3, f4);
    else if constexpr (INTERP == I_MDCD_LINEAR) return mdcd_linear(c, f0, f1, f2, f3, f4, f5);
    else return mdcd_hybrid(c, f0, f1, f2, f3, f4, f5, off);
}

// Both reconstructions of one cell from its 5-cell stencil f = cells c-2..c+2
// for the component-wise WCNS5 / WENO5-JS schemes: yL = interp(f0..f4) (the
// cell's value at its right face, the left state of face c+1/2) and yR =
// interp(f4..f0) (the value at its left face, the right state of face
// c-1/2).  The mirrored stencil has the same smoothness indicators with
// beta_0 and beta_2 swapped, so they are formed once; the values equal
// wcns5 / weno5js up to rounding.
template <int INTERP, class R>
NS3D_HD void recon_pair(const R *f, R &yL, R &yR) {
    static_assert(INTERP == I_WENO5 || INTERP == I_WENO5_JS, "recon_pair: WCNS5 or WENO5-JS");
    constexpr bool JS = INTERP == I_WENO5_JS;
    const R eps = R(1e-6);
    const R f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    const R t0 = f0 - R(2) * f1 + f2, s0 = f0 - R(4) * f1 + R(3) * f2;
    const R t1 = f1 - R(2) * f2 + f3, s1 = f1 - f3;
    const R t2 = f2 - R(2) * f3 + f4, s2 = R(3) * f2 - R(4) * f3 + f4;
    const R b0 = R(13.0 / 12.0) * t0 * t0 + R(0.25) * s0 * s0;
    const R b1 = R(13.0 / 12.0) * t1 * t1 + R(0.25) * s1 * s1;
    const R b2 = R(13.0 / 12.0) * t2 * t2 + R(0.25) * s2 * s2;
    const R q0 = (eps + b0) * (eps + b0), q1 = (eps + b1) * (eps + b1), q2 = (eps + b2) * (eps + b2);
    const R d0 = JS ? R(0.1) : R(1.0 / 16.0), d1 = JS ? R(0.6) : R(5.0 / 8.0), d2 = JS ? R(0.3) : R(5.0 / 16.0);
    const R K = JS ? R(6) : R(8);
    const R c0 = JS ? R(2) : R(3), c1 = JS ? R(-7) : R(-10), c2 = JS ? R(11) : R(15);
    const R m1 = JS ? R(5) : R(6), m2 = JS ? R(2) : R(3);
    const R q12 = q1 * q2, q02 = q0 * q2, q01 = q0 * q1;
    {
        const R e0 = d0 * q12, e1 = d1 * q02, e2 = d2 * q01;
        const R p0 = c0 * f0 + c1 * f1 + c2 * f2, p1 = -f1 + m1 * f2 + m2 * f3, p2 = m2 * f2 + m1 * f3 - f4;
        yL = fdiv(e0 * p0 + e1 * p1 + e2 * p2, K * (e0 + e1 + e2));
    }
    {
        const R e0 = d0 * q01, e1 = d1 * q02, e2 = d2 * q12;
        const R p0 = c0 * f4 + c1 * f3 + c2 * f2, p1 = -f3 + m1 * f2 + m2 * f1, p2 = m2 * f2 + m1 * f1 - f0;
        yR = fdiv(e0 * p0 + e1 * p1 + e2 * p2, K * (e0 + e1 + e2));
    }
}

// Left (+1) and right (-1) states of one 6-point line s[0..5] = cells i-2..i+3
// (interpolate_select, interpolation.cpp:13-77: the right state reverses the
// stencil).
template <int INTERP, class R>
NS3D_HD void interp_lr(const SchemeCoef<R> &c, const R *s, R &left, R &right, R off = R(0)) {
    left = interp6<INTERP>(c, s[0], s[1], s[2], s[3], s[4], s[5], off);
    right = interp6<INTERP>(c, s[5], s[4], s[3], s[2], s[1], s[0], off);
}

// ------------------------------------------------------------ Riemann solvers
// Energy and pressure enter the solvers as offsets from a reference state:
// E = E' + E_ref, p = p' + p_ref, with H_ref = E_ref + p_ref.  The plain
// method has E_ref = p_ref = H_ref = 0 (E' = E, p' = p).  The mixed method
// carries O(1) perturbations through FP32 arithmetic; the constant p_ref is
// dropped from the momentum flux (its half-node difference is exactly 0) and
// H_ref is kept explicitly in the energy flux.  The pressure is recovered
// from the total energy in FP64 (pressure_of).

// Full pressure p' + p_ref and total enthalpy density E' + p' + H_ref.
template <class R>
NS3D_HD R pfull(const SchemeCoef<R> &c, R p) {
    if constexpr (kPert<R>) return p + c.p_ref;
    else return p;
}
template <class R>
NS3D_HD R hfull(const SchemeCoef<R> &c, R E, R p) {
    if constexpr (kPert<R>) return E + p + c.H_ref;
    else return E + p;
}

// p' = (gamma - 1) (E' - |m|^2 / (2 rho)) of a rotated conserved state; the
// subtraction runs in FP64 for every arithmetic type R.
template <class R>
NS3D_HD R pressure_of(R rho, R un, R u1, R u2, R E, R g) {
    if constexpr (sizeof(R) == 4) {
        const double r = rho, a = un, b = u1, c = u2;
        return (R)(((double)E - 0.5 * r * (a * a + b * b + c * c)) * ((double)g - 1.0));
    } else {
        return (E - R(0.5) * rho * (un * un + u1 * u1 + u2 * u2)) * (g - R(1));
    }
}

// Primitive unpacking of a rotated conserved state (E, p perturbations).
template <class R>
struct Prim {
    R rho, un, u1, u2, E, p, a, ir;
};

template <class R>
NS3D_HD Prim<R> unpack(const SchemeCoef<R> &c, const R *U) {
    Prim<R> s;
    s.rho = U[0];
    s.ir = frcp(s.rho);
    s.un = U[1] * s.ir;
    s.u1 = U[2] * s.ir;
    s.u2 = U[3] * s.ir;
    s.E = U[4];
    s.p = pressure_of(s.rho, s.un, s.u1, s.u2, U[4], c.gamma);
    s.a = fsqrt(c.gamma * pfull(c, s.p) * s.ir);
    return s;
}

// Physical x-normal flux of a rotated state (momentum without p_ref).
template <class R>
NS3D_HD void phys_flux(const SchemeCoef<R> &c, const Prim<R> &s, R *F) {
    const R m = s.rho * s.un;
    F[0] = m;
    F[1] = m * s.un + s.p;
    F[2] = m * s.u1;
    F[3] = m * s.u2;
    F[4] = hfull(c, s.E, s.p) * s.un;
}

// Rusanov (InviscousFlux.cpp:270-311)
template <class R>
NS3D_HD void rusanov(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const Prim<R> L = unpack(c, UL), Rr = unpack(c, UR);
    R FL[5], FR[5];
    phys_flux(c, L, FL);
    phys_flux(c, Rr, FR);
    const R sl = fabs(L.un) + L.a, sr = fabs(Rr.un) + Rr.a;
    const R smax = sl > sr ? sl : sr;
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = R(0.5) * (FL[n] + FR[n]) - R(0.5) * smax * (UR[n] - UL[n]);
}

// HLL (InviscousFlux.cpp:384-447), Davis wave speeds.
template <class R>
NS3D_HD void hll(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const Prim<R> L = unpack(c, UL), Rr = unpack(c, UR);
    R FL[5], FR[5];
    phys_flux(c, L, FL);
    phys_flux(c, Rr, FR);
    const R SL = fmin(L.un - L.a, Rr.un - Rr.a);
    const R SR = fmax(L.un + L.a, Rr.un + Rr.a);
    if (SL >= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FL[n];
    } else if (SR <= R(0)) {
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = FR[n];
    } else {
        const R id = frcp(SR - SL);
        const R f1 = SR * id, f2 = SL * id, f3 = SL * SR * id;
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = f1 * FL[n] - f2 * FR[n] + f3 * (UR[n] - UL[n]);
    }
}

// HLLC star-region flux, shared by HLLC and HLLC_p (InviscousFlux.cpp:538-591).
// Tangential velocities are unchanged across the contact; the star energy
// E* = E/rho + (SM - un)(SM + p / (rho (S - un))) enters through
// rho* E* - E = E (r - 1) + rho* (SM - un)(SM + p / (rho (S - un))), r = rho*/rho
// = (S - un)/(S - SM), written for E = E' + E_ref, p = p' + p_ref as
// E' (r - 1) + rho* (SM - un)(SM + p' / (rho (S - un))) + H_ref (SM - un)/(S - SM).
template <class R>
NS3D_HD void hllc_star(const SchemeCoef<R> &c, const Prim<R> &L, const Prim<R> &Rr, const R *UL, const R *UR,
                       const R *FL, const R *FR, R SL, R SR, R SM, R *F) {
    const bool left = SM >= R(0);
    const Prim<R> &s = left ? L : Rr;
    const R S = left ? SL : SR;
    const R *U = left ? UL : UR;
    const R *Fs = left ? FL : FR;
    const R dS = S - s.un;
    if constexpr (kPert<R>) {
        const R iSM = frcp(S - SM);
        const R rs = s.rho * dS * iSM;
        const R q = (SM - s.un) * iSM;
        const R dE = (s.E + c.H_ref) * q + rs * (SM - s.un) * (SM + fdiv(s.p, s.rho * dS));
        F[0] = Fs[0] + S * (rs - U[0]);
        F[1] = Fs[1] + S * (rs * SM - U[1]);
        F[2] = Fs[2] + S * (rs * s.u1 - U[2]);
        F[3] = Fs[3] + S * (rs * s.u2 - U[3]);
        F[4] = Fs[4] + S * dE;
    } else {
        const R rs = fdiv(s.rho * dS, S - SM);
        const R Es = s.E * s.ir + (SM - s.un) * (SM + fdiv(s.p, s.rho * dS));
        const R Us[5] = {rs, rs * SM, rs * s.u1, rs * s.u2, rs * Es};
#pragma unroll
        for (int n = 0; n < 5; ++n) F[n] = Fs[n] + S * (Us[n] - U[n]);
    }
}

// HLLC (InviscousFlux.cpp:449-592): Davis SL/SR, supersonic early-outs,
// pressure-based contact speed with the Roe-averaged normal velocity as the
// fallback when |den| <= 1e-12.
template <class R>
NS3D_HD void hllc(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const Prim<R> L = unpack(c, UL), Rr = unpack(c, UR);
    R FL[5], FR[5];
    phys_flux(c, L, FL);
    phys_flux(c, Rr, FR);
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
        SM = fdiv(num, den);
    } else {
        const R sqL = fsqrt(L.rho), sqR = fsqrt(Rr.rho);
        SM = fdiv(sqL * L.un + sqR * Rr.un, sqL + sqR);
    }
    hllc_star(c, L, Rr, UL, UR, FL, FR, SL, SR, SM, F);
}

// HLLC with pressure-based wave speeds (HLLC_p_Riemann_solver,
// InviscousFlux.cpp:594-717).
template <class R>
NS3D_HD void hllc_p(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const R g = c.gamma;
    const Prim<R> L = unpack(c, UL), Rr = unpack(c, UR);
    R FL[5], FR[5];
    phys_flux(c, L, FL);
    phys_flux(c, Rr, FR);
    R ps = R(0.5) * (L.p + Rr.p) - R(0.125) * (Rr.un - L.un) * (L.rho + Rr.rho) * (L.a + Rr.a);
    if constexpr (kPert<R>) ps = fmax(-c.p_ref, ps);  // full p* >= 0
    else ps = fmax(R(0), ps);
    const R c1 = (g + R(1)) / (R(2) * g), c2 = (g - R(1)) / (R(2) * g);
    const R SL = ps <= L.p ? L.un - L.a : L.un - L.a * fsqrt(fdiv(c1 * pfull(c, ps), pfull(c, L.p)) + c2);
    const R SR = ps <= Rr.p ? Rr.un + Rr.a : Rr.un + Rr.a * fsqrt(fdiv(c1 * pfull(c, ps), pfull(c, Rr.p)) + c2);
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
    hllc_star(c, L, Rr, UL, UR, FL, FR, SL, SR, SM, F);
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
    const R ic = frcp(c);
    const R ic2 = ic * ic;
    const R irho = frcp(rho);
    const R a2 = R(0.70710678118654752440) * irho * ic;  // 1 / (sqrt(2) rho c)
    const R a3 = R(0.70710678118654752440) * rho * ic;   // rho / (sqrt(2) c)
    const R a4 = fdiv(phi + c * c, a1);
    const R a5 = R(1) - phi * ic2;
    const R a6 = fdiv(phi, a1);
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
NS3D_HD void roe(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const R g = c.gamma;
    const Prim<R> L = unpack(c, UL), Rr = unpack(c, UR);
    R FL[5], FR[5];
    phys_flux(c, L, FL);
    phys_flux(c, Rr, FR);
    const R sL = fsqrt(L.rho), sR = fsqrt(Rr.rho);
    const R rb = sL * sR;
    const R dn = frcp(sL + sR);
    const R ub = (sL * L.un + sR * Rr.un) * dn;
    const R vb = (sL * L.u1 + sR * Rr.u1) * dn;
    const R wb = (sL * L.u2 + sR * Rr.u2) * dn;
    const R HL = hfull(c, L.E, L.p) * L.ir, HR = hfull(c, Rr.E, Rr.p) * Rr.ir;
    const R hb = (sL * HL + sR * HR) * dn;
    const R cb = fsqrt((g - R(1)) * (hb - R(0.5) * (ub * ub + vb * vb + wb * wb)));
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
        if (al_m < delta) al_m = fabs(fdiv(lam[m] * lam[m] + delta * delta, R(2) * delta));
        al[m] *= al_m;
    }
    R D[5];
    proj_R(e, al, D);
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = R(0.5) * (FL[n] + FR[n]) - R(0.5) * D[n];
}

// AUSM-type flux (AUSM_Riemann_solver, InviscousFlux.cpp:719-846).  XS is the
// rotated slot of the x velocity: the reference's pressure-diffusion term uses
// (u_R - u_L) of the x velocity for every face direction (line 823).  The
// pressure floor 1e-12 applies to the full pressure.
template <int XS, class R>
NS3D_HD void ausm(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    const R g = c.gamma;
    const R Kp = R(0.25), Ku = R(0.75);
    const R rL = UL[0], irL = frcp(rL);
    const R vL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
    const R pfL = fmax(pfull(c, pressure_of(rL, vL[0], vL[1], vL[2], UL[4], g)), R(1e-12));
    const R pL = kPert<R> ? pfL - c.p_ref : pfL;
    const R aL = fsqrt(g * pfL * irL);
    const R HL = hfull(c, UL[4], pL) * irL;
    const R rR = UR[0], irR = frcp(rR);
    const R vR[3] = {UR[1] * irR, UR[2] * irR, UR[3] * irR};
    const R pfR = fmax(pfull(c, pressure_of(rR, vR[0], vR[1], vR[2], UR[4], g)), R(1e-12));
    const R pR = kPert<R> ? pfR - c.p_ref : pfR;
    const R aR = fsqrt(g * pfR * irR);
    const R HR = hfull(c, UR[4], pR) * irR;
    const R ab = R(0.5) * (aL + aR);
    const R rhob = R(0.5) * (rL + rR);
    const R iab = frcp(ab);
    const R ML = vL[0] * iab, MR = vR[0] * iab;
    const R Mb2 = fdiv(vL[0] * vL[0] + vR[0] * vR[0], R(2) * ab * ab);
    const R Mo2 = fmin(R(1), fmax(Mb2, R(0.5)));
    const R fa = Mo2 * (R(2) - Mo2);
    const R Mp = fabs(ML) >= R(1) ? R(0.5) * (ML + fabs(ML)) : R(0.25) * (ML + R(1)) * (ML + R(1));
    const R Mm = fabs(MR) >= R(1) ? R(0.5) * (MR - fabs(MR)) : R(-0.25) * (MR - R(1)) * (MR - R(1));
    const R Mh = Mp + Mm - fdiv(Kp, fa) * fmax(R(1) - Mb2, R(0)) * fdiv(pR - pL, rhob * ab * ab);
    const R mdot = ab * Mh * (Mh >= R(0) ? rL : rR);
    const R Pp = fabs(ML) >= R(1) ? R(0.5) * (R(1) + (ML > R(0) ? R(1) : R(-1)))
                                  : R(0.25) * (ML + R(1)) * (ML + R(1)) * (R(2) - ML);
    const R Pm = fabs(MR) >= R(1) ? R(0.5) * (R(1) - (MR > R(0) ? R(1) : R(-1)))
                                  : R(0.25) * (MR - R(1)) * (MR - R(1)) * (R(2) + MR);
    const R dux = vR[XS - 1] - vL[XS - 1];
    // p_half - p_ref = Pp pL' + Pm pR' + p_ref (Pp + Pm - 1) - Ku ...
    R ph = Pp * pL + Pm * pR - Ku * Pm * Pp * (rL + rR) * fa * dux * ab;
    if constexpr (kPert<R>) ph = ph + c.p_ref * (Pp + Pm - R(1));
    const bool up = mdot >= R(0);
    F[0] = mdot;
    F[1] = mdot * (up ? vL[0] : vR[0]) + ph;
    F[2] = mdot * (up ? vL[1] : vR[1]);
    F[3] = mdot * (up ? vL[2] : vR[2]);
    F[4] = mdot * (up ? HL : HR);
}

template <int RS, int XS, class R>
NS3D_HD void riemann(const SchemeCoef<R> &c, const R *UL, const R *UR, R *F) {
    if constexpr (RS == R_ROE) roe(c, UL, UR, F);
    else if constexpr (RS == R_RUSANOV) rusanov(c, UL, UR, F);
    else if constexpr (RS == R_HLLC) hllc(c, UL, UR, F);
    else if constexpr (RS == R_HLL) hll(c, UL, UR, F);
    else if constexpr (RS == R_HLLC_P) hllc_p(c, UL, UR, F);
    else ausm<XS>(c, UL, UR, F);
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
        for (int n = 0; n < 5; ++n) interp_lr<INTERP>(c, st[n], UL[n], UR[n], n == 4 ? c.E_ref : R(0));
    } else {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            // eigenvectors of the single cell state i (left) or i+1 (right)
            const R rho = pc[side][0], u = pc[side][1], v = pc[side][2], w = pc[side][3], p = pc[side][4];
            const R cs = fsqrt(fdiv(c.gamma * pfull(c, p), rho));
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
                // characteristic image of the removed (0, 0, 0, 0, E_ref)
                const R off = e.L[n][4] * c.E_ref;
                if (side == 0) Q[n] = interp6<INTERP>(c, W[n][0], W[n][1], W[n][2], W[n][3], W[n][4], W[n][5], off);
                else Q[n] = interp6<INTERP>(c, W[n][5], W[n][4], W[n][3], W[n][2], W[n][1], W[n][0], off);
            }
            proj_R(e, Q, side == 0 ? UL : UR);
        }
    }
    riemann<RS, XS>(c, UL, UR, F);
}

}  // namespace ns3d
