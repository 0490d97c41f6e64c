// ns3d_visc.h -- pointwise viscous numerics shared by the forward kernels,
// the tangent-linear model (on Dual) and the mixed method (on float).
#pragma once
#include "ns3d_common.h"

namespace ns3d {

// 6th-order central first derivative of a 7-point line without its centre,
// times 1/(60 h) (diff_6th_central, ViscousFlux.cpp:27-29).
template <class R>
NS3D_HD R d6c(R fm3, R fm2, R fm1, R fp1, R fp2, R fp3, R i60h) {
    return (fp3 - R(9) * fp2 + R(45) * fp1 - R(45) * fm1 + R(9) * fm2 - fm3) * i60h;
}

// Sutherland viscosity mu(T) = mu_ref T^1.5 (1 + S) / (T + S)
// (SolverParams::get_mu, field_structures.h:96-106).
template <class R>
NS3D_HD R sutherland(R mu, R S, R T) {
    return fdiv(mu * (T * fsqrt(T)) * (R(1) + S), T + S);
}

// Stresses and energy fluxes of one cell from its velocity, temperature and
// the twelve gradients g[q][d] = d q / d x_d, q = u, v, w, T
// (compute_viscous_flux, ViscousFlux.cpp:151-208): the nine distinct
// components in the slots VXX..VEZ.
template <class R>
NS3D_HD void visc_local(R mu_ref, R S_ref, R cp_pr, R u, R v, R w, R T, const R (&g)[4][3], R (&V)[NVISC]) {
    const R mu = sutherland(mu_ref, S_ref, T);
    const R divU = g[0][0] + g[1][1] + g[2][2];
    const R txx = R(2) * mu * g[0][0] - R(2.0 / 3.0) * mu * divU;
    const R tyy = R(2) * mu * g[1][1] - R(2.0 / 3.0) * mu * divU;
    const R tzz = R(2) * mu * g[2][2] - R(2.0 / 3.0) * mu * divU;
    const R txy = mu * (g[0][1] + g[1][0]);
    const R txz = mu * (g[0][2] + g[2][0]);
    const R tyz = mu * (g[1][2] + g[2][1]);
    const R kq = mu * cp_pr;
    V[VXX] = txx;
    V[VYY] = tyy;
    V[VZZ] = tzz;
    V[VXY] = txy;
    V[VXZ] = txz;
    V[VYZ] = tyz;
    V[VEX] = u * txx + v * txy + w * txz + kq * g[3][0];
    V[VEY] = u * txy + v * tyy + w * tyz + kq * g[3][1];
    V[VEZ] = u * txz + v * tyz + w * tzz + kq * g[3][2];
}

}  // namespace ns3d
