// ns3d_bounded.h -- the RHS + RK3 path of a box with non-periodic faces
// (Wall, Symmetry, Outflow, Inflow; boundary.cpp, ViscousFlux.cpp:15-50).
//
// It follows the reference's pass structure rather than the fused pencils of
// the periodic box: the face fluxes of an axis with physical faces include the
// two boundary faces and take their stencils from the ghost layers the
// boundary conditions fill, the ghost faces beyond them are set by the
// boundary type, and the viscous stencils drop to lower orders next to a
// physical face.  Face fluxes are stored in the reference's face layout
// (idx_fx / idx_fy / idx_fz, field_structures.h:73-82).  FP64 only.
#pragma once
#include "ns3d_kernels.h"

namespace ns3d {

enum { BC_PERIODIC = 0, BC_WALL = 1, BC_SYMMETRY = 2, BC_OUTFLOW = 3, BC_INFLOW = 4 };

// exchange + boundary fill of the primitives rho, u, v, w, p
struct BndArgs {
    double *P[5];  // ghosted rho, u, v, w, p
    Box b;
    int per[3];  // axis periodic (both faces Periodic)
    int face;    // fill: face 0..5 (xmin, xmax, ymin, ymax, zmin, zmax)
    int type;    // fill: its BC_* type
};

// ghost faces of one direction's reference-layout face arrays
struct GhostFaceArgs {
    double *F[5];
    Box b;
    int dir;
    int per;                // the axis is periodic
    int type_lo, type_hi;  // BC_* of its two faces
};

// gradients (choose_scheme orders) and viscous fluxes of the interior cells
struct BndViscArgs {
    const double *Q[4];  // ghosted u, v, w, T
    double *G[12];       // compact gradients (nullptr: not stored)
    double *V[NVISC];    // compact viscous fluxes (nullptr: not computed)
    Box b;
    int per[3];
    double h[3];
    double mu, S_ref, cp_pr;
};

// RHS of the interior cells: - half-node differences of the face fluxes
// (reference layout, ghost faces set) + viscous divergence (choose_scheme
// orders, compact viscous fluxes); mode: ASM_INVISCID / ASM_VISCOUS /
// ASM_ACCUMULATE bits as for the stepwise assembly
struct BndRhsArgs {
    const double *F[3][5];
    const double *V[NVISC];
    double *RHS[5];
    Box b;
    int per[3];
    double h[3];
    int mode;
};

// one Shu-Osher stage on all cells (RHS taken as 0 in the ghost layers)
struct RkAllArgs {
    double *U[5];
    const double *U0[5];
    const double *R[5];
    Box b;
    const double *dt;  // device scalar, or nullptr to use dtv
    double dtv;
    int stage;  // 1..3
};

void launch_bnd_exchange(const BndArgs &a, hipStream_t s);
void launch_bnd_fill(const BndArgs &a, hipStream_t s);
void launch_ghost_faces(const GhostFaceArgs &a, hipStream_t s);
void launch_bnd_visc(const BndViscArgs &a, hipStream_t s);
void launch_bnd_rhs(const BndRhsArgs &a, hipStream_t s);
void launch_rk_all(const RkAllArgs &a, hipStream_t s);

// reference-layout face fluxes of one direction, faces -1/2 .. n-1/2 of every
// interior line (face_<solver>.hip)
template <int RS>
void launch_face_ref(int dir, int interp, bool chr, const FaceArgs<double> &a, hipStream_t s);

// reference face index of the face between ghosted cells q and q+1 along dir,
// at ghosted tangential position (p1, p2) = (j, k) / (i, k) / (i, j)
__host__ __device__ inline long face_ref_index(const Box &b, int dir, int q, int p1, int p2) {
    if (dir == 0) return ((long)p2 * b.sy + p1) * (b.sx - 1) + q;
    if (dir == 1) return ((long)p2 * (b.sy - 1) + q) * b.sx + p1;
    return ((long)q * b.sy + p2) * b.sx + p1;
}

}  // namespace ns3d
