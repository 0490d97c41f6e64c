// ns3d_face_impl.h -- the per-direction face-flux kernel (compute_invis_flux +
// WCNS_Riemann_InviscidFlux, InviscousFlux.cpp:10-268), instantiated once per
// Riemann solver in face_<solver>.hip.
//
// One thread per face: the thread at interior cell c computes face c+1/2 along
// DIR from cells c-2..c+3, which lie inside the three ghost layers.  Lanes run
// along x in every direction, so each stencil load is one coalesced row.
#pragma once
#include "ns3d_bounded.h"
#include "ns3d_common.h"

namespace ns3d {

template <int DIR, int INTERP, int RS, bool CHAR, class R>
__global__ void __launch_bounds__(256) k_face(FaceArgs<R> a) {
    int i, j, k;
    if (!tile_cell<DIR == 1 ? TILE_Y_FAST : TILE_Z_FAST>(a.b, i, j, k)) return;
    if (i >= a.b.nx || j >= a.b.ny) return;
    // rotated slots: (n, t1, t2) = (u,v,w) / (v,w,u) / (w,u,v)
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
    const long sxy = (long)a.b.sx * a.b.sy;
    const long st = DIR == 0 ? 1 : (DIR == 1 ? (long)a.b.sx : sxy);
    const long c0 = a.b.gidx(i, j, k);
    R s[5][6];
    const bool mixed = sizeof(R) == 4;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const long id = c0 + (m - 2) * st;
        s[0][m] = (R)a.U[0][id];
        s[1][m] = (R)a.U[P1][id];
        s[2][m] = (R)a.U[P2][id];
        s[3][m] = (R)a.U[P3][id];
        s[4][m] = mixed ? (R)(a.U[4][id] - a.ref.E_ref) : (R)a.U[4][id];
    }
    R pc[2][5];
    if constexpr (CHAR) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const long id = c0 + side * st;
            pc[side][0] = (R)a.U[0][id];
            pc[side][1] = (R)a.Q[P1 - 1][id];
            pc[side][2] = (R)a.Q[P2 - 1][id];
            pc[side][3] = (R)a.Q[P3 - 1][id];
            pc[side][4] = (R)a.Q[3][id];
        }
    } else {
#pragma unroll
        for (int side = 0; side < 2; ++side)
#pragma unroll
            for (int n = 0; n < 5; ++n) pc[side][n] = R(0);
    }
    R F[5];
    face_flux<INTERP, RS, CHAR, XS>(a.c, s, pc, F);
    const long o = a.b.oidx(i, j, k);
    a.F[0][o] = (double)F[0];
    a.F[P1][o] = (double)F[1];
    a.F[P2][o] = (double)F[2];
    a.F[P3][o] = (double)F[3];
    a.F[4][o] = (double)F[4];
}

// Faces -1/2 .. n-1/2 of every interior line along DIR (compute_invis_flux's
// range [ng-1, ng+n-1], InviscousFlux.cpp:10-144, both boundary faces
// included), FP64, into the reference's face layout.  The stencils of the
// boundary faces reach into the ghost layers, which the boundary conditions
// fill.  Lanes run along x for the y and z faces (coalesced stencil rows).
template <int DIR, int INTERP, int RS, bool CHAR>
__global__ void __launch_bounds__(256) k_face_ref(FaceArgs<double> a) {
    const Box &b = a.b;
    const int nd = DIR == 0 ? b.nx : (DIR == 1 ? b.ny : b.nz);
    const long m = (long)(nd + 1) * (DIR == 0 ? b.ny : b.nx) * (DIR == 2 ? b.ny : b.nz);
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= m) return;
    int i, j, k, c;  // c: interior coordinate of the face's left cell, -1..nd-1
    if (DIR == 0) {
        c = (int)(id % (nd + 1)) - 1;
        const long r = id / (nd + 1);
        i = c;
        j = (int)(r % b.ny);
        k = (int)(r / b.ny);
    } else if (DIR == 1) {
        i = (int)(id % b.nx);
        const long r = id / b.nx;
        c = (int)(r % (nd + 1)) - 1;
        j = c;
        k = (int)(r / (nd + 1));
    } else {
        i = (int)(id % b.nx);
        const long r = id / b.nx;
        j = (int)(r % b.ny);
        c = (int)(r / b.ny) - 1;
        k = c;
    }
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
    const long st = DIR == 0 ? 1 : (DIR == 1 ? (long)b.sx : (long)b.sx * b.sy);
    const long c0 = b.gidx(i, j, k);
    double s[5][6];
#pragma unroll
    for (int m6 = 0; m6 < 6; ++m6) {
        const long g = c0 + (m6 - 2) * st;
        s[0][m6] = a.U[0][g];
        s[1][m6] = a.U[P1][g];
        s[2][m6] = a.U[P2][g];
        s[3][m6] = a.U[P3][g];
        s[4][m6] = a.U[4][g];
    }
    double pc[2][5];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const long g = c0 + side * st;
        pc[side][0] = CHAR ? a.U[0][g] : 0.0;
        pc[side][1] = CHAR ? a.Q[P1 - 1][g] : 0.0;
        pc[side][2] = CHAR ? a.Q[P2 - 1][g] : 0.0;
        pc[side][3] = CHAR ? a.Q[P3 - 1][g] : 0.0;
        pc[side][4] = CHAR ? a.Q[3][g] : 0.0;
    }
    double F[5];
    face_flux<INTERP, RS, CHAR, XS>(a.c, s, pc, F);
    const long o = DIR == 0 ? face_ref_index(b, 0, c + NG, j + NG, k + NG)
                            : (DIR == 1 ? face_ref_index(b, 1, c + NG, i + NG, k + NG)
                                        : face_ref_index(b, 2, c + NG, i + NG, j + NG));
    a.F[0][o] = F[0];
    a.F[P1][o] = F[1];
    a.F[P2][o] = F[2];
    a.F[P3][o] = F[3];
    a.F[4][o] = F[4];
}

template <int DIR, int INTERP, int RS, bool CHAR>
inline void launch_face_ref_t(const FaceArgs<double> &a, hipStream_t s) {
    const Box &b = a.b;
    const int nd = DIR == 0 ? b.nx : (DIR == 1 ? b.ny : b.nz);
    const long m = (long)(nd + 1) * (DIR == 0 ? b.ny : b.nx) * (DIR == 2 ? b.ny : b.nz);
    hipLaunchKernelGGL((k_face_ref<DIR, INTERP, RS, CHAR>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, a);
}

template <int RS, int INTERP, bool CHAR>
inline void launch_face_ref_dir(int dir, const FaceArgs<double> &a, hipStream_t s) {
    switch (dir) {
        case 0: launch_face_ref_t<0, INTERP, RS, CHAR>(a, s); break;
        case 1: launch_face_ref_t<1, INTERP, RS, CHAR>(a, s); break;
        default: launch_face_ref_t<2, INTERP, RS, CHAR>(a, s); break;
    }
}

template <int RS, bool CHAR>
inline void launch_face_ref_interp(int dir, int interp, const FaceArgs<double> &a, hipStream_t s) {
    switch (interp) {
        case I_WENO5: launch_face_ref_dir<RS, I_WENO5, CHAR>(dir, a, s); break;
        case I_MDCD_LINEAR: launch_face_ref_dir<RS, I_MDCD_LINEAR, CHAR>(dir, a, s); break;
        case I_MDCD_HYBRID: launch_face_ref_dir<RS, I_MDCD_HYBRID, CHAR>(dir, a, s); break;
        default: launch_face_ref_dir<RS, I_WENO5_JS, CHAR>(dir, a, s); break;
    }
}

template <int RS>
void launch_face_ref(int dir, int interp, bool chr, const FaceArgs<double> &a, hipStream_t s) {
    if (chr) launch_face_ref_interp<RS, true>(dir, interp, a, s);
    else launch_face_ref_interp<RS, false>(dir, interp, a, s);
}

template <int DIR, int INTERP, int RS, bool CHAR>
inline void launch_face_t(int precision, const FaceArgs<double> &a64, const FaceArgs<float> &a32, hipStream_t s) {
    (void)precision;
    (void)a32;
    hipLaunchKernelGGL((k_face<DIR, INTERP, RS, CHAR, double>), tile_grid(a64.b), cell_block(), 0, s, a64);
}

template <int RS, int INTERP, bool CHAR>
inline void launch_face_dir(int dir, int precision, const FaceArgs<double> &a64, const FaceArgs<float> &a32,
                            hipStream_t s) {
    switch (dir) {
        case 0: launch_face_t<0, INTERP, RS, CHAR>(precision, a64, a32, s); break;
        case 1: launch_face_t<1, INTERP, RS, CHAR>(precision, a64, a32, s); break;
        default: launch_face_t<2, INTERP, RS, CHAR>(precision, a64, a32, s); break;
    }
}

template <int RS, bool CHAR>
inline void launch_face_interp(int dir, int interp, int precision, const FaceArgs<double> &a64,
                               const FaceArgs<float> &a32, hipStream_t s) {
    switch (interp) {
        case I_WENO5: launch_face_dir<RS, I_WENO5, CHAR>(dir, precision, a64, a32, s); break;
        case I_MDCD_LINEAR: launch_face_dir<RS, I_MDCD_LINEAR, CHAR>(dir, precision, a64, a32, s); break;
        case I_MDCD_HYBRID: launch_face_dir<RS, I_MDCD_HYBRID, CHAR>(dir, precision, a64, a32, s); break;
        default: launch_face_dir<RS, I_WENO5_JS, CHAR>(dir, precision, a64, a32, s); break;
    }
}

template <int RS>
void launch_face(int dir, int interp, bool chr, int precision, const FaceArgs<double> &a64,
                 const FaceArgs<float> &a32, hipStream_t s) {
    if (chr) launch_face_interp<RS, true>(dir, interp, precision, a64, a32, s);
    else launch_face_interp<RS, false>(dir, interp, precision, a64, a32, s);
}

}  // namespace ns3d

#define NS3D_INSTANTIATE_FACE(RS)                                                                        \
    template void ns3d::launch_face<RS>(int, int, bool, int, const ns3d::FaceArgs<double> &,            \
                                        const ns3d::FaceArgs<float> &, hipStream_t);                    \
    template void ns3d::launch_face_ref<RS>(int, int, bool, const ns3d::FaceArgs<double> &, hipStream_t);
