// ns3d_adjoint.hip -- tangent-linear and adjoint kernels of the FP64 RHS
// (compute_rhs, range_kutta.cpp:11-43) on one periodic box.
//
// Tangent-linear: the forward numerics instantiated on Dual (ns3d_dual.h),
// so that any interpolation / Riemann / projection combination has its
// tangent:
//   k_tl_prim     (u, v, w, T) tangents of the interior cells;
//   k_tl_visc     tangents of the nine viscous flux components;
//   k_tl_face     tangent face fluxes of one direction (thread per face);
//   k_tl_assemble tangent RHS: half-node differences + viscous divergence.
// Adjoint (hand-written reverse mode, ns3d_adjoint.h; component-wise WCNS5 /
// WENO5-JS with HLLC / Rusanov):
//   k_adj_face    per face of one direction: the face adjoint from the
//                 transposed half-node difference of lambda, then the
//                 reverse of Riemann flux and interpolation, stored as the
//                 contributions to the face's six stencil cells;
//   k_adj_gather  per cell: sum of the contributions of its six faces;
//   k_adj_visc_a  per cell: viscous flux adjoints from the transposed
//                 divergence, reverse of stresses / heat flux -> gradient
//                 and (u, v, w, T) adjoints;
//   k_adj_visc_b  per cell: transposed gradient stencil and reverse of the
//                 conserved -> primitive map.
// All adjoint accumulation is by gathers (no atomics): results are
// deterministic.
#include "ns3d_adjoint.h"
#include "ns3d_dual.h"
#include "ns3d_kernels.h"

namespace ns3d {

namespace {

// Half-node difference weights of faces c-3..c+2 of cell c (times 1/h; the
// inviscid RHS is minus this sum).
__device__ constexpr double kW[6] = {-3.0 / 640.0, 25.0 / 384.0, -75.0 / 64.0, 75.0 / 64.0, -25.0 / 384.0,
                                     3.0 / 640.0};
// 6th-order central difference weights of cells c-3..c+3 (times 1/(60 h)).
__device__ constexpr double kD[7] = {-1.0, 9.0, -45.0, 0.0, 45.0, -9.0, 1.0};
__device__ constexpr int kSlot[3][4] = {{VXX, VXY, VXZ, VEX}, {VXY, VYY, VYZ, VEY}, {VXZ, VYZ, VZZ, VEZ}};

__device__ __forceinline__ bool cell_of(const Box &b, long id, int &i, int &j, int &k) {
    if (id >= (long)b.nx * b.ny * b.nz) return false;
    i = (int)(id % b.nx);
    const long r = id / b.nx;
    j = (int)(r % b.ny);
    k = (int)(r / b.ny);
    return true;
}

__device__ __forceinline__ long gstride(const Box &b, int dir) {
    return dir == 0 ? 1L : (dir == 1 ? (long)b.sx : (long)b.sx * b.sy);
}

// compact index of (i, j, k) moved by s cells along dir, periodically
__device__ __forceinline__ long owrap(const Box &b, int i, int j, int k, int dir, int s) {
    if (dir == 0) i = wrap(i + s, b.nx);
    else if (dir == 1) j = wrap(j + s, b.ny);
    else k = wrap(k + s, b.nz);
    return b.oidx(i, j, k);
}

}  // namespace

// ------------------------------------------------------------ tangent
__global__ void __launch_bounds__(256) k_tl_prim(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long g = a.b.gidx(i, j, k);
    const Dual r(a.U[0][g], a.X[0][g]);
    const Dual ir = Dual(1.0) / r;
    const Dual u = Dual(a.U[1][g], a.X[1][g]) * ir, v = Dual(a.U[2][g], a.X[2][g]) * ir,
               w = Dual(a.U[3][g], a.X[3][g]) * ir;
    const Dual kin = Dual(0.5) * (u * u + v * v + w * w);
    const Dual p = Dual(a.gamma - 1.0) * r * (Dual(a.U[4][g], a.X[4][g]) * ir - kin);
    const Dual T = p / (r * Dual(a.Rgas));
    a.dQ[0][g] = u.d;
    a.dQ[1][g] = v.d;
    a.dQ[2][g] = w.d;
    a.dQ[3][g] = T.d;
}

__global__ void __launch_bounds__(256) k_tl_visc(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const Box &b = a.b;
    const long c = b.gidx(i, j, k);
    const double *Qv[4] = {a.Q[0], a.Q[1], a.Q[2], a.Q[4]};
    Dual g[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const long st = gstride(b, d);
            Dual f[7];
#pragma unroll
            for (int m = 0; m < 7; ++m) f[m] = Dual(Qv[q][c + (m - 3) * st], a.dQ[q][c + (m - 3) * st]);
            g[q][d] = d6c(f[0], f[1], f[2], f[4], f[5], f[6], Dual(a.i60[d]));
        }
    Dual V[NVISC];
    visc_local(Dual(a.mu), Dual(a.S_ref), Dual(a.cp_pr), Dual(a.Q[0][c], a.dQ[0][c]), Dual(a.Q[1][c], a.dQ[1][c]),
               Dual(a.Q[2][c], a.dQ[2][c]), Dual(a.Q[4][c], a.dQ[3][c]), g, V);
    const long o = b.oidx(i, j, k);
#pragma unroll
    for (int n = 0; n < NVISC; ++n) a.dV[n][o] = V[n].d;
}

template <int DIR, int INTERP, int RS, bool CHAR>
__global__ void __launch_bounds__(256) k_tl_face(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
    const Box &b = a.b;
    const long st = gstride(b, DIR);
    const long g0 = b.gidx(i, j, k);
    const int var[5] = {0, P1, P2, P3, 4};
    Dual s[5][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const long g = g0 + (m - 2) * st;
#pragma unroll
        for (int n = 0; n < 5; ++n) s[n][m] = Dual(a.U[var[n]][g], a.X[var[n]][g]);
    }
    SchemeCoef<Dual> cd;
    cd.gamma = Dual(a.c.gamma);
    cd.p_ref = cd.H_ref = cd.E_ref = Dual(0.0);
#pragma unroll
    for (int q = 0; q < 6; ++q) cd.mdcd_lin[q] = Dual(a.c.mdcd_lin[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) cd.mdcd_d[q] = Dual(a.c.mdcd_d[q]);
    Dual pc[2][5];
    if constexpr (CHAR) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const Dual u[5] = {s[0][2 + side], s[1][2 + side], s[2][2 + side], s[3][2 + side], s[4][2 + side]};
            const Dual ir = Dual(1.0) / u[0];
            pc[side][0] = u[0];
            pc[side][1] = u[1] * ir;
            pc[side][2] = u[2] * ir;
            pc[side][3] = u[3] * ir;
            pc[side][4] = (cd.gamma - Dual(1.0)) *
                          (u[4] - Dual(0.5) * u[0] *
                                      (pc[side][1] * pc[side][1] + pc[side][2] * pc[side][2] + pc[side][3] * pc[side][3]));
        }
    }
    Dual F[5];
    face_flux<INTERP, RS, CHAR, XS>(cd, s, pc, F);
    const long o = b.oidx(i, j, k);
#pragma unroll
    for (int n = 0; n < 5; ++n) a.dF[DIR][var[n]][o] = F[n].d;
}

__global__ void __launch_bounds__(256) k_tl_assemble(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const Box &b = a.b;
    double r[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        long fo[6], vo[7];
#pragma unroll
        for (int m = 0; m < 6; ++m) fo[m] = owrap(b, i, j, k, d, m - 3);
#pragma unroll
        for (int m = 0; m < 7; ++m) vo[m] = owrap(b, i, j, k, d, m - 3);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) s += kW[m] * a.dF[d][n][fo[m]];
            r[n] -= s * a.ih[d];
        }
#pragma unroll
        for (int n = 1; n < 5; ++n) {
            const double *V = a.dV[kSlot[d][n - 1]];
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 7; ++m) s += kD[m] * V[vo[m]];
            r[n] += s * a.i60[d];
        }
    }
    const long g = b.gidx(i, j, k);
#pragma unroll
    for (int n = 0; n < 5; ++n) a.out[n][g] = r[n];
}

// ------------------------------------------------------------ adjoint
template <int DIR, int INTERP, int RS>
__global__ void __launch_bounds__(256) k_adj_face(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    const Box &b = a.b;
    const long st = gstride(b, DIR);
    const long g0 = b.gidx(i, j, k);  // face g0 + 1/2 along DIR
    const int var[5] = {0, P1, P2, P3, 4};
    // face adjoint: Fb[j] = -(1/h) sum_m kW[m] lambda[j + 3 - m]
    double Fb[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) s += kW[m] * a.X[var[n]][g0 + (3 - m) * st];
        Fb[n] = -s * a.ih[DIR];
    }
    double s[5][6], sb[5][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const long g = g0 + (m - 2) * st;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            s[n][m] = a.U[var[n]][g];
            sb[n][m] = 0.0;
        }
    }
    face_vjp<INTERP, RS>(a.c, s, Fb, sb);
    const long o = b.oidx(i, j, k);
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int n = 0; n < 5; ++n) a.C[m * 5 + var[n]][o] = sb[n][m];
}

template <int DIR>
__global__ void __launch_bounds__(256) k_adj_gather(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const Box &b = a.b;
    const long g = b.gidx(i, j, k);
    double acc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[n] = a.out[n][g];
    // cell p is stencil slot m of face p + 2 - m
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const long o = owrap(b, i, j, k, DIR, 2 - m);
#pragma unroll
        for (int n = 0; n < 5; ++n) acc[n] += a.C[m * 5 + n][o];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) a.out[n][g] = acc[n];
}

__global__ void __launch_bounds__(256) k_adj_visc_a(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const Box &b = a.b;
    const long c = b.gidx(i, j, k);
    const double *Qv[4] = {a.Q[0], a.Q[1], a.Q[2], a.Q[4]};
    double g[4][3], lg[4][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long st = gstride(b, d);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double *f = Qv[q];
            g[q][d] = d6c(f[c - 3 * st], f[c - 2 * st], f[c - st], f[c + st], f[c + 2 * st], f[c + 3 * st], a.i60[d]);
            const double *l = a.X[q + 1];  // lambda of rho u, rho v, rho w, E
            lg[q][d] = d6c(l[c - 3 * st], l[c - 2 * st], l[c - st], l[c + st], l[c + 2 * st], l[c + 3 * st], a.i60[d]);
        }
    }
    // transposed divergence (D6 is antisymmetric on the periodic box)
    double Vb[NVISC];
    Vb[VXX] = -lg[0][0];
    Vb[VYY] = -lg[1][1];
    Vb[VZZ] = -lg[2][2];
    Vb[VXY] = -(lg[0][1] + lg[1][0]);
    Vb[VXZ] = -(lg[0][2] + lg[2][0]);
    Vb[VYZ] = -(lg[1][2] + lg[2][1]);
    Vb[VEX] = -lg[3][0];
    Vb[VEY] = -lg[3][1];
    Vb[VEZ] = -lg[3][2];
    double gb[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, qb[4] = {0, 0, 0, 0};
    visc_local_vjp(a.mu, a.S_ref, a.cp_pr, a.Q[0][c], a.Q[1][c], a.Q[2][c], a.Q[4][c], g, Vb, gb, qb);
    const long o = b.oidx(i, j, k);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int d = 0; d < 3; ++d) a.Gb[3 * q + d][o] = gb[q][d];
        a.qb[q][o] = qb[q];
    }
}

__global__ void __launch_bounds__(256) k_adj_visc_b(AdjArgs a) {
    int i, j, k;
    if (!cell_of(a.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const Box &b = a.b;
    const long o = b.oidx(i, j, k);
    double qb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) qb[q] = a.qb[q][o];
    // transposed gradient stencil: qb -= D6_d(Gb[q][d])
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        long vo[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) vo[m] = owrap(b, i, j, k, d, m - 3);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double *G = a.Gb[3 * q + d];
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 7; ++m) s += kD[m] * G[vo[m]];
            qb[q] -= s * a.i60[d];
        }
    }
    const long g = b.gidx(i, j, k);
    double U[5], Ub[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        U[n] = a.U[n][g];
        Ub[n] = a.out[n][g];
    }
    c2p_vjp(a.gamma, a.Rgas, U, qb, Ub);
#pragma unroll
    for (int n = 0; n < 5; ++n) a.out[n][g] = Ub[n];
}

// out = alpha A + beta B + gamma P on the interior cells; a null operand
// counts as zero, and out may alias any operand.
__global__ void __launch_bounds__(256) k_lincomb(LinComb l) {
    int i, j, k;
    if (!cell_of(l.b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long g = l.b.gidx(i, j, k);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        double r = 0.0;
        if (l.A[n]) r += l.alpha * l.A[n][g];
        if (l.B[n]) r += l.beta * l.B[n][g];
        if (l.P[n]) r += l.gamma * l.P[n][g];
        l.out[n][g] = r;
    }
}

// ------------------------------------------------------------ launchers
namespace {
dim3 cgrid(const Box &b) { return dim3((unsigned)(((long)b.nx * b.ny * b.nz + 255) / 256)); }

template <int DIR, int INTERP, int RS>
void tl_face_c(const AdjArgs &a, bool chr, hipStream_t s) {
    if (chr) hipLaunchKernelGGL((k_tl_face<DIR, INTERP, RS, true>), cgrid(a.b), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_tl_face<DIR, INTERP, RS, false>), cgrid(a.b), dim3(256), 0, s, a);
}

template <int DIR, int INTERP>
void tl_face_r(const AdjArgs &a, int rs, bool chr, hipStream_t s) {
    switch (rs) {
        case R_ROE: tl_face_c<DIR, INTERP, R_ROE>(a, chr, s); break;
        case R_RUSANOV: tl_face_c<DIR, INTERP, R_RUSANOV>(a, chr, s); break;
        case R_HLLC: tl_face_c<DIR, INTERP, R_HLLC>(a, chr, s); break;
        case R_HLL: tl_face_c<DIR, INTERP, R_HLL>(a, chr, s); break;
        case R_HLLC_P: tl_face_c<DIR, INTERP, R_HLLC_P>(a, chr, s); break;
        default: tl_face_c<DIR, INTERP, R_AUSM>(a, chr, s); break;
    }
}

template <int DIR>
void tl_face_i(const AdjArgs &a, int interp, int rs, bool chr, hipStream_t s) {
    switch (interp) {
        case I_WENO5: tl_face_r<DIR, I_WENO5>(a, rs, chr, s); break;
        case I_MDCD_LINEAR: tl_face_r<DIR, I_MDCD_LINEAR>(a, rs, chr, s); break;
        case I_MDCD_HYBRID: tl_face_r<DIR, I_MDCD_HYBRID>(a, rs, chr, s); break;
        default: tl_face_r<DIR, I_WENO5_JS>(a, rs, chr, s); break;
    }
}

template <int DIR>
void adj_face_d(const AdjArgs &a, int interp, int rs, hipStream_t s) {
    const bool js = interp == I_WENO5_JS, rus = rs == R_RUSANOV;
    if (js && rus) hipLaunchKernelGGL((k_adj_face<DIR, I_WENO5_JS, R_RUSANOV>), cgrid(a.b), dim3(256), 0, s, a);
    else if (js) hipLaunchKernelGGL((k_adj_face<DIR, I_WENO5_JS, R_HLLC>), cgrid(a.b), dim3(256), 0, s, a);
    else if (rus) hipLaunchKernelGGL((k_adj_face<DIR, I_WENO5, R_RUSANOV>), cgrid(a.b), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_adj_face<DIR, I_WENO5, R_HLLC>), cgrid(a.b), dim3(256), 0, s, a);
    hipLaunchKernelGGL((k_adj_gather<DIR>), cgrid(a.b), dim3(256), 0, s, a);
}
}  // namespace

bool adjoint_supported(int interp, int rs, bool chr) {
    return !chr && (interp == I_WENO5 || interp == I_WENO5_JS) && (rs == R_HLLC || rs == R_RUSANOV);
}

void launch_tl_prim(const AdjArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_tl_prim, cgrid(a.b), dim3(256), 0, s, a);
}

void launch_tl_rhs(const AdjArgs &a, int interp, int rs, bool chr, hipStream_t s) {
    hipLaunchKernelGGL(k_tl_visc, cgrid(a.b), dim3(256), 0, s, a);
    tl_face_i<0>(a, interp, rs, chr, s);
    tl_face_i<1>(a, interp, rs, chr, s);
    tl_face_i<2>(a, interp, rs, chr, s);
    hipLaunchKernelGGL(k_tl_assemble, cgrid(a.b), dim3(256), 0, s, a);
}

void launch_adj_rhs(const AdjArgs &a, int interp, int rs, hipStream_t s) {
    // out is zeroed by the caller; the three directions accumulate in turn
    adj_face_d<0>(a, interp, rs, s);
    adj_face_d<1>(a, interp, rs, s);
    adj_face_d<2>(a, interp, rs, s);
    hipLaunchKernelGGL(k_adj_visc_a, cgrid(a.b), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_adj_visc_b, cgrid(a.b), dim3(256), 0, s, a);
}

void launch_lincomb(const LinComb &l, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb, cgrid(l.b), dim3(256), 0, s, l);
}

}  // namespace ns3d
