// ns3d_bounded.hip -- kernels of the non-periodic box (ns3d_bounded.h):
// ghost exchange and physical boundary fill (boundary.cpp:14-205), ghost
// faces (boundary.cpp:207-300), gradients and viscous fluxes with the
// near-face orders of choose_scheme (ViscousFlux.cpp:15-208), RHS assembly
// (InviscousFlux.cpp:1027-1056, ViscousFlux.cpp:210-255) and the all-cell
// Shu-Osher stage (range_kutta.cpp:46-92).
#include "ns3d_bounded.h"

namespace ns3d {

namespace {

constexpr int BT = 256;

__device__ __forceinline__ unsigned blocks_for(long n) { return (unsigned)((n + BT - 1) / BT); }

// Interior extent of an axis and the ghosted strides along it.
__device__ __forceinline__ int ext(const Box &b, int a) { return a == 0 ? b.nx : (a == 1 ? b.ny : b.nz); }
__device__ __forceinline__ int sext(const Box &b, int a) { return a == 0 ? b.sx : (a == 1 ? b.sy : b.sz); }

// ------------------------------------------------------------------ exchange
// The one-rank face-ghost exchange of rho, u, v, w, p
// (exchange_halos_physical, field_structures.h:936-994): the three layers of
// each face over the interior tangential range; a periodic axis takes the
// periodic image, a non-periodic one the zero its MPI_PROC_NULL receive
// leaves.  blockIdx.y = axis.
__global__ void __launch_bounds__(BT) k_bnd_exchange(BndArgs a) {
    const Box &b = a.b;
    const int ax = blockIdx.y;
    const int n = ext(b, ax);
    const int t1 = ax == 0 ? b.ny : b.nx, t2 = ax == 2 ? b.ny : b.nz;
    const long m = (long)t1 * t2 * 2 * NG;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int g = (int)(id % (2 * NG));  // 0..2: low side, 3..5: high side
    const long r = id / (2 * NG);
    const int p1 = (int)(r % t1), p2 = (int)(r / t1);
    const int gd = g < NG ? g - NG : n + (g - NG);  // interior coordinate of the ghost
    const int gs = g < NG ? gd + n : gd - n;        // its periodic image
    int di[3], si[3];
    di[ax] = gd;
    si[ax] = gs;
    const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
    di[o1] = si[o1] = p1;
    di[o2] = si[o2] = p2;
    const long d = b.gidx(di[0], di[1], di[2]);
    const long s = b.gidx(si[0], si[1], si[2]);
    const bool per = a.per[ax] != 0;
#pragma unroll
    for (int f = 0; f < 5; ++f) a.P[f][d] = per ? a.P[f][s] : 0.0;
}

// ------------------------------------------------------------------ boundary fill
// One physical face over its full planes (edges and corners included), the
// three ghost layers: Wall / Symmetry mirror the cell about the face and
// negate the normal velocity (apply_wall_bc, boundary.cpp:58-141), Outflow
// copies the nearest interior cell (apply_outflow_bc, 144-176), Inflow sets
// rho = 1, u = 1, v = w = 0, p = 1 (apply_inflow_bc, 179-198; its energy is
// replaced by the primitiveToConserved that follows every fill).
__global__ void __launch_bounds__(BT) k_bnd_fill(BndArgs a) {
    const Box &b = a.b;
    const int ax = a.face / 2;
    const bool hi = (a.face & 1) != 0;
    const int s1 = ax == 0 ? b.sy : b.sx, s2 = ax == 2 ? b.sy : b.sz;
    const long m = (long)s1 * s2 * NG;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int l = (int)(id % NG);
    const long r = id / NG;
    const int q1 = (int)(r % s1), q2 = (int)(r / s1);
    const int sa = sext(b, ax);
    const int g = hi ? sa - NG + l : l;  // ghosted coordinate of the ghost along the axis
    int c[3];
    const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
    c[ax] = g;
    c[o1] = q1;
    c[o2] = q2;
    auto gi = [&](const int (&x)[3]) { return ((long)x[2] * b.sy + x[1]) * b.sx + x[0]; };
    const long d = gi(c);
    if (a.type == BC_INFLOW) {
        a.P[0][d] = 1.0;
        a.P[1][d] = 1.0;
        a.P[2][d] = 0.0;
        a.P[3][d] = 0.0;
        a.P[4][d] = 1.0;
        return;
    }
    int m3[3] = {c[0], c[1], c[2]};
    if (a.type == BC_OUTFLOW) m3[ax] = hi ? sa - NG - 1 : NG;
    else m3[ax] = hi ? 2 * (sa - NG) - 1 - g : 2 * NG - 1 - g;
    const long s = gi(m3);
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const double v = a.P[f][s];
        a.P[f][d] = (f == 1 + ax && (a.type == BC_WALL || a.type == BC_SYMMETRY)) ? -v : v;
    }
}

// ------------------------------------------------------------------ ghost faces
// The ng-1 = 2 ghost faces beyond each boundary face of one direction, on the
// interior lines: the periodic images (exchange_halos_halfnode_flux,
// field_structures.h:773-874, 998-1056), or on a physical face the zero of
// the MPI_PROC_NULL receive, replaced on an Outflow face by the boundary face
// itself (apply_outflow_bc_halfnode_flux, boundary.cpp:246-300).
__global__ void __launch_bounds__(BT) k_ghost_faces(GhostFaceArgs a) {
    const Box &b = a.b;
    const int ax = a.dir;
    const int n = ext(b, ax);
    const int t1 = ax == 0 ? b.ny : b.nx, t2 = ax == 2 ? b.ny : b.nz;
    const long m = (long)t1 * t2 * 2 * (NG - 1);
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int g = (int)(id % (2 * (NG - 1)));
    const long r = id / (2 * (NG - 1));
    const int p1 = (int)(r % t1) + NG, p2 = (int)(r / t1) + NG;
    const bool hi = g >= NG - 1;
    const int q = hi ? NG + n + (g - (NG - 1)) : g;  // ghost face index
    int src;
    if (a.per) src = hi ? q - n : q + n;
    else if ((hi ? a.type_hi : a.type_lo) == BC_OUTFLOW) src = hi ? NG + n - 1 : NG - 1;
    else src = -1;
    const long d = face_ref_index(b, ax, q, p1, p2);
    const long s = src >= 0 ? face_ref_index(b, ax, src, p1, p2) : 0;
#pragma unroll
    for (int f = 0; f < 5; ++f) a.F[f][d] = src >= 0 ? a.F[f][s] : 0.0;
}

// ------------------------------------------------------------------ viscous
// choose_scheme (ViscousFlux.cpp:33-50): 6 on a periodic axis; next to a
// physical face 1 / -1 (2nd-order one-sided, forward / backward) on the
// first / last interior cell, 2 and 4 (central) one and two cells in.
__device__ __forceinline__ int order_at(int c, int n, int per) {
    if (per) return 6;
    if (c == 0) return 1;
    if (c == n - 1) return -1;
    const int d = min(c, n - 1 - c);
    return d == 1 ? 2 : (d == 2 ? 4 : 6);
}

// The derivative stencils of ViscousFlux.cpp:15-29 on f(o) = the value o
// cells along the axis; only the points the order uses are read.
template <class Fn>
__device__ __forceinline__ double dstencil(Fn f, int order, double h) {
    if (order == 6) return (f(3) - 9.0 * f(2) + 45.0 * f(1) - 45.0 * f(-1) + 9.0 * f(-2) - f(-3)) / (60.0 * h);
    if (order == 4) return (-f(2) + 8.0 * f(1) - 8.0 * f(-1) + f(-2)) / (12.0 * h);
    if (order == 2) return (f(1) - f(-1)) / (2.0 * h);
    return (-3.0 * f(0) + 4.0 * f(order) - f(2 * order)) / (2.0 * h);
}

__global__ void __launch_bounds__(BT) k_bnd_visc(BndViscArgs a) {
    const Box &b = a.b;
    const long m = (long)b.nx * b.ny * b.nz;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int i = (int)(id % b.nx), j = (int)((id / b.nx) % b.ny), k = (int)(id / ((long)b.nx * b.ny));
    const long c = b.gidx(i, j, k);
    const long st[3] = {1, (long)b.sx, (long)b.sx * b.sy};
    const int ord[3] = {order_at(i, b.nx, a.per[0]), order_at(j, b.ny, a.per[1]), order_at(k, b.nz, a.per[2])};
    double g[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double *f = a.Q[q];
            const long s = st[d];
            g[q][d] = dstencil([&](int o) { return f[c + o * s]; }, ord[d], a.h[d]);
        }
    if (a.G[0]) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) a.G[3 * q + d][id] = g[q][d];
    }
    if (!a.V[0]) return;
    double V[NVISC];
    visc_local(a.mu, a.S_ref, a.cp_pr, a.Q[0][c], a.Q[1][c], a.Q[2][c], a.Q[3][c], g, V);
#pragma unroll
    for (int n = 0; n < NVISC; ++n) a.V[n][id] = V[n];
}

// ------------------------------------------------------------------ RHS
// diff_6th_central_half (InviscousFlux.cpp:990-992) on faces i-5/2..i+5/2
__device__ __forceinline__ double d6half(double f0, double f1, double f2, double f3, double f4, double f5, double h) {
    return ((f3 - f2) * (75.0 / 64.0) + (f4 - f1) * (-25.0 / 384.0) + (f5 - f0) * (3.0 / 640.0)) / h;
}

__global__ void __launch_bounds__(BT) k_bnd_rhs(BndRhsArgs a) {
    const Box &b = a.b;
    const long m = (long)b.nx * b.ny * b.nz;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int i = (int)(id % b.nx), j = (int)((id / b.nx) % b.ny), k = (int)(id / ((long)b.nx * b.ny));
    const long c = b.gidx(i, j, k);
    const int I = i + NG, J = j + NG, K = k + NG;  // ghosted coordinates
    const int ord[3] = {order_at(i, b.nx, a.per[0]), order_at(j, b.ny, a.per[1]), order_at(k, b.nz, a.per[2])};
    // the viscous flux of cell (i, j, k) shifted by o along axis d (periodic
    // axes wrap; along a physical one the orders stay inside the box)
    auto vat = [&](const double *v, int d, int o) {
        int x = i, y = j, z = k;
        if (d == 0) x = wrap(i + o, b.nx);
        else if (d == 1) y = wrap(j + o, b.ny);
        else z = wrap(k + o, b.nz);
        return v[b.oidx(x, y, z)];
    };
    // viscous flux slots per (direction, variable): F = x, H = y, G = z
    // (ViscousFlux.cpp:187-206)
    constexpr int VS[3][5] = {{-1, VXX, VXY, VXZ, VEX}, {-1, VXY, VYY, VYZ, VEY}, {-1, VXZ, VYZ, VZZ, VEZ}};
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        double r = (a.mode & ASM_ACCUMULATE) ? a.RHS[n][c] : 0.0;
        if (a.mode & ASM_INVISCID) {
            const double *X = a.F[0][n], *Y = a.F[1][n], *Z = a.F[2][n];
            r -= d6half(X[face_ref_index(b, 0, I - 3, J, K)], X[face_ref_index(b, 0, I - 2, J, K)],
                        X[face_ref_index(b, 0, I - 1, J, K)], X[face_ref_index(b, 0, I, J, K)],
                        X[face_ref_index(b, 0, I + 1, J, K)], X[face_ref_index(b, 0, I + 2, J, K)], a.h[0]);
            r -= d6half(Y[face_ref_index(b, 1, J - 3, I, K)], Y[face_ref_index(b, 1, J - 2, I, K)],
                        Y[face_ref_index(b, 1, J - 1, I, K)], Y[face_ref_index(b, 1, J, I, K)],
                        Y[face_ref_index(b, 1, J + 1, I, K)], Y[face_ref_index(b, 1, J + 2, I, K)], a.h[1]);
            r -= d6half(Z[face_ref_index(b, 2, K - 3, I, J)], Z[face_ref_index(b, 2, K - 2, I, J)],
                        Z[face_ref_index(b, 2, K - 1, I, J)], Z[face_ref_index(b, 2, K, I, J)],
                        Z[face_ref_index(b, 2, K + 1, I, J)], Z[face_ref_index(b, 2, K + 2, I, J)], a.h[2]);
        }
        if ((a.mode & ASM_VISCOUS) && n > 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double *v = a.V[VS[d][n]];
                r += dstencil([&](int o) { return vat(v, d, o); }, ord[d], a.h[d]);
            }
        }
        a.RHS[n][c] = r;
    }
}

// ------------------------------------------------------------------ stage
// One Shu-Osher stage on every cell of the ghosted box (range_kutta.cpp:46-92):
// the RHS is zero outside the interior there.
__global__ void __launch_bounds__(BT) k_rk_all(RkAllArgs a) {
    const Box &b = a.b;
    const long m = (long)b.sx * b.sy * b.sz;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= m) return;
    const int I = (int)(id % b.sx), J = (int)((id / b.sx) % b.sy), K = (int)(id / ((long)b.sx * b.sy));
    const bool inner = I >= NG && I < NG + b.nx && J >= NG && J < NG + b.ny && K >= NG && K < NG + b.nz;
    const double dt = a.dt ? *a.dt : a.dtv;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        const double r = inner ? a.R[n][id] : 0.0;
        const double u0 = a.U0[n][id];
        double u;
        if (a.stage == 1) u = u0 + dt * r;
        else if (a.stage == 2) u = 0.75 * u0 + 0.25 * (a.U[n][id] + dt * r);
        else u = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (a.U[n][id] + dt * r);
        a.U[n][id] = u;
    }
}

}  // namespace

void launch_bnd_exchange(const BndArgs &a, hipStream_t s) {
    const Box &b = a.b;
    const long m = (long)std::max(std::max(b.ny * b.nz, b.nx * b.nz), b.nx * b.ny) * 2 * NG;
    hipLaunchKernelGGL(k_bnd_exchange, dim3((unsigned)((m + BT - 1) / BT), 3), dim3(BT), 0, s, a);
}

void launch_bnd_fill(const BndArgs &a, hipStream_t s) {
    const Box &b = a.b;
    const int ax = a.face / 2;
    const long m = (long)(ax == 0 ? b.sy : b.sx) * (ax == 2 ? b.sy : b.sz) * NG;
    hipLaunchKernelGGL(k_bnd_fill, dim3((unsigned)((m + BT - 1) / BT)), dim3(BT), 0, s, a);
}

void launch_ghost_faces(const GhostFaceArgs &a, hipStream_t s) {
    const Box &b = a.b;
    const long m = (long)(a.dir == 0 ? b.ny : b.nx) * (a.dir == 2 ? b.ny : b.nz) * 2 * (NG - 1);
    hipLaunchKernelGGL(k_ghost_faces, dim3((unsigned)((m + BT - 1) / BT)), dim3(BT), 0, s, a);
}

void launch_bnd_visc(const BndViscArgs &a, hipStream_t s) {
    const long m = (long)a.b.nx * a.b.ny * a.b.nz;
    hipLaunchKernelGGL(k_bnd_visc, dim3((unsigned)((m + BT - 1) / BT)), dim3(BT), 0, s, a);
}

void launch_bnd_rhs(const BndRhsArgs &a, hipStream_t s) {
    const long m = (long)a.b.nx * a.b.ny * a.b.nz;
    hipLaunchKernelGGL(k_bnd_rhs, dim3((unsigned)((m + BT - 1) / BT)), dim3(BT), 0, s, a);
}

void launch_rk_all(const RkAllArgs &a, hipStream_t s) {
    const long m = (long)a.b.sx * a.b.sy * a.b.sz;
    hipLaunchKernelGGL(k_rk_all, dim3((unsigned)((m + BT - 1) / BT)), dim3(BT), 0, s, a);
}

}  // namespace ns3d
