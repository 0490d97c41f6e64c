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
//   k_adj_lamq_march  z-marching plane tiles of lambda and of u, v, w, T:
//                 face adjoints of the three directions (the transposed
//                 half-node difference), viscous flux adjoints (the
//                 transposed divergence), then the gradients and the reverse
//                 of stresses / heat flux -> gradient and (u, v, w, T)
//                 adjoints;
//   per direction, marching along the direction (lanes across it):
//   k_adj_riemann both reconstructions of each cell (shared smoothness
//                 indicators) and the reverse of the Riemann flux of each
//                 face from the face adjoint -> left / right state adjoints;
//   k_adj_weno    reverse of each cell's two reconstructions, accumulated in
//                 a register ring of its five stencil cells and added to the
//                 result once complete; the last direction also adds the
//                 transposed gradient stencil and the reverse of the
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
// Column marching along DIR: lanes run along LA (x, or y for DIR = x), block
// rows along OA, segments of a.seg cells along DIR in grid z.

template <int DIR>
struct March {
    static constexpr int LA = DIR == 0 ? 1 : 0;
    static constexpr int OA = DIR == 2 ? 1 : 2;
    int N, c0, c1;
    long gbase, gst, obase, ost;
    bool ok;
    __device__ March(const Box &b, int seg) {
        const int n[3] = {b.nx, b.ny, b.nz};
        int ijk[3];
        ijk[LA] = blockIdx.x * 64 + threadIdx.x;
        ijk[OA] = blockIdx.y * 4 + threadIdx.y;
        ijk[DIR] = 0;
        ok = ijk[LA] < n[LA] && ijk[OA] < n[OA];
        if (!ok) ijk[LA] = ijk[OA] = 0;
        N = n[DIR];
        c0 = blockIdx.z * seg;
        c1 = min(c0 + seg, N);
        gbase = b.gidx(ijk[0], ijk[1], ijk[2]);
        obase = b.oidx(ijk[0], ijk[1], ijk[2]);
        gst = DIR == 0 ? 1L : (DIR == 1 ? (long)b.sx : (long)b.sx * b.sy);
        ost = DIR == 0 ? 1L : (DIR == 1 ? (long)b.nx : (long)b.nx * b.ny);
    }
    __device__ long g(int c) const { return gbase + (long)wrap(c, N) * gst; }
    __device__ long o(int c) const { return obase + (long)wrap(c, N) * ost; }
};

// Launch of kernel K marching along DIR, with K's segment length.
template <int DIR, auto K>
void launch_adj_march(AdjArgs a, hipStream_t s) {
    const int n[3] = {a.b.nx, a.b.ny, a.b.nz};
    constexpr int LA = March<DIR>::LA, OA = March<DIR>::OA;
    const int gx = (n[LA] + 63) / 64, gy = (n[OA] + 3) / 4;
    a.seg = march_seg<K>((long)gx * gy, n[DIR], ADJ_SEG_MIN);
    hipLaunchKernelGGL(K, dim3(gx, gy, (n[DIR] + a.seg - 1) / a.seg), dim3(64, 4), 0, s, a);
}

// Per face of DIR (marching): both reconstructions of every cell (shared
// smoothness indicators), then the reverse of the Riemann flux of face
// c-1/2 from the left state of cell c-1 and the right state of cell c, with
// the face adjoint FB[DIR] -> left / right state adjoints SB[0..4], SB[5..9].
template <int DIR, int INTERP, int RS>
__global__ void __launch_bounds__(256) k_adj_riemann(AdjArgs a) {
    const March<DIR> M(a.b, a.seg);
    if (!M.ok) return;
    constexpr bool JS = INTERP == I_WENO5_JS;
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    const int var[5] = {0, P1, P2, P3, 4};
    double w[5][5];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const long g = M.g(M.c0 - 2 + m);
#pragma unroll
        for (int n = 0; n < 5; ++n) w[n][m] = a.U[var[n]][g];
    }
    double ULp[5];
    for (int c = M.c0; c <= M.c1; ++c) {
        {
            const long g = M.g(c + 2);
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][4] = a.U[var[n]][g];
        }
        double UL[5], UR[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) weno_pair<JS>(w[n], UL[n], UR[n]);
        if (c > M.c0) {  // face (c-1) + 1/2
            const long of = M.o(c - 1);
            double Fb[5], ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int n = 0; n < 5; ++n) Fb[n] = a.FB[DIR][var[n]][of];
            riemann_vjp<RS, DIR == 0 ? 1 : (DIR == 1 ? 3 : 2)>(a.c, ULp, UR, Fb, ULb, URb);
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                a.SB[var[n]][of] = ULb[n];
                a.SB[5 + var[n]][of] = URb[n];
            }
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            ULp[n] = UL[n];
#pragma unroll
            for (int m = 0; m < 4; ++m) w[n][m] = w[n][m + 1];
        }
    }
}

// Per cell along DIR (marching): reverse of the cell's two reconstructions
// (shared indicators), accumulated in a ring of the five stencil cells; a
// cell is complete once the reconstructions of its four neighbours are in,
// and is added to out (FIRST: written).  The ring lives in LDS, one private
// column of 25 doubles per thread (slot of cell x: x mod 5, uniform over the
// block), which keeps the registers to the stencil window and the reverse.
template <int DIR, bool JS, bool FIRST>
__global__ void __launch_bounds__(256) k_adj_weno(AdjArgs a) {
    __shared__ double ring[5][5][256];  // [cell mod 5][var][thread]
    const March<DIR> M(a.b, a.seg);
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if (!M.ok) return;  // no barrier below: each thread owns its ring column
    double w[5][5];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const long g = M.g(M.c0 - 4 + m);
#pragma unroll
        for (int n = 0; n < 5; ++n) w[n][m] = a.U[n][g];
    }
#pragma unroll
    for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int n = 0; n < 5; ++n) ring[m][n][tid] = 0.0;
    // cell c's reconstructions reach cells c-2..c+2
    for (int c = M.c0 - 2; c <= M.c1 + 1; ++c) {
        {
            const long g = M.g(c + 2);
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][4] = a.U[n][g];
        }
        const long oc = M.o(c), om = M.o(c - 1);
        const int base = c + 5;  // non-negative for the mod (c >= c0 - 2)
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            double fb[5] = {0, 0, 0, 0, 0};
            weno_pair_vjp<JS>(w[n], a.SB[n][oc], a.SB[5 + n][om], fb);
            // only cells c0..c1-1 are accumulated: the slots of cells below
            // c0 are those of c0 + 1.. (uniform condition over the block)
#pragma unroll
            for (int m = 0; m < 5; ++m)
                if (c - 2 + m >= M.c0 && c - 2 + m < M.c1) ring[(base - 2 + m) % 5][n][tid] += fb[m];
        }
        const int done = c - 2;  // all of its five contributors are in
        if (done >= M.c0) {
            const int slot = (base - 2) % 5;
            const long g = M.g(done);
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                a.out[n][g] = (FIRST ? 0.0 : a.out[n][g]) + ring[slot][n][tid];
                ring[slot][n][tid] = 0.0;  // reused by cell done + 5
            }
        }
        if (done + 1 >= M.c1) break;  // cells c0..c1-1 are out
#pragma unroll
        for (int n = 0; n < 5; ++n)
#pragma unroll
            for (int m = 0; m < 4; ++m) w[n][m] = w[n][m + 1];
    }
}

// The y / z direction of the MDCD adjoint (marching): per face f + 1/2 both
// states from the six cells f-2..f+3 (the right one on the mirrored
// stencil), the reverse of the Riemann flux with the face adjoint FB[DIR],
// then the reverse of the two reconstructions, accumulated in a ring of the
// six stencil cells held in LDS (one private column per thread, slot of cell
// x: x mod 6).  A cell is complete once the six faces whose stencils hold it
// are in, and is added to out (FIRST: written).
template <int DIR, int INTERP, int RS, bool FIRST>
__global__ void __launch_bounds__(256) k_adj_mdcd(AdjArgs a) {
    __shared__ double ring[6][5][256];  // [cell mod 6][rotated var][thread]
    const March<DIR> M(a.b, a.seg);
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if (!M.ok) return;  // no barrier below: each thread owns its ring column
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
    const int var[5] = {0, P1, P2, P3, 4};
    // faces c0-3 .. c1+1 hold the cells c0 .. c1-1; face f reads cells f-2..f+3
    double w[5][6];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const long g = M.g(M.c0 - 5 + m);
#pragma unroll
        for (int n = 0; n < 5; ++n) w[n][m] = a.U[var[n]][g];
    }
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int n = 0; n < 5; ++n) ring[m][n][tid] = 0.0;
    for (int f = M.c0 - 3; f <= M.c1 + 1; ++f) {
        {
            const long g = M.g(f + 3);
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][5] = a.U[var[n]][g];
        }
        double UL[5], UR[5], Fb[5], ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
        const long of = M.o(f);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            interp_lr<INTERP>(a.c, w[n], UL[n], UR[n]);
            Fb[n] = a.FB[DIR][var[n]][of];
        }
        riemann_vjp<RS, XS>(a.c, UL, UR, Fb, ULb, URb);
        const int base = f + 6;  // non-negative for the mod (f >= c0 - 3)
#pragma unroll 1
        for (int n = 0; n < 5; ++n) {
            double fb[6] = {0, 0, 0, 0, 0, 0}, gb[6] = {0, 0, 0, 0, 0, 0};
            const double fr[6] = {w[n][5], w[n][4], w[n][3], w[n][2], w[n][1], w[n][0]};
            mdcd_vjp<INTERP>(a.c, w[n], ULb[n], fb);
            mdcd_vjp<INTERP>(a.c, fr, URb[n], gb);
            // only cells c0..c1-1 are accumulated: the slots of cells below c0
            // are those of c0 + 1.. (uniform condition over the block)
#pragma unroll
            for (int m = 0; m < 6; ++m)
                if (f - 2 + m >= M.c0 && f - 2 + m < M.c1) ring[(base - 2 + m) % 6][n][tid] += fb[m] + gb[5 - m];
        }
        const int done = f - 2;  // its faces done-3 .. done+2 are in
        if (done >= M.c0) {
            const int slot = (base - 2) % 6;
            const long g = M.g(done);
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                a.out[var[n]][g] = (FIRST ? 0.0 : a.out[var[n]][g]) + ring[slot][n][tid];
                ring[slot][n][tid] = 0.0;  // reused by cell done + 6
            }
        }
        if (done + 1 >= M.c1) break;  // cells c0..c1-1 are out
#pragma unroll
        for (int n = 0; n < 5; ++n)
#pragma unroll
            for (int m = 0; m < 5; ++m) w[n][m] = w[n][m + 1];
    }
}

// alpha A + beta B + gamma p at ghosted cell g, component n (null operands
// count as zero; p enters where l.P is set), in this order.
__device__ __forceinline__ double lin_at(const LinComb &l, int n, long g, double p) {
    double r = 0.0;
    if (l.A[n]) r += l.alpha * l.A[n][g];
    if (l.B[n]) r += l.beta * l.B[n][g];
    if (l.P[n]) r += l.gamma * p;
    return r;
}

// The viscous part of the x-direction adjoint at cell (i, j, k): the
// transposed gradient stencil and the reverse of the conserved -> primitive
// map added to the cell's face part acc, and the total added to out, or
// handed to the tail's combinations.
// Uc: the cell's state (read from the kernel's LDS copy of the line).
__device__ __forceinline__ void xline_tail(const AdjArgs &a, const LinTail &tl, int i, int j, int k, double *acc,
                                           const double *Uc) {
    const Box &b = a.b;
    const long g = b.gidx(i, j, k), o = b.oidx(i, j, k);
    double qb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) qb[q] = a.qb[q][o];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        long vo[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) vo[m] = owrap(b, i, j, k, d, m - 3);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double *G = a.Gb[3 * q + d];
            double sum = 0.0;
#pragma unroll
            for (int m = 0; m < 7; ++m) sum += kD[m] * G[vo[m]];
            qb[q] -= sum * a.i60[d];
        }
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[n] += a.out[n][g];
    c2p_vjp(a.gamma, a.Rgas, Uc, qb, acc);
    if (tl.n == 0) {
#pragma unroll
        for (int n = 0; n < 5; ++n) a.out[n][g] = acc[n];
        return;
    }
    // the stage combinations (only this cell of their operands is read, and
    // no kernel of this RHS reads what they write)
    double v[2][5];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 5; ++n) v[m][n] = m < tl.n ? lin_at(tl.c[m], n, g, acc[n]) : 0.0;
#pragma unroll
    for (int m = 0; m < 2; ++m)
        if (m < tl.n)
#pragma unroll
            for (int n = 0; n < 5; ++n) tl.c[m].out[n][g] = v[m][n];
}

// Longest x line held whole by one block of the x-direction adjoint; longer
// lines go in segments (k_adj_xseg).
constexpr int XSEG = 256;

// The x direction for lines of up to XSEG cells, one block per x line
// (threads = cells): the line and its 2-cell periodic halo in LDS; both
// reconstructions of every cell; the reverse of the Riemann flux of every face (state adjoints
// written over the states they replace: each slot is read and written by
// one thread); per variable the reverse of the reconstructions into LDS and
// the gather of the five contributions of each cell (two contribution
// buffers, used in turn, so that one barrier per variable orders the
// writes and reads); then the viscous part (transposed gradient stencil and
// the reverse of the conserved -> primitive map), added to out; with a tail,
// the cell's total goes into the tail's combinations instead of out.
template <bool JS, int RS>
__global__ void __launch_bounds__(512) k_adj_xline(AdjArgs a, LinTail tl) {
    extern __shared__ double lds[];
    const Box &b = a.b;
    const int N = b.nx, W = N + 4;
    const int t = threadIdx.x;
    const int j = blockIdx.x % b.ny, k = blockIdx.x / b.ny;
    double *Ul = lds;           // [5][W]: cells -2..N+1
    double *A = Ul + 5 * W;     // [10][N]: left / right states, then their adjoints
    double *Cc = A + 10 * N;    // [2][5][N]: contributions of one variable, buffer v % 2
    const long row = b.gidx(0, j, k) - NG;  // ghosted index of x = -NG
    for (int x = t; x < W; x += blockDim.x) {
        const long g = row + NG + wrap(x - 2, N);
#pragma unroll
        for (int n = 0; n < 5; ++n) Ul[n * W + x] = a.U[n][g];
    }
    __syncthreads();
    const bool own = t < N;
    if (own) {
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const double f[5] = {Ul[n * W + t], Ul[n * W + t + 1], Ul[n * W + t + 2], Ul[n * W + t + 3],
                                 Ul[n * W + t + 4]};
            weno_pair<JS>(f, A[n * N + t], A[(5 + n) * N + t]);
        }
    }
    __syncthreads();
    const long o = b.oidx(0, j, k) + t;
    if (own) {  // face t + 1/2: left state of cell t, right state of cell t + 1
        const int tp = t + 1 < N ? t + 1 : 0;
        double UL[5], UR[5], Fb[5], ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            UL[n] = A[n * N + t];
            UR[n] = A[(5 + n) * N + tp];
            Fb[n] = a.FB[0][n][o];
        }
        riemann_vjp<RS, 1>(a.c, UL, UR, Fb, ULb, URb);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            A[n * N + t] = ULb[n];          // adjoint of the left state of face t + 1/2
            A[(5 + n) * N + tp] = URb[n];   // slot tp: adjoint of the right state of face tp - 1/2
        }
    }
    __syncthreads();
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        // buffer v % 2 was last read (variable v - 2) before the previous barrier
        double *C = Cc + (v & 1) * 5 * N;
        if (own) {
            const double f[5] = {Ul[v * W + t], Ul[v * W + t + 1], Ul[v * W + t + 2], Ul[v * W + t + 3],
                                 Ul[v * W + t + 4]};
            double fb[5] = {0, 0, 0, 0, 0};
            weno_pair_vjp<JS>(f, A[v * N + t], A[(5 + v) * N + t], fb);
#pragma unroll
            for (int m = 0; m < 5; ++m) C[m * N + t] = fb[m];
        }
        __syncthreads();
        if (own) {
            double sum = 0.0;
#pragma unroll
            for (int m = 0; m < 5; ++m) sum += C[m * N + wrap(t + 2 - m, N)];
            acc[v] = sum;
        }
    }
    if (!own) return;
    const double Uc[5] = {Ul[t + 2], Ul[W + t + 2], Ul[2 * W + t + 2], Ul[3 * W + t + 2], Ul[4 * W + t + 2]};
    xline_tail(a, tl, t, j, k, acc, Uc);
}

// The x direction of the MDCD adjoint, for x lines of any length: one block
// of XSEG threads per segment of S <= XSEG cells (one segment when N <= XSEG;
// cells beyond the line are periodic images).  Per face f + 1/2 of the
// faces (s0-3)+1/2 .. (s0+S+1)+1/2, whose six-cell stencils reach the
// segment: both states, the reverse of the Riemann flux; then per variable
// the reverse of the two reconstructions of every face into LDS, and the
// gather of the six contributions of each cell; then the viscous part, as in
// k_adj_xline.
__host__ __device__ inline long xmdcd_doubles(int S) { return 5L * (S + 10) + 16L * (S + 5); }

template <int INTERP, int RS>
__global__ void __launch_bounds__(XSEG) k_adj_mdcd_x(AdjArgs a, LinTail tl, int nseg, int S) {
    extern __shared__ double lds[];
    const Box &b = a.b;
    const int N = b.nx, W = S + 10, F = S + 5;
    const int t = threadIdx.x;
    const int seg = blockIdx.x % nseg;
    const long line = blockIdx.x / nseg;
    const int j = (int)(line % b.ny), k = (int)(line / b.ny);
    const int s0 = seg * S, Sl = min(S, N - s0);  // this segment: cells s0..s0+Sl-1
    double *Ul = lds;          // [5][W]: slot x = cell s0 - 5 + x
    double *A = Ul + 5 * W;    // [10][F]: slot q = face (s0 - 3 + q) + 1/2, left / right state adjoints
    double *Cc = A + 10 * F;   // [6][F]: face q's contribution to cell s0 - 5 + q + m
    const long grow = b.gidx(0, j, k);
    for (int x = t; x < Sl + 10; x += XSEG) {
        const long g = grow + wrap(s0 - 5 + x, N);
#pragma unroll
        for (int n = 0; n < 5; ++n) Ul[n * W + x] = a.U[n][g];
    }
    __syncthreads();
    for (int q = t; q < Sl + 5; q += XSEG) {  // stencil: state slots q..q+5
        const long of = b.oidx(wrap(s0 - 3 + q, N), j, k);
        double UL[5], UR[5], Fb[5], ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            interp_lr<INTERP>(a.c, Ul + n * W + q, UL[n], UR[n]);
            Fb[n] = a.FB[0][n][of];
        }
        riemann_vjp<RS, 1>(a.c, UL, UR, Fb, ULb, URb);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            A[n * F + q] = ULb[n];
            A[(5 + n) * F + q] = URb[n];
        }
    }
    __syncthreads();
    double acc[5];
#pragma unroll 1
    for (int v = 0; v < 5; ++v) {
        for (int q = t; q < Sl + 5; q += XSEG) {
            const double *fl = Ul + v * W + q;
            const double fr[6] = {fl[5], fl[4], fl[3], fl[2], fl[1], fl[0]};
            double fb[6] = {0, 0, 0, 0, 0, 0}, gb[6] = {0, 0, 0, 0, 0, 0};
            mdcd_vjp<INTERP>(a.c, fl, A[v * F + q], fb);
            mdcd_vjp<INTERP>(a.c, fr, A[(5 + v) * F + q], gb);
#pragma unroll
            for (int m = 0; m < 6; ++m) Cc[m * F + q] = fb[m] + gb[5 - m];
        }
        __syncthreads();
        if (t < Sl) {  // cell s0+t gathers output m of face s0+t+2-m (slot t+5-m)
            double sum = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) sum += Cc[m * F + t + 5 - m];
            acc[v] = sum;
        }
        __syncthreads();  // Cc is rewritten by the next variable
    }
    if (t >= Sl) return;
    const double Uc[5] = {Ul[t + 5], Ul[W + t + 5], Ul[2 * W + t + 5], Ul[3 * W + t + 5], Ul[4 * W + t + 5]};
    xline_tail(a, tl, s0 + t, j, k, acc, Uc);
}

// The x direction for lines longer than XSEG cells: one block of XSEG
// threads per segment of S <= XSEG cells, which recomputes the reach of the
// reconstructions on either side (periodic images): the state of cells
// s0-5..s0+S+4, both reconstructions of cells s0-3..s0+S+2, the Riemann
// reverse of faces (s0-3)+1/2..(s0+S+1)+1/2 and the reconstruction reverse of
// cells s0-2..s0+S+1, whose contributions reach the segment.  Per cell the
// arithmetic and its order are those of k_adj_xline.

__host__ __device__ inline long xseg_doubles(int S) { return 5L * (S + 10) + 10L * (S + 6) + 5L * (S + 4); }

template <bool JS, int RS>
__global__ void __launch_bounds__(XSEG) k_adj_xseg(AdjArgs a, LinTail tl, int nseg, int S) {
    extern __shared__ double lds[];
    const Box &b = a.b;
    const int N = b.nx, W = S + 10, R = S + 6, C = S + 4;
    const int t = threadIdx.x;
    const int seg = blockIdx.x % nseg;
    const long line = blockIdx.x / nseg;
    const int j = (int)(line % b.ny), k = (int)(line / b.ny);
    const int s0 = seg * S, Sl = min(S, N - s0);  // this segment: cells s0..s0+Sl-1
    double *Ul = lds;          // [5][W]: slot x = cell s0 - 5 + x
    double *A = Ul + 5 * W;    // [10][R]: slot r = cell s0 - 3 + r
    double *Cc = A + 10 * R;   // [5][C]: slot q = cell s0 - 2 + q
    const long grow = b.gidx(0, j, k);
    for (int x = t; x < Sl + 10; x += XSEG) {
        const long g = grow + wrap(s0 - 5 + x, N);
#pragma unroll
        for (int n = 0; n < 5; ++n) Ul[n * W + x] = a.U[n][g];
    }
    __syncthreads();
    for (int r = t; r < Sl + 6; r += XSEG) {
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const double f[5] = {Ul[n * W + r], Ul[n * W + r + 1], Ul[n * W + r + 2], Ul[n * W + r + 3],
                                 Ul[n * W + r + 4]};
            weno_pair<JS>(f, A[n * R + r], A[(5 + n) * R + r]);
        }
    }
    __syncthreads();
    // face (s0-3+q)+1/2: left state of slot q, right state of slot q + 1; each
    // slot is read and written by one face
    for (int q = t; q < Sl + 5; q += XSEG) {
        const long of = b.oidx(wrap(s0 - 3 + q, N), j, k);
        double UL[5], UR[5], Fb[5], ULb[5] = {0, 0, 0, 0, 0}, URb[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            UL[n] = A[n * R + q];
            UR[n] = A[(5 + n) * R + q + 1];
            Fb[n] = a.FB[0][n][of];
        }
        riemann_vjp<RS, 1>(a.c, UL, UR, Fb, ULb, URb);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            A[n * R + q] = ULb[n];
            A[(5 + n) * R + q + 1] = URb[n];
        }
    }
    __syncthreads();
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        for (int q = t; q < Sl + 4; q += XSEG) {  // cell s0-2+q: state slots q+1..q+5, recon slot q+1
            const double f[5] = {Ul[v * W + q + 1], Ul[v * W + q + 2], Ul[v * W + q + 3], Ul[v * W + q + 4],
                                 Ul[v * W + q + 5]};
            double fb[5] = {0, 0, 0, 0, 0};
            weno_pair_vjp<JS>(f, A[v * R + q + 1], A[(5 + v) * R + q + 1], fb);
#pragma unroll
            for (int m = 0; m < 5; ++m) Cc[m * C + q] = fb[m];
        }
        __syncthreads();
        if (t < Sl) {  // cell s0+t gathers output m of cell s0+t+2-m (slot t+4-m)
            double sum = 0.0;
#pragma unroll
            for (int m = 0; m < 5; ++m) sum += Cc[m * C + t + 4 - m];
            acc[v] = sum;
        }
        __syncthreads();  // Cc is rewritten by the next variable
    }
    if (t >= Sl) return;
    const double Uc[5] = {Ul[t + 5], Ul[W + t + 5], Ul[2 * W + t + 5], Ul[3 * W + t + 5], Ul[4 * W + t + 5]};
    xline_tail(a, tl, s0 + t, j, k, acc, Uc);
}

// The face and viscous adjoints of one adjoint RHS, as two z-marching
// kernels over 64 x 4 column tiles: each
// thread keeps planes k-3..k+3 of its own column in registers for the z
// stencils; each plane's tile with a 3-cell x/y halo is staged in LDS for the
// x and y stencils.  Cells beyond the box come from their periodic image, so
// neither lambda nor the primitives need ghost layers.

// Index of the periodic image of interior cell (x, y, z) for x, y in any
// range and z in [-3, nz + 3).
struct PlaneMap {
    Box b;
    long col, hc[2];
    int hi[2], hj[2];
    bool hv[2];
    __device__ PlaneMap(const Box &bb, int i0, int j0, int tx, int ty) : b(bb) {
        auto per = [](int x, int n) { return ((x % n) + n) % n; };
        const int tid = ty * 64 + tx;
        col = b.gidx(per(i0 + tx, b.nx), per(j0 + ty, b.ny), 0);
        // halo cells of a plane: 6 rows x 64 above/below, then 4 rows x 6 left/right
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + 256 * h;
            int x, y;
            if (e < 6 * 64) {
                const int r = e / 64;
                y = r < 3 ? r - 3 : r + 1;
                x = e % 64;
            } else {
                const int e2 = e - 6 * 64, r = e2 / 6, c6 = e2 % 6;
                y = r;
                x = c6 < 3 ? c6 - 3 : 64 + c6 - 3;
            }
            hi[h] = x;
            hj[h] = y;
            hv[h] = e < 6 * 64 + 6 * 4;
            hc[h] = b.gidx(per(i0 + x, b.nx), per(j0 + y, b.ny), 0);
        }
    }
    __device__ long plane(int k) const { return (long)wrap(k, b.nz) * ((long)b.sx * b.sy); }
};

// z-marching plane tiles of lambda and of (u, v, w, T), one plane per
// iteration.  Lambda part: the face adjoints FB = -(1/h) sum_m kW[m]
// lambda[c + 3 - m] of the three directions (the transposed half-node
// difference) and the viscous-flux adjoints Vb = -(transposed divergence of
// lambda).  Primitive part: the gradients of u, v, w, T (6th-order central)
// and the reverse of the local viscous-flux map with Vb -> gradient adjoints
// Gb and (u, v, w, T) adjoints qb.  Vb stays in registers between the two
// parts; two waves per SIMD (238 VGPRs) keep both register rings unspilled.
__global__ void __launch_bounds__(256, 2) k_adj_lamq_march(AdjArgs a) {
    constexpr int HX = 70, HY = 10;
    __shared__ double tl[5][HY][HX];
    __shared__ double tq[4][HY][HX];
    const Box &b = a.b;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4, i = i0 + tx, j = j0 + ty;
    const bool own = i < b.nx && j < b.ny;
    const int k0 = blockIdx.z * a.seg, k1 = min(k0 + a.seg, b.nz);
    const PlaneMap pm(b, i0, j0, tx, ty);
    const double *Qv[4] = {a.Q[0], a.Q[1], a.Q[2], a.Q[4]};
    double wl[5][7], wq[4][7];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
#pragma unroll
        for (int n = 0; n < 5; ++n) wl[n][m] = a.X[n][pm.col + pm.plane(k0 - 3 + m)];
#pragma unroll
        for (int q = 0; q < 4; ++q) wq[q][m] = Qv[q][pm.col + pm.plane(k0 - 3 + m)];
    }
    for (int k = k0; k < k1; ++k) {
#pragma unroll
        for (int n = 0; n < 5; ++n) wl[n][6] = a.X[n][pm.col + pm.plane(k + 3)];
#pragma unroll
        for (int q = 0; q < 4; ++q) wq[q][6] = Qv[q][pm.col + pm.plane(k + 3)];
        double hl[2][5], hq[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int n = 0; n < 5; ++n) hl[h][n] = pm.hv[h] ? a.X[n][pm.hc[h] + pm.plane(k)] : 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) hq[h][q] = pm.hv[h] ? Qv[q][pm.hc[h] + pm.plane(k)] : 0.0;
        }
        __syncthreads();  // the previous plane's tiles have been read
#pragma unroll
        for (int n = 0; n < 5; ++n) tl[n][ty + 3][tx + 3] = wl[n][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) tq[q][ty + 3][tx + 3] = wq[q][3];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (pm.hv[h]) {
#pragma unroll
                for (int n = 0; n < 5; ++n) tl[n][pm.hj[h] + 3][pm.hi[h] + 3] = hl[h][n];
#pragma unroll
                for (int q = 0; q < 4; ++q) tq[q][pm.hj[h] + 3][pm.hi[h] + 3] = hq[h][q];
            }
        __syncthreads();
        if (own) {
            const long o = b.oidx(i, j, k);
            double lg[4][3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int n = 0; n < 5; ++n) {
                    double lv[7];
#pragma unroll
                    for (int m = 0; m < 7; ++m)
                        lv[m] = d == 0 ? tl[n][ty + 3][tx + m] : d == 1 ? tl[n][ty + m][tx + 3] : wl[n][m];
                    double sum = 0.0;
#pragma unroll
                    for (int m = 0; m < 6; ++m) sum += kW[m] * lv[6 - m];
                    __builtin_nontemporal_store(-sum * a.ih[d], &a.FB[d][n][o]);
                    if (n > 0) lg[n - 1][d] = d6c(lv[0], lv[1], lv[2], lv[4], lv[5], lv[6], a.i60[d]);
                }
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
            double g[4][3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                g[q][0] = d6c(tq[q][ty + 3][tx], tq[q][ty + 3][tx + 1], tq[q][ty + 3][tx + 2], tq[q][ty + 3][tx + 4],
                              tq[q][ty + 3][tx + 5], tq[q][ty + 3][tx + 6], a.i60[0]);
                g[q][1] = d6c(tq[q][ty][tx + 3], tq[q][ty + 1][tx + 3], tq[q][ty + 2][tx + 3], tq[q][ty + 4][tx + 3],
                              tq[q][ty + 5][tx + 3], tq[q][ty + 6][tx + 3], a.i60[1]);
                g[q][2] = d6c(wq[q][0], wq[q][1], wq[q][2], wq[q][4], wq[q][5], wq[q][6], a.i60[2]);
            }
            double gb[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, qb[4] = {0, 0, 0, 0};
            visc_local_vjp(a.mu, a.S_ref, a.cp_pr, wq[0][3], wq[1][3], wq[2][3], wq[3][3], g, Vb, gb, qb);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int d = 0; d < 3; ++d) a.Gb[3 * q + d][o] = gb[q][d];
                a.qb[q][o] = qb[q];
            }
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) {
#pragma unroll
            for (int n = 0; n < 5; ++n) wl[n][m] = wl[n][m + 1];
#pragma unroll
            for (int q = 0; q < 4; ++q) wq[q][m] = wq[q][m + 1];
        }
    }
}

// out = alpha A + beta B + gamma P on the interior cells; a null operand
// counts as zero, and out may alias any operand.
__device__ __forceinline__ double lincomb_at(const LinComb &l, int n, long g) {
    return lin_at(l, n, g, l.P[n] ? l.P[n][g] : 0.0);
}

// Interior cell of a thread (32-bit index arithmetic: the box has fewer than
// 2^31 cells).
__device__ __forceinline__ bool cell_of32(const Box &b, int id, int &i, int &j, int &k) {
    if (id >= b.nx * b.ny * b.nz) return false;
    i = id % b.nx;
    const int r = id / b.nx;
    j = r % b.ny;
    k = r / b.ny;
    return true;
}

__global__ void __launch_bounds__(256) k_lincomb(LinComb l) {
    int i, j, k;
    if (!cell_of32(l.b, blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long g = l.b.gidx(i, j, k);
#pragma unroll
    for (int n = 0; n < 5; ++n) l.out[n][g] = lincomb_at(l, n, g);
}

// Zero on the ghost layers of l.out (the interior untouched), over the whole
// ghosted box.
__global__ void __launch_bounds__(256) k_zero_ghosts(LinComb l) {
    const Box &b = l.b;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= b.sx * b.sy * b.sz) return;
    const int x = id % b.sx, r = id / b.sx, y = r % b.sy, z = r / b.sy;
    if (x >= NG && x < NG + b.nx && y >= NG && y < NG + b.ny && z >= NG && z < NG + b.nz) return;
#pragma unroll
    for (int n = 0; n < 5; ++n) l.out[n][id] = 0.0;
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

template <int DIR, bool JS, int RS>
void adj_dir_t(const AdjArgs &a, hipStream_t s) {
    launch_adj_march<DIR, k_adj_riemann<DIR, JS ? I_WENO5_JS : I_WENO5, RS>>(a, s);
    launch_adj_march<DIR, k_adj_weno<DIR, JS, DIR == 1>>(a, s);
}

template <int DIR, bool JS>
void adj_dir_r(const AdjArgs &a, int rs, hipStream_t s) {
    switch (rs) {
        case R_ROE: adj_dir_t<DIR, JS, R_ROE>(a, s); break;
        case R_RUSANOV: adj_dir_t<DIR, JS, R_RUSANOV>(a, s); break;
        case R_HLLC: adj_dir_t<DIR, JS, R_HLLC>(a, s); break;
        case R_HLL: adj_dir_t<DIR, JS, R_HLL>(a, s); break;
        case R_HLLC_P: adj_dir_t<DIR, JS, R_HLLC_P>(a, s); break;
        default: adj_dir_t<DIR, JS, R_AUSM>(a, s); break;
    }
}

template <int DIR, int INTERP>
void adj_mdcd_r(const AdjArgs &a, int rs, hipStream_t s) {
    constexpr bool FIRST = DIR == 1;
    switch (rs) {
        case R_ROE: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_ROE, FIRST>>(a, s); break;
        case R_RUSANOV: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_RUSANOV, FIRST>>(a, s); break;
        case R_HLLC: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_HLLC, FIRST>>(a, s); break;
        case R_HLL: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_HLL, FIRST>>(a, s); break;
        case R_HLLC_P: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_HLLC_P, FIRST>>(a, s); break;
        default: launch_adj_march<DIR, k_adj_mdcd<DIR, INTERP, R_AUSM, FIRST>>(a, s); break;
    }
}

template <int DIR>
void adj_dir(const AdjArgs &a, int interp, int rs, hipStream_t s) {
    if (interp == I_MDCD_LINEAR) adj_mdcd_r<DIR, I_MDCD_LINEAR>(a, rs, s);
    else if (interp == I_MDCD_HYBRID) adj_mdcd_r<DIR, I_MDCD_HYBRID>(a, rs, s);
    else if (interp == I_WENO5_JS) adj_dir_r<DIR, true>(a, rs, s);
    else adj_dir_r<DIR, false>(a, rs, s);
}

template <bool JS, int RS>
void adj_x_t(const AdjArgs &a, hipStream_t s, const LinTail &t) {
    const int N = a.b.nx;
    if (N > XSEG) {  // segments of at most XSEG cells, balanced over the line
        const int nseg = (N + XSEG - 1) / XSEG, S = (N + nseg - 1) / nseg;
        const dim3 g((unsigned)((long)a.b.ny * a.b.nz * nseg)), bl(XSEG);
        const size_t shm = (size_t)xseg_doubles(S) * sizeof(double);
        hipLaunchKernelGGL((k_adj_xseg<JS, RS>), g, bl, shm, s, a, t, nseg, S);
        return;
    }
    const dim3 g((unsigned)((long)a.b.ny * a.b.nz)), bl((unsigned)(((N + 63) / 64) * 64));
    const size_t shm = (size_t)(5 * (N + 4) + 20 * N) * sizeof(double);
    hipLaunchKernelGGL((k_adj_xline<JS, RS>), g, bl, shm, s, a, t);
}

template <bool JS>
void adj_x_r(const AdjArgs &a, int rs, hipStream_t s, const LinTail &t) {
    switch (rs) {
        case R_ROE: adj_x_t<JS, R_ROE>(a, s, t); break;
        case R_RUSANOV: adj_x_t<JS, R_RUSANOV>(a, s, t); break;
        case R_HLLC: adj_x_t<JS, R_HLLC>(a, s, t); break;
        case R_HLL: adj_x_t<JS, R_HLL>(a, s, t); break;
        case R_HLLC_P: adj_x_t<JS, R_HLLC_P>(a, s, t); break;
        default: adj_x_t<JS, R_AUSM>(a, s, t); break;
    }
}

template <int INTERP, int RS>
void adj_mdcd_x_t(const AdjArgs &a, hipStream_t s, const LinTail &t) {
    const int N = a.b.nx;
    const int nseg = (N + XSEG - 1) / XSEG, S = (N + nseg - 1) / nseg;
    const dim3 g((unsigned)((long)a.b.ny * a.b.nz * nseg)), bl(XSEG);
    const size_t shm = (size_t)xmdcd_doubles(S) * sizeof(double);
    hipLaunchKernelGGL((k_adj_mdcd_x<INTERP, RS>), g, bl, shm, s, a, t, nseg, S);
}

template <int INTERP>
void adj_mdcd_x_r(const AdjArgs &a, int rs, hipStream_t s, const LinTail &t) {
    switch (rs) {
        case R_ROE: adj_mdcd_x_t<INTERP, R_ROE>(a, s, t); break;
        case R_RUSANOV: adj_mdcd_x_t<INTERP, R_RUSANOV>(a, s, t); break;
        case R_HLLC: adj_mdcd_x_t<INTERP, R_HLLC>(a, s, t); break;
        case R_HLL: adj_mdcd_x_t<INTERP, R_HLL>(a, s, t); break;
        case R_HLLC_P: adj_mdcd_x_t<INTERP, R_HLLC_P>(a, s, t); break;
        default: adj_mdcd_x_t<INTERP, R_AUSM>(a, s, t); break;
    }
}

void adj_x(const AdjArgs &a, int interp, int rs, hipStream_t s, const LinTail &t) {
    if (interp == I_MDCD_LINEAR) adj_mdcd_x_r<I_MDCD_LINEAR>(a, rs, s, t);
    else if (interp == I_MDCD_HYBRID) adj_mdcd_x_r<I_MDCD_HYBRID>(a, rs, s, t);
    else if (interp == I_WENO5_JS) adj_x_r<true>(a, rs, s, t);
    else adj_x_r<false>(a, rs, s, t);
}
}  // namespace

bool adjoint_supported(int interp, int rs, bool chr) {
    return !chr && (interp == I_WENO5 || interp == I_WENO5_JS || interp == I_MDCD_LINEAR || interp == I_MDCD_HYBRID) &&
           rs >= R_ROE && rs <= R_AUSM;
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

void launch_adj_rhs(const AdjArgs &a, int interp, int rs, hipStream_t s, const LinTail &tail) {
    // the face and viscous gradient adjoints first, then the y and z
    // directions (y writes the interior of out, z accumulates), then the x
    // direction with the viscous part (ghosts of out are not written)
    launch_adj_march<2, k_adj_lamq_march>(a, s);
    adj_dir<1>(a, interp, rs, s);
    adj_dir<2>(a, interp, rs, s);
    adj_x(a, interp, rs, s, tail);
}

void launch_lincomb(const LinComb &l, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb, cgrid(l.b), dim3(256), 0, s, l);
}

void launch_zero_ghosts(const LinComb &l, hipStream_t s) {
    const long cells = (long)l.b.sx * l.b.sy * l.b.sz;
    hipLaunchKernelGGL(k_zero_ghosts, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, l);
}

}  // namespace ns3d
