// ns3d_kernels.hip -- cell-wise kernels of the RHS + RK3 path: viscous fluxes,
// RHS assembly fused with the SSP-RK3 stage update and the C2P/P2C round trip,
// periodic halo fill, CFL reduction and the kinetic-energy sum.
#include "ns3d_kernels.h"

namespace ns3d {

// 6th-order midpoint-to-node difference of faces i-5/2..i+5/2 times 1/h
// (diff_6th_central_half, InviscousFlux.cpp:990-992).
__device__ __forceinline__ double d6h(double f0, double f1, double f2, double f3, double f4, double f5, double ih) {
    return ((f3 - f2) * (75.0 / 64.0) + (f4 - f1) * (-25.0 / 384.0) + (f5 - f0) * (3.0 / 640.0)) * ih;
}

// ---------------------------------------------------------------- viscous
// Velocity/temperature gradients, Sutherland mu, stresses and heat flux
// (compute_gradients + compute_viscous_flux, ViscousFlux.cpp:109-208) in one
// pass; writes the nine distinct viscous flux components in compact layout.
__global__ void __launch_bounds__(256) k_visc(ViscArgs a) {
    int i, j, k;
    if (!tile_cell<TILE_Z_FAST>(a.b, i, j, k)) return;
    if (i >= a.b.nx || j >= a.b.ny) return;
    const long c = a.b.gidx(i, j, k);
    const long sx = a.b.sx, sxy = (long)a.b.sx * a.b.sy;
    double g[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double *f = a.Q[q];
        g[q][0] = d6c(f[c - 3], f[c - 2], f[c - 1], f[c + 1], f[c + 2], f[c + 3], a.i60[0]);
        g[q][1] = d6c(f[c - 3 * sx], f[c - 2 * sx], f[c - sx], f[c + sx], f[c + 2 * sx], f[c + 3 * sx], a.i60[1]);
        g[q][2] = d6c(f[c - 3 * sxy], f[c - 2 * sxy], f[c - sxy], f[c + sxy], f[c + 2 * sxy], f[c + 3 * sxy], a.i60[2]);
    }
    const long o = a.b.oidx(i, j, k);
    if (a.G[0]) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) a.G[3 * q + d][o] = g[q][d];
    }
    if (!a.V[0]) return;
    double V[NVISC];
    visc_local(a.mu, a.S_ref, a.cp_pr, a.Q[0][c], a.Q[1][c], a.Q[2][c], a.Q[3][c], g, V);
#pragma unroll
    for (int n = 0; n < NVISC; ++n) a.V[n][o] = V[n];
}

// The same viscous fluxes from a z-marching 64 x 8 column tile: each thread
// keeps u, v, w, T of planes k-3..k+3 of its own column in registers (z
// derivatives); each plane's tile with a 3-cell x/y halo is staged in LDS
// (x and y derivatives).  Cells beyond the box are read from their periodic
// image.
// 64 x 8 tiles marching 32 planes: the plane tile with its halo is 70 x 14
// cells per 512 outputs (1.91 loads per cell against 2.73 for 64 x 4), and
// the 6-plane window prologue is paid once per 32 planes.  The FP64 kernel is
// held to 128 registers, four waves per SIMD: two blocks per CU.  A box with
// too few tiles per plane for two blocks per CU marches shorter segments,
// down to VZ_SEG_MIN planes (64^3: 4 planes, 128 blocks instead of 16).
constexpr int VZ_SEG = 32;
constexpr int VZ_SEG_MIN = 4;
constexpr int VT_Y = 8;      // tile rows (64 x VT_Y threads per block)
constexpr int VT_WAVES = 4;  // minimum waves per SIMD asked of the register allocation

// R = double: the plain method.  R = float: the mixed method's FP32 viscous
// stencils; the temperature enters as T - T_ref (subtracted in FP64, T_ref =
// 1, the nondimensional reference temperature) and the nine fluxes are
// stored as FP32 in the first half of each FP64 buffer.
//
// Q32 (mixed only): the primitives are the compact FP32 u, v, w, T - T_ref
// that the previous stage's x pencil stored (a.Q32; the same values this
// kernel otherwise rounds from the ghosted FP64 ones).
template <class R, bool Q32 = false>
__global__ void __launch_bounds__(64 * VT_Y, VT_WAVES) k_visc_march(ViscArgs a) {
    constexpr int TX = 64, TY = VT_Y, HX = TX + 6, HY = TY + 6, NT = TX * TY;
    constexpr int NHALO = 6 * TX + 6 * TY, NH = (NHALO + NT - 1) / NT;
    __shared__ R tile[4][HY][HX];
    const Box &b = a.b;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < b.nx && j < b.ny;
    // every tile slot holds the periodic image of its cell (no ghost layer is
    // read); slots beyond the first three columns / rows past the box are
    // staged but never read
    auto per = [](int x, int n) { return ((x % n) + n) % n; };
    const int k0 = blockIdx.z * a.zseg, k1 = min(k0 + a.zseg, b.nz);
    // plane stride and (i, j, 0) index of the primitives' layout
    const long sxy = Q32 ? (long)b.nx * b.ny : (long)b.sx * b.sy;
    auto at = [&](int x, int y) { return Q32 ? b.oidx(x, y, 0) : b.gidx(x, y, 0); };
    const double tref = sizeof(R) == 4 ? 1.0 : 0.0;
    auto ld = [&](int q, long g) -> R {
        if constexpr (Q32) return a.Q32[q][g];
        else return q == 3 ? (R)(a.Q[3][g] - tref) : (R)a.Q[q][g];
    };
    const R i60[3] = {(R)a.i60[0], (R)a.i60[1], (R)a.i60[2]};
    const R mu_ref = (R)a.mu, S_ref = (R)a.S_ref, cp_pr = (R)a.cp_pr;
    // index of the periodic image of (i, j, 0)
    const long col = at(per(i, b.nx), per(j, b.ny));
    R wq[4][7];
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) wq[q][m] = ld(q, col + (long)wrap(k0 - 3 + m, b.nz) * sxy);
    // halo cells of a plane: 6 rows x 64 above/below, then TY rows x 6 left/right
    int hi[NH], hj[NH];
    bool hv[NH];
    long hc[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        const int e = tid + NT * h;
        int x, y;
        if (e < 6 * TX) {
            const int r = e / TX;
            y = r < 3 ? r - 3 : TY + r - 3;
            x = e % TX;
        } else {
            const int e2 = e - 6 * TX, r = e2 / 6, c6 = e2 % 6;
            y = r;
            x = c6 < 3 ? c6 - 3 : TX + c6 - 3;
        }
        hi[h] = x;
        hj[h] = y;
        hv[h] = e < NHALO;
        hc[h] = at(per(i0 + x, b.nx), per(j0 + y, b.ny));
    }
    for (int k = k0; k < k1; ++k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) wq[q][6] = ld(q, col + (long)wrap(k + 3, b.nz) * sxy);
        R hvv[NH][4];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const long g = hc[h] + (long)k * sxy;
#pragma unroll
            for (int q = 0; q < 4; ++q) hvv[h][q] = hv[h] ? ld(q, g) : R(0);
        }
        __syncthreads();  // the previous plane's tile has been read
#pragma unroll
        for (int q = 0; q < 4; ++q) tile[q][ty + 3][tx + 3] = wq[q][3];
#pragma unroll
        for (int h = 0; h < NH; ++h)
            if (hv[h])
#pragma unroll
                for (int q = 0; q < 4; ++q) tile[q][hj[h] + 3][hi[h] + 3] = hvv[h][q];
        __syncthreads();
        if (own) {
            R g[4][3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const R(*tq)[HX] = tile[q];
                g[q][0] = d6c(tq[ty + 3][tx], tq[ty + 3][tx + 1], tq[ty + 3][tx + 2], tq[ty + 3][tx + 4],
                              tq[ty + 3][tx + 5], tq[ty + 3][tx + 6], i60[0]);
                g[q][1] = d6c(tq[ty][tx + 3], tq[ty + 1][tx + 3], tq[ty + 2][tx + 3], tq[ty + 4][tx + 3],
                              tq[ty + 5][tx + 3], tq[ty + 6][tx + 3], i60[1]);
                g[q][2] = d6c(wq[q][0], wq[q][1], wq[q][2], wq[q][4], wq[q][5], wq[q][6], i60[2]);
            }
            const R T = sizeof(R) == 4 ? wq[3][3] + R(1) : wq[3][3];
            R Vc[NVISC];
            visc_local(mu_ref, S_ref, cp_pr, wq[0][3], wq[1][3], wq[2][3], T, g, Vc);
            const long o = b.oidx(i, j, k);
            R *const *V = reinterpret_cast<R *const *>(a.V);
#pragma unroll
            for (int n = 0; n < NVISC; ++n) __builtin_nontemporal_store(Vc[n], &V[n][o]);  // read once, next kernel
        }
#pragma unroll
        for (int m = 0; m < 6; ++m)
#pragma unroll
            for (int q = 0; q < 4; ++q) wq[q][m] = wq[q][m + 1];
    }
}

// Viscous fluxes from stored gradients (compute_viscous_flux alone,
// ViscousFlux.cpp:151-208), for the stepwise call surface.
__global__ void __launch_bounds__(256) k_visc_from_grad(ViscArgs a) {
    int i, j, k;
    if (!tile_cell<TILE_Z_FAST>(a.b, i, j, k)) return;
    if (i >= a.b.nx || j >= a.b.ny) return;
    const long c = a.b.gidx(i, j, k);
    const long o = a.b.oidx(i, j, k);
    double g[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) g[q][d] = a.G[3 * q + d][o];
    double V[NVISC];
    visc_local(a.mu, a.S_ref, a.cp_pr, a.Q[0][c], a.Q[1][c], a.Q[2][c], a.Q[3][c], g, V);
#pragma unroll
    for (int n = 0; n < NVISC; ++n) a.V[n][o] = V[n];
}

// ---------------------------------------------------------------- assembly
// Per interior cell: rhs = - d/dx F - d/dy G - d/dz H (per variable x, y, z as
// in compute_invis_dflux, InviscousFlux.cpp:1037-1055) + viscous divergence
// (compute_vis_flux, ViscousFlux.cpp:239-253).  Face and viscous-flux indices
// wrap periodically, which reads the values the reference's half-node and
// viscous-flux halo exchanges copy (field_structures.h:773-932).
//
// stage 0 stores the RHS; stage 1..3 applies the Shu-Osher update
// (range_kutta.cpp:53-87) and the conserved -> primitive -> conserved round
// trip (field_structures.h:444-488) in registers and stores U and Q.
__global__ void __launch_bounds__(256) k_assemble(AsmArgs a) {
    int i, j, k;
    if (!tile_cell<TILE_Z_FAST>(a.b, i, j, k)) return;
    if (i >= a.b.nx || j >= a.b.ny) return;
    const Box &b = a.b;
    const long c = b.gidx(i, j, k);
    long ox[7], oy[7], oz[7];  // compact indices of offsets -3..+3 along each axis
#pragma unroll
    for (int m = 0; m < 7; ++m) {
        ox[m] = b.oidx(wrap(i + m - 3, b.nx), j, k);
        oy[m] = b.oidx(i, wrap(j + m - 3, b.ny), k);
        oz[m] = b.oidx(i, j, wrap(k + m - 3, b.nz));
    }
    double r[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) r[n] = (a.mode & ASM_ACCUMULATE) ? a.RHS[n][c] : 0.0;
    if (a.mode & ASM_INVISCID) {
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const double *X = a.F[0][n], *Y = a.F[1][n], *Z = a.F[2][n];
            // faces i-3 .. i+2 are stored at cells i-3 .. i+2
            r[n] -= d6h(X[ox[0]], X[ox[1]], X[ox[2]], X[ox[3]], X[ox[4]], X[ox[5]], a.ih[0]);
            r[n] -= d6h(Y[oy[0]], Y[oy[1]], Y[oy[2]], Y[oy[3]], Y[oy[4]], Y[oy[5]], a.ih[1]);
            r[n] -= d6h(Z[oz[0]], Z[oz[1]], Z[oz[2]], Z[oz[3]], Z[oz[4]], Z[oz[5]], a.ih[2]);
        }
    }
    if (a.mode & ASM_VISCOUS) {
        // momentum n = 1..3 and energy n = 4: d/dx Fv + d/dy Hv + d/dz Gv
        const int vx[4] = {VXX, VXY, VXZ, VEX};
        const int vy[4] = {VXY, VYY, VYZ, VEY};
        const int vz[4] = {VXZ, VYZ, VZZ, VEZ};
#pragma unroll
        for (int n = 1; n < 5; ++n) {
            const double *X = a.V[vx[n - 1]], *Y = a.V[vy[n - 1]], *Z = a.V[vz[n - 1]];
            r[n] += d6c(X[ox[0]], X[ox[1]], X[ox[2]], X[ox[4]], X[ox[5]], X[ox[6]], a.i60[0]);
            r[n] += d6c(Y[oy[0]], Y[oy[1]], Y[oy[2]], Y[oy[4]], Y[oy[5]], Y[oy[6]], a.i60[1]);
            r[n] += d6c(Z[oz[0]], Z[oz[1]], Z[oz[2]], Z[oz[4]], Z[oz[5]], Z[oz[6]], a.i60[2]);
        }
    }
    if (a.stage == 0) {
#pragma unroll
        for (int n = 0; n < 5; ++n) a.RHS[n][c] = r[n];
        return;
    }
    const double dt = a.dt ? *a.dt : a.dtv;
    double un[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        const double u0 = a.U0[n][c];
        if (a.stage == 1) un[n] = u0 + dt * r[n];
        else if (a.stage == 2) un[n] = 0.75 * u0 + 0.25 * (a.U[n][c] + dt * r[n]);
        else un[n] = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (a.U[n][c] + dt * r[n]);
    }
    // conservedToPrimitive then primitiveToConserved
    double rr = un[0];
    if (rr <= 0.0) rr = 1e-12;
    const double ir = 1.0 / rr;
    const double u = un[1] * ir, v = un[2] * ir, w = un[3] * ir;
    const double kin = 0.5 * (u * u + v * v + w * w);
    const double p = (a.gamma - 1.0) * rr * (un[4] * ir - kin);
    const double T = p / (rr * a.Rgas);
    a.U[0][c] = rr;
    a.U[1][c] = rr * u;
    a.U[2][c] = rr * v;
    a.U[3][c] = rr * w;
    a.U[4][c] = rr * (p / ((a.gamma - 1.0) * rr) + kin);
    a.Q[0][c] = u;
    a.Q[1][c] = v;
    a.Q[2][c] = w;
    a.Q[3][c] = p;
    a.Q[4][c] = T;
}

// ---------------------------------------------------------------- conversions
// Field3D::conservedToPrimitive over all cells (field_structures.h:444-465),
// without contraction: the same operation sequence rounds to the reference's bits.
__global__ void k_c2p(ConvArgs a) {
#pragma clang fp contract(off)
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= a.n) return;
    double rr = a.U[0][id];
    if (rr <= 0.0) {
        rr = 1e-12;
        if (!a.keep) a.U[0][id] = rr;
    }
    const double u = a.U[1][id] / rr, v = a.U[2][id] / rr, w = a.U[3][id] / rr;
    const double kin = 0.5 * (u * u + v * v + w * w);
    const double p = (a.gamma - 1.0) * rr * (a.U[4][id] / rr - kin);
    a.Q[0][id] = u;
    a.Q[1][id] = v;
    a.Q[2][id] = w;
    a.Q[3][id] = p;
    a.Q[4][id] = p / (rr * a.Rgas);
}

// Field3D::primitiveToConserved over all cells (field_structures.h:468-488).
__global__ void k_p2c(ConvArgs a) {
#pragma clang fp contract(off)
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= a.n) return;
    double rr = a.U[0][id];
    if (rr <= 0.0) {
        rr = 1e-12;
        a.U[0][id] = rr;
    }
    const double u = a.Q[0][id], v = a.Q[1][id], w = a.Q[2][id], p = a.Q[3][id];
    a.U[1][id] = rr * u;
    a.U[2][id] = rr * v;
    a.U[3][id] = rr * w;
    const double kin = 0.5 * (u * u + v * v + w * w);
    a.U[4][id] = rr * (p / ((a.gamma - 1.0) * rr) + kin);
    a.Q[4][id] = p / (rr * a.Rgas);
}

__global__ void k_fill(double *p, long n, double v) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) p[id] = v;
}

// ---------------------------------------------------------------- halo
// Periodic copy into the NG ghost layers of the six faces, tangential range
// interior only (exchange_halos_physical on one periodic rank,
// field_structures.h:511-649, 936-994).  blockIdx.y selects the x, y or z
// faces; element e = ((side*NG + layer)*B + b)*A + a with a the fastest
// tangential index (x for the y and z faces).
__global__ void k_halo(HaloArgs h) {
    const Box &b = h.b;
    const int grp = blockIdx.y;
    const int A = grp == 0 ? b.ny : b.nx;
    const int B = grp == 2 ? b.ny : b.nz;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)A * B;
    if (e >= 2 * NG * per) return;
    const int a = (int)(e % A);
    const int bb = (int)((e / A) % B);
    const int layer = (int)((e / per) % NG);
    const int side = (int)(e / (per * NG));
    const int n = grp == 0 ? b.nx : (grp == 1 ? b.ny : b.nz);
    // ghosted coordinate along the normal: left ghosts take cell + n, right ghosts cell - n
    const int gd = side == 0 ? layer : NG + n + layer;
    const int gs = side == 0 ? layer + n : NG + layer;
    long dst, src;
    const long sx = b.sx, sxy = (long)b.sx * b.sy;
    if (grp == 0) {
        const long base = (long)(bb + NG) * sxy + (long)(a + NG) * sx;
        dst = base + gd;
        src = base + gs;
    } else if (grp == 1) {
        const long base = (long)(bb + NG) * sxy + (a + NG);
        dst = base + (long)gd * sx;
        src = base + (long)gs * sx;
    } else {
        const long base = (long)(bb + NG) * sx + (a + NG);
        dst = base + (long)gd * sxy;
        src = base + (long)gs * sxy;
    }
    for (int f = 0; f < h.nf; ++f) h.f[f][dst] = h.f[f][src];
}

// ---------------------------------------------------------------- CFL
// max over the interior of |u| + a (cfl_speed, ns3d_common.h); NaN
// candidates are skipped, as std::min skips them in timestep.cpp.

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// The block's maximum CFL speed over the interior cells: block-strided rows
// (y, z) of the interior with the threads along x, four rows of each thread
// in flight (a maximum does not depend on the order of the cells).  The row
// index is uniform over the block, so no per-cell index division is formed.
__global__ void __launch_bounds__(256) k_cfl(CflArgs a) {
    const Box &b = a.b;
    const int rows = b.ny * b.nz, G = (int)gridDim.x;
    double s = 0.0;
    for (int r0 = blockIdx.x; r0 < rows; r0 += 4 * G) {
        long row[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = r0 + u * G;
            row[u] = r < rows ? b.gidx(0, r % b.ny, r / b.ny) : -1;
        }
        for (int x = threadIdx.x; x < b.nx; x += 256) {
            double rho[4], rhou[4], E[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                rho[u] = rhou[u] = E[u] = 0.0;
                if (row[u] >= 0) {
                    const long c = row[u] + x;
                    rho[u] = a.U[0][c];
                    rhou[u] = a.U[1][c];
                    E[u] = a.U[4][c];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (row[u] >= 0) {
                    const double v = cfl_speed(rho[u], rhou[u], E[u], a.gamma);
                    if (v == v) s = fmax(s, v);  // skip NaN
                }
        }
    }
    s = wave_max(s);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

// Second pass: max of the block maxima, then dt = cfl * min(1e9, dx / smax).
__global__ void __launch_bounds__(256) k_cfl_final(CflArgs a, long nblk, bool clear) {
    __shared__ double part[4];
    double s = 0.0;
    for (long q = threadIdx.x; q < nblk; q += 256) {
        s = fmax(s, a.partial[q]);
        // slots of the stage-3 x pencils' atomic maxima start the next step at 0
        if (clear) a.partial[q] = 0.0;
    }
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma clang fp contract(off)
        const double smax = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
        double dtl = a.dx / smax;  // +inf when every candidate was 0 or NaN
        dtl = dtl < 1e9 ? dtl : 1e9;
        *a.dt_out = a.cfl * dtl;
    }
}

// ---------------------------------------------------------------- kinetic energy
// Deterministic two-pass sum of 0.5 rho |u|^2 dV over the interior
// (compute_total_energy, postprocess.cpp:103-128): fixed block count, fixed
// row assignment, fixed reduction tree.
__global__ void __launch_bounds__(256) k_ke_partial(KeArgs a) {
    const Box &b = a.b;
    const long rows = (long)b.ny * b.nz;
    double s = 0.0;
    for (long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int j = (int)(r % b.ny), k = (int)(r / b.ny);
        for (int i = threadIdx.x; i < b.nx; i += 256) {
            const long c = b.gidx(i, j, k);
            const double u = a.u[c], v = a.v[c], w = a.w[c];
            s += 0.5 * a.rho[c] * (u * u + v * v + w * w) * a.vol;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ void __launch_bounds__(256) k_sum_final(const double *partial, long n, double *out) {
    __shared__ double part[4];
    double s = 0.0;
    for (long q = threadIdx.x; q < n; q += 256) s += partial[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (part[0] + part[1]) + (part[2] + part[3]);
}

// ---------------------------------------------------------------- scalar batches
template <int INTERP>
__global__ void k_interp_batch(SchemeCoef<double> c, const double *st, int n, double flag, double *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double *s = st + 6 * (long)t;
    double l, r;
    interp_lr<INTERP>(c, s, l, r);
    out[t] = flag < 0.0 ? r : l;
}

template <int RS, int DIR>
__global__ void k_riemann_batch(SchemeCoef<double> c, const double *UL, const double *UR, int n, double *F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 0 ? 1 : (DIR == 1 ? 3 : 2);
    const double *l = UL + 5 * (long)t, *r = UR + 5 * (long)t;
    const double a[5] = {l[0], l[P1], l[P2], l[P3], l[4]};
    const double b[5] = {r[0], r[P1], r[P2], r[P3], r[4]};
    double f[5];
    riemann<RS, XS>(c, a, b, f);
    double *o = F + 5 * (long)t;
    o[0] = f[0];
    o[P1] = f[1];
    o[P2] = f[2];
    o[P3] = f[3];
    o[4] = f[4];
}

// ---------------------------------------------------------------- launchers
long ke_partial_count(const Box &) { return REDUCE_BLOCKS; }
static unsigned reduce_grid(long rows) { return (unsigned)(rows < REDUCE_BLOCKS ? rows : REDUCE_BLOCKS); }
void launch_visc(const ViscArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_visc, tile_grid(a.b), cell_block(), 0, s, a);
}
void launch_visc_march(const ViscArgs &args, int precision, hipStream_t s) {
    // VZ_SEG planes per segment where the box fills the device; smaller boxes
    // take shorter segments (down to VZ_SEG_MIN) for the two blocks per CU
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            n = 0;
        return n;
    }();
    ViscArgs a = args;
    const long xy = (long)((a.b.nx + 63) / 64) * ((a.b.ny + VT_Y - 1) / VT_Y);
    const long want = 2L * (cus > 0 ? cus : 256);
    const long segs = (want + xy - 1) / xy;
    a.zseg = (int)std::min<long>(VZ_SEG, std::max<long>(VZ_SEG_MIN, (a.b.nz + segs - 1) / segs));
    const dim3 g((a.b.nx + 63) / 64, (a.b.ny + VT_Y - 1) / VT_Y, (a.b.nz + a.zseg - 1) / a.zseg);
    if (precision == 1 && a.Q32[0]) hipLaunchKernelGGL((k_visc_march<float, true>), g, dim3(64, VT_Y), 0, s, a);
    else if (precision == 1) hipLaunchKernelGGL(k_visc_march<float>, g, dim3(64, VT_Y), 0, s, a);
    else hipLaunchKernelGGL(k_visc_march<double>, g, dim3(64, VT_Y), 0, s, a);
}
void launch_visc_from_grad(const ViscArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_visc_from_grad, tile_grid(a.b), cell_block(), 0, s, a);
}
void launch_assemble(const AsmArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_assemble, tile_grid(a.b), cell_block(), 0, s, a);
}
void launch_c2p(const ConvArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_c2p, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
}
void launch_p2c(const ConvArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_p2c, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
}
void launch_fill(double *p, long n, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
void launch_halo(const HaloArgs &h, hipStream_t s) {
    const long m = (long)std::max(std::max(h.b.ny * h.b.nz, h.b.nx * h.b.nz), h.b.nx * h.b.ny) * 2 * NG;
    hipLaunchKernelGGL(k_halo, dim3((unsigned)((m + 255) / 256), 3), dim3(256), 0, s, h);
}
// blocks of the first CFL pass: one per four interior rows, at most REDUCE_BLOCKS
long cfl_block_count(const Box &b) { return reduce_grid(((long)b.ny * b.nz + 3) / 4); }
void launch_cfl(const CflArgs &a, hipStream_t s) {
    const unsigned g = (unsigned)cfl_block_count(a.b);
    hipLaunchKernelGGL(k_cfl, dim3(g), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, s, a, (long)g, false);
}
// dt from the CFL maxima that the previous step's stage-3 x pencils folded
// into a.partial (CFL_SLOTS slots), which are cleared for the next fold
void launch_cfl_final_slots(const CflArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, s, a, (long)CFL_SLOTS, true);
}
void launch_ke(const KeArgs &a, double *out, hipStream_t s) {
    const unsigned g = reduce_grid((long)a.b.ny * a.b.nz);
    hipLaunchKernelGGL(k_ke_partial, dim3(g), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, (const double *)a.partial, (long)g, out);
}
void launch_interp_batch(int interp, const SchemeCoef<double> &c, const double *st, int n, double flag, double *out,
                         hipStream_t s) {
    const dim3 g((n + 255) / 256), bl(256);
    switch (interp) {
        case I_WENO5: hipLaunchKernelGGL(k_interp_batch<I_WENO5>, g, bl, 0, s, c, st, n, flag, out); break;
        case I_MDCD_LINEAR: hipLaunchKernelGGL(k_interp_batch<I_MDCD_LINEAR>, g, bl, 0, s, c, st, n, flag, out); break;
        case I_MDCD_HYBRID: hipLaunchKernelGGL(k_interp_batch<I_MDCD_HYBRID>, g, bl, 0, s, c, st, n, flag, out); break;
        default: hipLaunchKernelGGL(k_interp_batch<I_WENO5_JS>, g, bl, 0, s, c, st, n, flag, out); break;
    }
}

template <int RS>
static void riemann_dir(int dir, double g, const double *UL, const double *UR, int n, double *F, hipStream_t s) {
    const dim3 gr((n + 255) / 256), bl(256);
    SchemeCoef<double> c{};
    c.gamma = g;
    if (dir == 0) hipLaunchKernelGGL((k_riemann_batch<RS, 0>), gr, bl, 0, s, c, UL, UR, n, F);
    else if (dir == 1) hipLaunchKernelGGL((k_riemann_batch<RS, 1>), gr, bl, 0, s, c, UL, UR, n, F);
    else hipLaunchKernelGGL((k_riemann_batch<RS, 2>), gr, bl, 0, s, c, UL, UR, n, F);
}

void launch_riemann_batch(int rs, int dir, double g, const double *UL, const double *UR, int n, double *F,
                          hipStream_t s) {
    switch (rs) {
        case R_ROE: riemann_dir<R_ROE>(dir, g, UL, UR, n, F, s); break;
        case R_RUSANOV: riemann_dir<R_RUSANOV>(dir, g, UL, UR, n, F, s); break;
        case R_HLLC: riemann_dir<R_HLLC>(dir, g, UL, UR, n, F, s); break;
        case R_HLL: riemann_dir<R_HLL>(dir, g, UL, UR, n, F, s); break;
        case R_HLLC_P: riemann_dir<R_HLLC_P>(dir, g, UL, UR, n, F, s); break;
        default: riemann_dir<R_AUSM>(dir, g, UL, UR, n, F, s); break;
    }
}

}  // namespace ns3d
