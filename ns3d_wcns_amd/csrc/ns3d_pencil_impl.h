// ns3d_pencil_impl.h -- fused per-direction pencil kernels of the RHS + RK3
// stage: each computes the WCNS face fluxes of its direction
// (WCNS_Riemann_InviscidFlux, InviscousFlux.cpp:146-268), their 6th-order
// half-node difference (compute_invis_dflux, InviscousFlux.cpp:990-1056) and
// the viscous-flux divergence of that direction (compute_vis_flux,
// ViscousFlux.cpp:210-255) without writing face fluxes to memory.
//
//   x pencil: one block per (group of) periodic x line(s); the line with
//             three periodically wrapped cells on each side is staged in LDS,
//             one thread per face, exactly N faces per line.
//   y / z pencils: one thread per (x, other, segment) column marching along
//             the line with a register window of the stencil cells (five
//             with the paired reconstruction, else six) and seven running
//             cell sums; lanes run along x, so every load is one coalesced
//             row.  Segments of L cells compute L + 5 faces; the segment
//             count follows the kernel's occupancy (march_seg).
// None of them reads a ghost layer: stencil cells beyond the box are taken
// from the periodic image, so the state's ghost layers need to be current
// only where the state is handed out.
//
// Accumulation: the z and y pencils store their parts (-dH/dz + dGv/dz and
// -dG/dy + dHv/dy) in compact partial arrays; the x pencil sums them with its
// own part and then either stores the RHS
// (stage 0) or applies the Shu-Osher stage update and the C2P/P2C round trip
// (range_kutta.cpp:53-91, field_structures.h:444-488) in place: its block
// owns the whole x line, which no other block of that kernel reads.
#pragma once
#include "ns3d_common.h"




namespace ns3d {

// Cell-sum coefficient of face f for cell c = f - 2 + q, q = 0..5 (the
// half-node difference of cell c reads faces c-3..c+2 with weights
// -3/640, 25/384, -75/64, 75/64, -25/384, 3/640).
__device__ constexpr double kHalfW[6] = {3.0 / 640.0, -25.0 / 384.0, 75.0 / 64.0, -75.0 / 64.0, 25.0 / 384.0,
                                         -3.0 / 640.0};

// Viscous-flux slots differentiated along each direction, per momentum/energy
// component (Fv for x, Hv for y, Gv for z; ViscousFlux.cpp:239-253).
__device__ constexpr int kViscSlot[3][4] = {{VXX, VXY, VXZ, VEX}, {VXY, VYY, VYZ, VEY}, {VXZ, VYZ, VZZ, VEZ}};

// Primitives (rho, un, u1, u2, p) of a rotated conserved state, for the
// characteristic projection (the reference projects with the cell's own
// primitive state, InviscousFlux.cpp:174-176, 204-206).
template <class R>
__device__ __forceinline__ void prim_of(const R *U, R g, R *pc) {
    const R ir = frcp(U[0]);
    pc[0] = U[0];
    pc[1] = U[1] * ir;
    pc[2] = U[2] * ir;
    pc[3] = U[3] * ir;
    if constexpr (kPert<R>) pc[4] = pressure_of(U[0], pc[1], pc[2], pc[3], U[4], g);  // p' in FP64
    else pc[4] = (g - R(1)) * (U[4] - R(0.5) * U[0] * (pc[1] * pc[1] + pc[2] * pc[2] + pc[3] * pc[3]));
}

// Viscous flux component of compact cell o: FP64 for the plain method, FP32
// (the mixed method's viscous kernel output, same buffer) otherwise.
template <class R>
__device__ __forceinline__ double visc_at(const double *V, long o) {
    if constexpr (kPert<R>) return (double)reinterpret_cast<const float *>(V)[o];
    else return V[o];
}

// Partial RHS of the z and y pencils at compact cell o: FP64 for the plain
// method; for the mixed method FP32, in the first half of the same buffer.
// The mixed partials are sums of differences of FP32 face fluxes, whose
// rounding (about 1e-7 of the fluxes before cancellation) dominates the 6e-8
// relative rounding of the stored sum.
template <class R>
__device__ __forceinline__ void put_partial(double *P, long o, double v) {
    if constexpr (kPert<R>) __builtin_nontemporal_store((float)v, reinterpret_cast<float *>(P) + o);
    else __builtin_nontemporal_store(v, P + o);
}
template <class R>
__device__ __forceinline__ double get_partial(const double *P, long o) {
    if constexpr (kPert<R>) return (double)__builtin_nontemporal_load(reinterpret_cast<const float *>(P) + o);
    else return __builtin_nontemporal_load(P + o);
}

// Shu-Osher stage update of one cell from the complete RHS r, the recorded
// state u0v and the cell's current state ucur, then the conserved -> primitive -> conserved round
// trip; writes the new state and the primitives (range_kutta.cpp:53-91,
// field_structures.h:444-488).
template <class R>
__device__ __forceinline__ double stage_update(const PencilArgs &a, long g, long o, const double *u0v,
                                               const double *ucur, const double *r) {
    const double dt = a.dt ? *a.dt : a.dtv;
    double un[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        const double u0 = u0v[n];
        if (a.stage == 1) un[n] = u0 + dt * r[n];
        else if (a.stage == 2) un[n] = 0.75 * u0 + 0.25 * (ucur[n] + dt * r[n]);
        else un[n] = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (ucur[n] + dt * r[n]);
    }
    double rr = un[0];
    if (rr <= 0.0) rr = 1e-12;
    const double ir = 1.0 / rr;
    const double u = un[1] * ir, v = un[2] * ir, w = un[3] * ir;
    const double kin = 0.5 * (u * u + v * v + w * w);
    const double p = (a.gamma - 1.0) * rr * (un[4] * ir - kin);
    const double ru = rr * u, En = rr * (p / ((a.gamma - 1.0) * rr) + kin);
    a.Unew[0][g] = rr;
    a.Unew[1][g] = ru;
    a.Unew[2][g] = rr * v;
    a.Unew[3][g] = rr * w;
    a.Unew[4][g] = En;
    // the CFL speed of the stored state, as the CFL reduction forms it from
    // these same values (step 3 of a device time loop folds it)
    const double cs = a.cfl_slots ? cfl_speed(rr, ru, En, a.gamma) : 0.0;
    // the primitives are next read by the following stage's viscous kernel:
    // streamed past the caches (nontemporal), as are the partial sums and U0
    // below and the viscous fluxes, each read once by one later kernel.  The
    // pressure is read by no kernel of the next stage; it is stored with the
    // step's last stage, whose primitives are handed out (an adjoint step,
    // which stops after stage 2, recomputes its primitives by C2P).
    const double T = p / (rr * a.Rgas);
    if constexpr (kPert<R>) {
        // mixed stages whose next kernel is another stage's viscous kernel:
        // it reads u, v, w and T - T_ref (T_ref = 1) in FP32, which it would
        // round from these same FP64 values; stored compact as FP32 at
        // compact cell o (a stage whose primitives are handed out stores
        // them in FP64)
        if (a.Q32[0]) {
            __builtin_nontemporal_store((float)u, a.Q32[0] + o);
            __builtin_nontemporal_store((float)v, a.Q32[1] + o);
            __builtin_nontemporal_store((float)w, a.Q32[2] + o);
            __builtin_nontemporal_store((float)(T - 1.0), a.Q32[3] + o);
            return cs;
        }
    }
    __builtin_nontemporal_store(u, &a.Q[0][g]);
    __builtin_nontemporal_store(v, &a.Q[1][g]);
    __builtin_nontemporal_store(w, &a.Q[2][g]);
    if (a.stage == 3) __builtin_nontemporal_store(p, &a.Q[3][g]);
    __builtin_nontemporal_store(T, &a.Q[4][g]);
    return cs;
}

// The wave's maximum CFL speed (NaN skipped, as the CFL reduction does)
// folded into its slot by an atomic maximum of the bit patterns.  Called by
// every lane of the wave.
__device__ __forceinline__ void cfl_fold(unsigned long long *slots, double v) {
    double s = v == v ? v : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o, 64));
    if ((threadIdx.x & 63) == 0) {
        const unsigned wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        atomicMax(&slots[wave % CFL_SLOTS], (unsigned long long)__double_as_longlong(s));
    }
}

// Component-wise WCNS5 / WENO5-JS: the marches form both reconstructions of a
// cell at once (recon_pair), in FP64 and in the FP32 faces of the mixed
// method (which then still fits three waves per SIMD without scratch).
template <int INTERP, bool CHAR, class R>
constexpr bool kPairRecon = !CHAR && (INTERP == I_WENO5 || INTERP == I_WENO5_JS);

// The x pencil forms both reconstructions of a cell at once for
// component-wise WCNS5 / WENO5-JS when one thread owns one face and one cell.
template <int INTERP, bool CHAR, class R>
constexpr bool kPairReconX = !CHAR && (INTERP == I_WENO5 || INTERP == I_WENO5_JS);

// ------------------------------------------------------------------ x pencil
// Last of the three: adds the x part to the partial RHS of the z and y
// pencils and stores the RHS (stage 0) or updates the line in place (stage
// 1..3; the block holds its whole line in LDS and no other block reads it).
// With a.record == U0_RECORD at stage 1 the current line is also stored as U0.
// ONE: every thread owns exactly one face and one cell (N == threads per
// line); its partial RHS and U0 are then loaded into registers before the
// face work, so that their latency is hidden by it.
template <int INTERP, int RS, bool CHAR, bool ONE, class R>
__global__ void __launch_bounds__(512) k_pencil_x(PencilArgs a, SchemeCoef<R> cf) {
    extern __shared__ double lds[];
    const Box &b = a.b;
    const int N = b.nx, W = N + 2 * NG;
    const int P = a.line_threads, Lb = blockDim.x / P;
    // P is a multiple of 64, so each wave lies within one line: the line
    // index is wave-uniform and its divisions run on the scalar unit
    const int ll = wave_first_thread() / P, t = (int)threadIdx.x - ll * P;
    const int line = (int)blockIdx.x * Lb + ll;
    const bool valid = line < b.ny * b.nz;
    const int j = valid ? line % b.ny : 0, k = valid ? line / b.ny : 0;
    double *Ul = lds + (long)ll * (5 * W + 9 * N);  // U[5][W], F[5][N], V[4][N]
    double *Fl = Ul + 5 * W;
    double *Vl = Fl + 5 * N;
    const long row = ((long)(k + NG) * b.sy + (j + NG)) * b.sx;
    const long orow = b.oidx(0, j, k);
    const bool need_u0 = a.stage > 1 || (a.stage == 1 && a.record == U0_READ);
    double sp[5], zp[5];
    if constexpr (ONE) {
        if (valid) {
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                sp[n] = get_partial<R>(a.Rz[n], orow + t) + get_partial<R>(a.Ry[n], orow + t);
                zp[n] = need_u0 ? __builtin_nontemporal_load(&a.U0[n][row + NG + t]) : 0.0;
            }
        }
    }
    if (valid) {
        if constexpr (ONE) {
            // all loads of the thread in flight at once, then the LDS stores:
            // x = t and x = t + P (< W) of the five fields, x = t of the four
            // viscous fluxes
            double uv[5][2], vv[4];
            const bool tail = t + P < W;
            // line slot x holds cell x - NG of the periodic line (the ghost
            // layers are not read)
            const long g0 = row + NG + wrap(t - NG, N), g1 = row + NG + wrap(t + P - NG, N);
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                uv[n][0] = a.U[n][g0];
                uv[n][1] = tail ? a.U[n][g1] : 0.0;
            }
#pragma unroll
            for (int n = 0; n < 4; ++n) vv[n] = visc_at<R>(a.V[kViscSlot[0][n]], orow + t);
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                Ul[n * W + t] = uv[n][0];
                if (tail) Ul[n * W + t + P] = uv[n][1];
            }
#pragma unroll
            for (int n = 0; n < 4; ++n) Vl[n * N + t] = vv[n];
        } else {
            for (int n = 0; n < 5; ++n)
                for (int x = t; x < W; x += P) Ul[n * W + x] = a.U[n][row + NG + wrap(x - NG, N)];
            for (int n = 0; n < 4; ++n)
                for (int x = t; x < N; x += P) Vl[n * N + x] = visc_at<R>(a.V[kViscSlot[0][n]], orow + x);
        }
    }
    __syncthreads();
    if (valid) {
        for (int f = t; f < N; f += P) {
            R s[5][6];
#pragma unroll
            for (int m = 0; m < 6; ++m)
#pragma unroll
                for (int n = 0; n < 4; ++n) s[n][m] = (R)Ul[n * W + f + 1 + m];
            // E' = E - E_ref in FP64 before the FP32 cast (E_ref = 0 when plain)
#pragma unroll
            for (int m = 0; m < 6; ++m) s[4][m] = kPert<R> ? (R)(Ul[4 * W + f + 1 + m] - a.ref.E_ref)
                                                         : (R)Ul[4 * W + f + 1 + m];
            R pc[2][5];
            if constexpr (CHAR) {
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const R u[5] = {s[0][2 + side], s[1][2 + side], s[2][2 + side], s[3][2 + side], s[4][2 + side]};
                    prim_of(u, cf.gamma, pc[side]);
                }
            }
            R F[5];
            face_flux<INTERP, RS, CHAR, 1>(cf, s, pc, F);
#pragma unroll
            for (int n = 0; n < 5; ++n) Fl[n * N + f] = (double)F[n];
            if constexpr (ONE) break;
        }
    }
    __syncthreads();
    if (!valid) return;
    double cmax = 0.0;  // CFL speed maximum of the thread's cells (stage 3 folding)
    for (int c = t; c < N; c += P) {
        int fi[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) fi[m] = wrap(c + m - 3, N);
        const long g = row + c + NG;
        double r[5], u0[5], u[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            u[n] = Ul[n * W + c + NG];
            double part;
            if constexpr (ONE) {
                part = sp[n];
                u0[n] = need_u0 ? zp[n] : u[n];
            } else {
                part = get_partial<R>(a.Rz[n], orow + c) + get_partial<R>(a.Ry[n], orow + c);
                u0[n] = need_u0 ? a.U0[n][g] : u[n];
            }
            const double *Fn = Fl + n * N;
            r[n] = part - ((Fn[fi[3]] - Fn[fi[2]]) * (75.0 / 64.0) + (Fn[fi[4]] - Fn[fi[1]]) * (-25.0 / 384.0) +
                           (Fn[fi[5]] - Fn[fi[0]]) * (3.0 / 640.0)) *
                              a.ih;
        }
#pragma unroll
        for (int n = 1; n < 5; ++n) {
            const double *V = Vl + (n - 1) * N;
            r[n] += (V[fi[6]] - 9.0 * V[fi[5]] + 45.0 * V[fi[4]] - 45.0 * V[fi[2]] + 9.0 * V[fi[1]] - V[fi[0]]) * a.i60;
        }
        if (a.stage == 0) {
#pragma unroll
            for (int n = 0; n < 5; ++n) a.RHS[n][g] = r[n];
        } else {
            if (a.stage == 1 && a.record == U0_RECORD) {
#pragma unroll
                for (int n = 0; n < 5; ++n) a.U0[n][g] = u[n];
            }
            const double cs = stage_update<R>(a, g, orow + c, u0, u, r);
            cmax = cs == cs ? fmax(cmax, cs) : cmax;
        }
        if constexpr (ONE) break;
    }
    if (a.cfl_slots) cfl_fold(a.cfl_slots, cmax);
}

// LDS doubles per line of the paired x pencil: U and V on the line with three
// periodically wrapped cells on each side (W = N + 6 slots, slot x = cell
// x - 3), the faces -3..N+1 (N + 5 slots, slot s = face s - 3) and the right
// states of faces -1..N-1 (N + 1 slots, slot t = face t - 1).
__host__ __device__ inline long xpair_line_doubles(int N) { return 9L * (N + 2 * NG) + 5L * (N + 5) + 5L * (N + 1); }

// The x pencil with one thread per face and cell (N threads per line, N a
// multiple of 64, at most 512) for component-wise WCNS5 / WENO5-JS.  Thread t
// forms both reconstructions of cell t from cells t-2..t+2 (recon_pair): the
// left state of face t, kept in registers, and the right state of face t - 1,
// handed over in LDS.  The face and viscous-flux arrays are padded with their
// periodic images, so the differences index them without wrapping.  The rest
// is k_pencil_x's ONE path: partials and U0 in registers before the face
// work, then the sum, the stage update and the C2P/P2C round trip in place.
template <int INTERP, int RS, class R>
__global__ void __launch_bounds__(512) k_pencil_xp(PencilArgs a, SchemeCoef<R> cf) {
    extern __shared__ double lds[];
    const Box &b = a.b;
    const int N = b.nx, W = N + 2 * NG, NF = N + 5, NX = N + 1;
    const int Lb = blockDim.x / N;
    // N is a multiple of 64: the line index is wave-uniform (scalar unit)
    const int ll = wave_first_thread() / N, t = (int)threadIdx.x - ll * N;
    const int line = (int)blockIdx.x * Lb + ll;
    const bool valid = line < b.ny * b.nz;
    const int j = valid ? line % b.ny : 0, k = valid ? line / b.ny : 0;
    double *Ul = lds + (long)ll * xpair_line_doubles(N);  // U[5][W], V[4][W], F[5][NF], X[5][NX]
    double *Vl = Ul + 5 * W;
    double *Fl = Vl + 4 * W;
    double *Xl = Fl + 5 * NF;
    const long row = ((long)(k + NG) * b.sy + (j + NG)) * b.sx;
    const long orow = b.oidx(0, j, k);
    const bool need_u0 = a.stage > 1 || (a.stage == 1 && a.record == U0_READ);
    double sp[5], zp[5];
    if (valid) {
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            sp[n] = get_partial<R>(a.Rz[n], orow + t) + get_partial<R>(a.Ry[n], orow + t);
            zp[n] = need_u0 ? __builtin_nontemporal_load(&a.U0[n][row + NG + t]) : 0.0;
        }
        // all loads in flight, then the LDS stores: slots t and t + N (< W)
        double uv[5][2], vv[4][2];
        const bool tail = t + N < W;
        const int x0 = wrap(t - NG, N), x1 = wrap(t + N - NG, N);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            uv[n][0] = a.U[n][row + NG + x0];
            uv[n][1] = tail ? a.U[n][row + NG + x1] : 0.0;
        }
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            vv[n][0] = visc_at<R>(a.V[kViscSlot[0][n]], orow + x0);
            vv[n][1] = tail ? visc_at<R>(a.V[kViscSlot[0][n]], orow + x1) : 0.0;
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            Ul[n * W + t] = uv[n][0];
            if (tail) Ul[n * W + t + N] = uv[n][1];
        }
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            Vl[n * W + t] = vv[n][0];
            if (tail) Vl[n * W + t + N] = vv[n][1];
        }
    }
    __syncthreads();
    R yl[5];
    if (valid) {
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            R f[5];
            // E' = E - E_ref in FP64 before an FP32 cast (E_ref = 0 when plain)
#pragma unroll
            for (int m = 0; m < 5; ++m)
                f[m] = (kPert<R> && n == 4) ? (R)(Ul[n * W + t + 1 + m] - a.ref.E_ref) : (R)Ul[n * W + t + 1 + m];
            R yr;
            recon_pair<INTERP>(f, yl[n], yr);
            Xl[n * NX + t] = (double)yr;
            if (t == 0) Xl[n * NX + N] = (double)yr;
        }
    }
    __syncthreads();
    if (valid) {
        R UR[5], F[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) UR[n] = (R)Xl[n * NX + t + 1];
        riemann<RS, 1>(cf, yl, UR, F);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const double fv = (double)F[n];
            Fl[n * NF + t + 3] = fv;
            if (t < 2) Fl[n * NF + t + 3 + N] = fv;
            if (t >= N - 3) Fl[n * NF + t + 3 - N] = fv;
        }
    }
    __syncthreads();
    if (!valid) return;
    const int c = t;
    const long g = row + c + NG;
    double r[5], u0[5], u[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        u[n] = Ul[n * W + c + NG];
        u0[n] = need_u0 ? zp[n] : u[n];
        const double *Fn = Fl + n * NF + c;  // faces c-3..c+2
        r[n] = sp[n] - ((Fn[3] - Fn[2]) * (75.0 / 64.0) + (Fn[4] - Fn[1]) * (-25.0 / 384.0) +
                        (Fn[5] - Fn[0]) * (3.0 / 640.0)) *
                           a.ih;
    }
#pragma unroll
    for (int n = 1; n < 5; ++n) {
        const double *V = Vl + (n - 1) * W + c;  // cells c-3..c+3
        r[n] += (V[6] - 9.0 * V[5] + 45.0 * V[4] - 45.0 * V[2] + 9.0 * V[1] - V[0]) * a.i60;
    }
    if (a.stage == 0) {
#pragma unroll
        for (int n = 0; n < 5; ++n) a.RHS[n][g] = r[n];
    } else {
        if (a.stage == 1 && a.record == U0_RECORD) {
#pragma unroll
            for (int n = 0; n < 5; ++n) a.U0[n][g] = u[n];
        }
        const double cs = stage_update<R>(a, g, orow + c, u0, u, r);
        if (a.cfl_slots) cfl_fold(a.cfl_slots, cs);
    }
}

// ------------------------------------------------------------ y / z pencils
// Thread per (x, other, segment) column marching along DIR.  Iteration f
// computes face f (its stencil, cells f-2..f+3, in the register window) and
// loads the viscous fluxes of cell f; seven running sums hold cells f-3..f+3, and cell
// f-3 is complete (faces f-6..f-1, viscous fluxes f-6..f) and is emitted
// into the pencil's compact partial-RHS array.
template <int DIR, int INTERP, int RS, bool CHAR, class R>
__device__ __forceinline__ void pencil_m(PencilArgs a, SchemeCoef<R> cf) {
    const Box &b = a.b;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int other = blockIdx.y * 4 + threadIdx.y;  // z for the y pencil, y for the z pencil
    const int nother = DIR == 1 ? b.nz : b.ny;
    if (i >= b.nx || other >= nother) return;
    constexpr int P1 = 1 + DIR, P2 = 1 + (DIR + 1) % 3, P3 = 1 + (DIR + 2) % 3;
    constexpr int XS = DIR == 1 ? 3 : 2;
    const int N = DIR == 1 ? b.ny : b.nz;
    const int c0 = blockIdx.z * a.seg, c1 = min(c0 + a.seg, N);
    const long sxy = (long)b.sx * b.sy;
    const long gst = DIR == 1 ? (long)b.sx : sxy;               // ghosted stride along DIR
    const long ost = DIR == 1 ? (long)b.nx : (long)b.nx * b.ny;  // compact stride along DIR
    const long gbase = DIR == 1 ? b.gidx(i, 0, other) : b.gidx(i, other, 0);
    const long obase = DIR == 1 ? b.oidx(i, 0, other) : b.oidx(i, other, 0);
    constexpr bool mixed = kPert<R>;
    auto load = [&](int c, R *u) {
        const long g = gbase + (long)wrap(c, N) * gst;
        u[0] = (R)a.U[0][g];
        u[1] = (R)a.U[P1][g];
        u[2] = (R)a.U[P2][g];
        u[3] = (R)a.U[P3][g];
        u[4] = mixed ? (R)(a.U[4][g] - a.ref.E_ref) : (R)a.U[4][g];
    };
    // Cell-sum weights: face f -> cells f-2..f+3 (-kHalfW, slots 1..6),
    // viscous flux of cell f -> cells f-3..f+3 (central weights / 60, slots
    // 0..6, none for the cell itself).  FP64 component-wise MDCD: pre-scaled
    // by 1/h.  Otherwise unscaled compile-time constants, which occupy no
    // vector registers (the kernel then fits two waves per SIMD in FP64,
    // three with FP32 faces component-wise), and the cell sum is scaled when
    // emitted.
    constexpr bool PAIR = kPairRecon<INTERP, CHAR, R>;
    constexpr bool pre = sizeof(R) == 8 && !CHAR && !PAIR;
    const double sc = pre ? a.ih : 1.0, se = pre ? 1.0 : a.ih;
    double wf[7], wv[7];
    wf[0] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) wf[q + 1] = -sc * kHalfW[q];
    {
        const double d[7] = {1.0, -9.0, 45.0, 0.0, -45.0, 9.0, -1.0};
        const double s60 = pre ? a.i60 : 1.0 / 60.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) wv[q] = d[q] * s60;
    }
    // Component-wise WCNS5 / WENO5-JS: the two reconstructions of a cell
    // share their smoothness indicators (recon_pair), so iteration f forms
    // both values of cell f + 1 from cells f-1..f+3: the right state of face
    // f and the left state of face f + 1, carried to the next iteration in
    // yl.  The window then holds five cells (W = 5), else the six cells
    // f-2..f+3 of the face's stencil.
    constexpr int W = PAIR ? 5 : 6;
    R w[5][6];
    R yl[5];
    {
        R u[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            load(c0 - 5 + m, u);
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][m] = u[n];
        }
    }
    if constexpr (PAIR) {
        // cell c0 - 3 from cells c0-5..c0-1: the left state of face c0 - 3
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            R yr;
            recon_pair<INTERP>(w[n], yl[n], yr);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][m] = w[n][m + 1];
    }
    double acc[7][5];
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
        for (int n = 0; n < 5; ++n) acc[q][n] = 0.0;
    for (int f = c0 - 3; f <= c1 + 2; ++f) {
        // The window's new cell is loaded in the iteration that uses it: a
        // prefetch one iteration ahead kept the next cell's five values live
        // through the face work, which spilled the FP64 marches at two waves
        // per SIMD; the other resident waves cover the load's latency.
        {
            R nxt[5];
            load(f + 3, nxt);
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][W - 1] = nxt[n];
        }
        double vv[4];
        {
            const long o = obase + (long)wrap(f, N) * ost;
#pragma unroll
            for (int n = 0; n < 4; ++n) vv[n] = visc_at<R>(a.V[kViscSlot[DIR][n]], o);
        }
        // Wave priorities: a wave drops to the lowest issue priority once its
        // loads for this iteration are out, while it waits for them and
        // reconstructs the new cell, and takes a higher one for the Riemann
        // flux, the sums and the next iteration's loads.  The waves of a SIMD
        // then alternate instead of stalling on their loads at the same time.
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
        __builtin_amdgcn_sched_barrier(0);
        if (f <= c1 + 1) {  // the last iteration only completes cell c1 - 1
            R pc[2][5];
            if constexpr (CHAR) {
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const R u[5] = {w[0][2 + side], w[1][2 + side], w[2][2 + side], w[3][2 + side], w[4][2 + side]};
                    prim_of(u, cf.gamma, pc[side]);
                }
            }
            R Fr[5];
            if constexpr (PAIR) {
                R UR[5], yn[5];
#pragma unroll
                for (int n = 0; n < 5; ++n) recon_pair<INTERP>(w[n], yn[n], UR[n]);  // cell f + 1
                // the wave whose new cell has arrived and been reconstructed
                // takes the issue priority for the rest of its iteration
                // (see the loop head)
                __builtin_amdgcn_sched_barrier(0);
                __builtin_amdgcn_s_setprio(2);
                __builtin_amdgcn_sched_barrier(0);
                riemann<RS, XS>(cf, yl, UR, Fr);
#pragma unroll
                for (int n = 0; n < 5; ++n) yl[n] = yn[n];
            } else {
                face_flux<INTERP, RS, CHAR, XS>(cf, w, pc, Fr);
                // (the face's interpolation and flux are one call here: the
                // priority rises after both)
                __builtin_amdgcn_sched_barrier(0);
                __builtin_amdgcn_s_setprio(2);
                __builtin_amdgcn_sched_barrier(0);
            }
            double F[5];
            F[0] = (double)Fr[0];
            F[P1] = (double)Fr[1];
            F[P2] = (double)Fr[2];
            F[P3] = (double)Fr[3];
            F[4] = (double)Fr[4];
#pragma unroll
            for (int q = 1; q < 7; ++q)
#pragma unroll
                for (int n = 0; n < 5; ++n) acc[q][n] += wf[q] * F[n];
        }
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (pre || q != 3)
#pragma unroll
                for (int n = 1; n < 5; ++n) acc[q][n] += wv[q] * vv[n - 1];
        const int c = f - 3;
        if (c >= c0) {
            const long o = obase + (long)c * ost;
            double *const *dst = DIR == 2 ? a.Rz : a.Ry;
#pragma unroll
            for (int n = 0; n < 5; ++n) put_partial<R>(dst[n], o, acc[0][n] * se);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int n = 0; n < 5; ++n) acc[q][n] = acc[q + 1][n];
#pragma unroll
        for (int n = 0; n < 5; ++n) acc[6][n] = 0.0;
#pragma unroll
        for (int m = 0; m < W - 1; ++m)
#pragma unroll
            for (int n = 0; n < 5; ++n) w[n][m] = w[n][m + 1];
    }
}

// The compiler's own register budget: two waves per SIMD for FP64 MDCD
// component-wise and for FP32 faces characteristic-wise.
template <int DIR, int INTERP, int RS, bool CHAR, class R>
__global__ void __launch_bounds__(256) k_pencil_m(PencilArgs a, SchemeCoef<R> cf) {
    pencil_m<DIR, INTERP, RS, CHAR, R>(a, cf);
}

// FP64 characteristic-wise and FP64 paired reconstruction: held to two waves
// per SIMD (256 registers, a few values in scratch) instead of the compiler's
// one-wave allocation.
template <int DIR, int INTERP, int RS, bool CHAR, class R>
__global__ void __launch_bounds__(256, 2) k_pencil_m2(PencilArgs a, SchemeCoef<R> cf) {
    pencil_m<DIR, INTERP, RS, CHAR, R>(a, cf);
}

// FP32 face arithmetic, component-wise: fits three waves per SIMD (168
// registers) without spilling.
template <int DIR, int INTERP, int RS, bool CHAR, class R>
__global__ void __launch_bounds__(256, 3) k_pencil_m3(PencilArgs a, SchemeCoef<R> cf) {
    pencil_m<DIR, INTERP, RS, CHAR, R>(a, cf);
}

// Launch of the DIR march over gx x gy columns of length N.
template <auto K, class R>
inline void launch_march(int gx, int gy, int N, PencilArgs a, const SchemeCoef<R> &cf, hipStream_t s) {
    a.seg = march_seg<K>((long)gx * gy, N);
    hipLaunchKernelGGL(K, dim3(gx, gy, (N + a.seg - 1) / a.seg), dim3(64, 4), 0, s, a, cf);
}

template <int DIR, int INTERP, int RS, bool CHAR, class R>
inline void launch_pencil_m(int gx, int gy, int N, const PencilArgs &a, const SchemeCoef<R> &cf, hipStream_t s) {
    if constexpr (sizeof(R) == 4 && !CHAR) launch_march<k_pencil_m3<DIR, INTERP, RS, CHAR, R>>(gx, gy, N, a, cf, s);
    else if constexpr (sizeof(R) == 8 && (CHAR || kPairRecon<INTERP, CHAR, R>))
        launch_march<k_pencil_m2<DIR, INTERP, RS, CHAR, R>>(gx, gy, N, a, cf, s);
    else launch_march<k_pencil_m<DIR, INTERP, RS, CHAR, R>>(gx, gy, N, a, cf, s);
}

// ------------------------------------------------------------------ launch
template <int INTERP, int RS, bool CHAR, class R>
inline void launch_pencils_t(int dir, const PencilArgs &a, const SchemeCoef<R> &cf, hipStream_t s) {
    const Box &b = a.b;
    if (dir == 2) {
        PencilArgs az = a;
        az.ih = a.ihd[2];
        az.i60 = a.ihd[2] / 60.0;
        launch_pencil_m<2, INTERP, RS, CHAR, R>((b.nx + 63) / 64, (b.ny + 3) / 4, b.nz, az, cf, s);
    } else if (dir == 1) {
        PencilArgs ay = a;
        ay.ih = a.ihd[1];
        ay.i60 = a.ihd[1] / 60.0;
        launch_pencil_m<1, INTERP, RS, CHAR, R>((b.nx + 63) / 64, (b.nz + 3) / 4, b.ny, ay, cf, s);
    } else {
        // one thread per face of a line (N rounded up to 64 lanes, up to
        // 512); lines shorter than 256 share a 256-thread block; longer lines loop
        PencilArgs ax = a;
        const int npad = ((b.nx + 63) / 64) * 64;
        const bool one = npad == b.nx && b.nx <= 512;
        ax.line_threads = npad <= 512 ? npad : 256;
        const int bs = ax.line_threads >= 256 ? ax.line_threads : 256;
        const int lb = bs / ax.line_threads;
        const long lines = (long)b.ny * b.nz;
        const size_t shm = (size_t)lb * (5 * (b.nx + 2 * NG) + 9 * b.nx) * sizeof(double);
        ax.ih = a.ihd[0];
        ax.i60 = a.ihd[0] / 60.0;
        const dim3 gx((unsigned)((lines + lb - 1) / lb));
        if constexpr (kPairReconX<INTERP, CHAR, R>) {
            if (one) {
                const size_t shp = (size_t)lb * xpair_line_doubles(b.nx) * sizeof(double);
                hipLaunchKernelGGL((k_pencil_xp<INTERP, RS, R>), gx, dim3(bs), shp, s, ax, cf);
                return;
            }
        }
        if (one) hipLaunchKernelGGL((k_pencil_x<INTERP, RS, CHAR, true, R>), gx, dim3(bs), shm, s, ax, cf);
        else hipLaunchKernelGGL((k_pencil_x<INTERP, RS, CHAR, false, R>), gx, dim3(bs), shm, s, ax, cf);
    }
}

template <int RS, bool CHAR, class R>
inline void launch_pencils_i(int dir, int interp, const PencilArgs &a, const SchemeCoef<R> &c, hipStream_t s) {
    switch (interp) {
        case I_WENO5: launch_pencils_t<I_WENO5, RS, CHAR>(dir, a, c, s); break;
        case I_MDCD_LINEAR: launch_pencils_t<I_MDCD_LINEAR, RS, CHAR>(dir, a, c, s); break;
        case I_MDCD_HYBRID: launch_pencils_t<I_MDCD_HYBRID, RS, CHAR>(dir, a, c, s); break;
        default: launch_pencils_t<I_WENO5_JS, RS, CHAR>(dir, a, c, s); break;
    }
}

// precision 0: FP64 arithmetic; 1: the mixed method (FP32 face arithmetic in
// the perturbation form, a.ref set, FP32 viscous fluxes in a.V).
template <int RS>
void launch_pencils(int dir, int interp, bool chr, int precision, const PencilArgs &a, const SchemeCoef<double> &c64,
                    const SchemeCoef<float> &c32, hipStream_t s) {
    if (precision == 1) {
        if (chr) launch_pencils_i<RS, true>(dir, interp, a, c32, s);
        else launch_pencils_i<RS, false>(dir, interp, a, c32, s);
    } else {
        if (chr) launch_pencils_i<RS, true>(dir, interp, a, c64, s);
        else launch_pencils_i<RS, false>(dir, interp, a, c64, s);
    }
}

}  // namespace ns3d

#define NS3D_INSTANTIATE_PENCILS(RS)                                                                            \
    template void ns3d::launch_pencils<RS>(int, int, bool, int, const ns3d::PencilArgs &,                   \
                                           const ns3d::SchemeCoef<double> &, const ns3d::SchemeCoef<float> &,  \
                                           hipStream_t);
