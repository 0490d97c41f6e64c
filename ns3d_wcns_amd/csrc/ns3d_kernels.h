// ns3d_kernels.h -- argument blocks and launchers of the cell-wise kernels.
#pragma once
#include <algorithm>

#include "ns3d_common.h"
#include "ns3d_visc.h"

namespace ns3d {

struct ViscArgs {
    const double *Q[4];  // ghosted u, v, w, T
    double *G[12];       // compact gradients (optional: nullptr for the fused path)
    double *V[NVISC];    // compact viscous fluxes (nullptr: gradients only)
    Box b;
    double i60[3];       // 1 / (60 h)
    double mu, S_ref, cp_pr;
    const float *Q32[4];  // mixed: compact FP32 u, v, w, T - 1 read instead of Q (else null)
    int zseg;             // visc march: planes per z segment (set by the launch)
};

enum { ASM_INVISCID = 1, ASM_VISCOUS = 2, ASM_ACCUMULATE = 4 };

struct AsmArgs {
    const double *F[3][5];    // compact face fluxes per direction
    const double *V[NVISC];   // compact viscous fluxes
    double *U[5];             // ghosted conserved (read at stage 2/3, written at 1..3)
    const double *U0[5];      // ghosted recorded state
    double *Q[5];             // ghosted u, v, w, p, T (written at stage 1..3)
    double *RHS[5];           // ghosted RHS (stage 0)
    Box b;
    double ih[3], i60[3];
    double gamma, Rgas;
    const double *dt;         // device scalar, or nullptr to use dtv
    double dtv;
    int stage;                // 0: store RHS; 1..3: RK3 stage
    int mode;                 // ASM_* bits
};

struct ConvArgs {
    double *U[5];
    double *Q[5];
    long n;
    double gamma, Rgas;
    int keep;  // C2P: leave U untouched (no store of the clamped density)
};

struct HaloArgs {
    double *f[10];
    int nf;
    Box b;
};

// Fixed block count of the grid-stride reductions (partials per reduction).
constexpr long REDUCE_BLOCKS = 8192;

struct CflArgs {
    const double *U[5];
    double *partial;
    double *dt_out;
    Box b;
    double gamma, dx, cfl;
};

struct KeArgs {
    const double *rho, *u, *v, *w;
    double *partial;
    Box b;
    double vol;
};

// Tangent-linear / adjoint RHS arguments (ns3d_adjoint.hip).  Cell fields
// are ghosted with filled face ghosts; work arrays are compact.
struct AdjArgs {
    const double *U[5];   // state
    const double *Q[5];   // its primitives u, v, w, p, T
    const double *X[5];   // tangent dU (tangent) or lambda (adjoint)
    double *out[5];       // tangent RHS / adjoint (interior written)
    double *FB[3][5];     // adjoint: face adjoints per direction
    double *SB[10];       // adjoint: left / right state adjoints of one direction's faces
    double *Gb[12];       // adjoint: gradient adjoints
    double *qb[4];        // adjoint: (u, v, w, T) adjoints
    double *dQ[4];        // tangent: ghosted (u, v, w, T) tangents
    double *dF[3][5];     // tangent: face flux tangents per direction
    double *dV[NVISC];    // tangent: viscous flux tangents
    Box b;
    SchemeCoef<double> c;
    double ih[3], i60[3];
    double gamma, Rgas, mu, S_ref, cp_pr;
    int seg;  // cells (planes) per segment of the marching launch at hand
};

// out = alpha A + beta B + gamma P over the interior (null operands are 0).
struct LinComb {
    double *out[5];
    const double *A[5], *B[5], *P[5];
    double alpha, beta, gamma;
    Box b;
};

bool adjoint_supported(int interp, int rs, bool chr);
void launch_tl_prim(const AdjArgs &a, hipStream_t s);
void launch_tl_rhs(const AdjArgs &a, int interp, int rs, bool chr, hipStream_t s);
// Combinations formed from the finished adjoint RHS value r of each cell in
// place of storing r: c[m].out = alpha A + beta B (+ gamma r where c[m].P is
// set), m < n, both formed before either is stored; n = 0 stores r into out.
struct LinTail {
    LinComb c[2];
    int n;
};
void launch_adj_rhs(const AdjArgs &a, int interp, int rs, hipStream_t s, const LinTail &tail);
void launch_lincomb(const LinComb &l, hipStream_t s);
void launch_zero_ghosts(const LinComb &l, hipStream_t s);  // l.out's ghost layers set to 0

long ke_partial_count(const Box &b);
void launch_visc(const ViscArgs &a, hipStream_t s);
// precision 0: FP64 viscous fluxes; 1: FP32 stencils and FP32 fluxes (the
// mixed method's layout of the same buffers).
void launch_visc_march(const ViscArgs &a, int precision, hipStream_t s);
void launch_visc_from_grad(const ViscArgs &a, hipStream_t s);
void launch_assemble(const AsmArgs &a, hipStream_t s);
void launch_c2p(const ConvArgs &a, hipStream_t s);
void launch_p2c(const ConvArgs &a, hipStream_t s);
void launch_fill(double *p, long n, double v, hipStream_t s);
void launch_halo(const HaloArgs &h, hipStream_t s);
// Blocks of the CFL reduction's first pass (block maxima written to a.partial).
long cfl_block_count(const Box &b);
void launch_cfl(const CflArgs &a, hipStream_t s);
void launch_cfl_final_slots(const CflArgs &a, hipStream_t s);
void launch_ke(const KeArgs &a, double *out, hipStream_t s);
void launch_interp_batch(int interp, const SchemeCoef<double> &c, const double *st, int n, double flag, double *out,
                         hipStream_t s);
void launch_riemann_batch(int rs, int dir, double g, const double *UL, const double *UR, int n, double *F,
                          hipStream_t s);

}  // namespace ns3d
