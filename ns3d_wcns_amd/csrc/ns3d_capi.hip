// ns3d_capi.hip -- the extern "C" boundary of libns3d_hip (include/ns3d_hip.h):
// context lifetime, device memory layout, data movement and the orchestration
// of the kernels behind each reference entry point.
//
// Device memory of one context (doubles):
//   20 ghosted cell fields (U, U0, Q = u v w p T, RHS), sx*sy*sz each, in the
//      reference's layout so that upload/download are plain copies;
//   15 compact face-flux fields (3 directions x 5 components), nx*ny*nz each;
//    9 compact viscous-flux fields, nx*ny*nz each;
//   12 compact gradient fields, allocated on first use of compute_gradients;
//   reduction partials (one per 64x4 block) and a few device scalars.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/ns3d_hip.h"
#include "ns3d_bounded.h"
#include "ns3d_internal.h"
#include "ns3d_kernels.h"
#include "ns3d_observe.h"

using namespace ns3d;

// Block maxima of the CFL reduction for ns3d_compute_timestep_noalloc: one
// slot per block of the first pass (at most REDUCE_BLOCKS blocks, whatever
// the box).  Each device holds its own copy until the process ends; every
// call on a device shares it, so calls that overlap on one device race.
// Nothing in it needs clearing: each block writes its slot before the
// second pass reads it.
static_assert(REDUCE_BLOCKS <= NS3D_NOALLOC_SCRATCH_DOUBLES, "CFL block maxima exceed the no-allocation scratch");
static __device__ __align__(256) double scratch[NS3D_NOALLOC_SCRATCH_DOUBLES];

struct ns3d_ctx {
    Box b;
    ns3d_grid grid;
    ns3d_params P;
    int precision;
    hipStream_t stream;
    size_t ncell, nint;
    double *slab;
    double *cell[NS3D_NUM_CELL_FIELDS];
    double *face[3][5];
    double *visc[NVISC];
    double *rz[5], *ry[5];  // compact partial RHS of the z and y pencils
    double *grad_slab;
    double *grad[12];
    // tangent / adjoint workspace (allocated on first use)
    double *adj_slab;
    double *xg[5];   // ghosted copy of the tangent / adjoint input, ghosts filled
    double *dq[4];   // ghosted (u, v, w, T) tangents
    double *s1[5];   // stage-1 state of the step being differentiated
    double *q2[4];   // ghosted u, v, w, T of its stage-2 state (adjoint step)
    double *a0[5], *a1[5], *pr[5];  // stage accumulators and the RHS derivative
    double *cw[25];  // compact: face adjoints and state adjoints, or face and viscous tangents
    double *gb[12], *qb[4];
    // observables work memory (allocated on first use) and the FFT plan
    double *obs_work;
    long obs_cap;
    void *fft_plan;
    double *partial;
    double *scal;  // [0] dt, [1] reduction result
    // CFL_SLOTS maxima that a device time loop's stage-3 x pencils fold the
    // new state's CFL speeds into (zero between uses)
    unsigned long long *cfl_slots;
    bool timing;
    bool visc_fp32;  // visc[] holds the mixed method's FP32 fluxes (last writer: a mixed fused stage)
    // the primitives of the state are the compact FP32 u, v, w, T - 1 of
    // prims32(): only between the stages of one rk3 / step_stages call of
    // the mixed method, and between the steps of one advance_steps call
    // (the last stage that call runs stores the FP64 primitives)
    bool prims_fp32;
    // a face is not periodic: the stepwise path of the bounded box runs
    // (ns3d_bounded.h), with the reference-layout face arrays fref
    bool bounded;
    int per[3];  // axis periodic
    // second stream of the fused stage, with NS3D_CONCURRENT_PENCILS=1 in the
    // environment at create: the z pencil runs on it beside the y pencil
    // (both read the stage's state and write partials of their own), forked
    // from and joined back into `stream` by the two events.  Null (the
    // default): the pencils run in order on `stream`, so each launch's time
    // is its own.
    hipStream_t side;
    hipEvent_t ev_fork, ev_join;
    double *fref_slab;
    double *fref[3][5];
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;  // (class, (start, stop))
};

namespace {

thread_local std::string g_err;

// Event bracket around one launch of a timed kernel class (no-op unless the
// context's kernel timing is enabled).
struct Timed {
    ns3d_ctx *c;
    int cls;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(ns3d_ctx *ctx, int k, hipStream_t stream) : c(ctx), cls(k), s(stream) {
        if (!c->timing) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, s);
    }
    Timed(ns3d_ctx *ctx, int k) : Timed(ctx, k, ctx->stream) {}
    ~Timed() {
        if (!a) return;
        (void)hipEventRecord(b, s);
        c->events.push_back({cls, {a, b}});
    }
};

// The side stream and its events (those created so far).
void destroy_side(ns3d_ctx *c) {
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    c->side = nullptr;
    c->ev_fork = c->ev_join = nullptr;
}

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define NS3D_HIP(call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(NS3D_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define NS3D_LAUNCHED(what)                                                                          \
    do {                                                                                             \
        hipError_t e_ = hipGetLastError();                                                           \
        if (e_ != hipSuccess) return fail(NS3D_E_HIP, "%s launch: %s", what, hipGetErrorString(e_)); \
    } while (0)

#define NS3D_CTX(c)                                                     \
    do {                                                                \
        if (!(c)) return fail(NS3D_E_ARG, "%s: null context", __func__); \
    } while (0)

int check_box(const ns3d_grid *grid) {
    if (grid->ng != NG) return fail(NS3D_E_UNSUPPORTED, "ghost layers must be %d (got %d)", NG, grid->ng);
    if (grid->nx < 6 || grid->ny < 6 || grid->nz < 6)
        return fail(NS3D_E_ARG, "each interior extent must be >= 6 (got %d x %d x %d)", grid->nx, grid->ny,
                    grid->nz);
    const double ghosted = (double)(grid->nx + 2 * NG) * (grid->ny + 2 * NG) * (grid->nz + 2 * NG);
    if (ghosted >= 2147483647.0)
        return fail(NS3D_E_ARG, "box of %g cells exceeds the reference's 32-bit cell index", ghosted);
    if (!(grid->dx > 0 && grid->dy > 0 && grid->dz > 0)) return fail(NS3D_E_ARG, "grid spacing must be positive");
    return NS3D_OK;
}

Box box_of(const ns3d_grid &g) {
    Box b;
    b.nx = g.nx; b.ny = g.ny; b.nz = g.nz;
    b.sx = b.nx + 2 * NG; b.sy = b.ny + 2 * NG; b.sz = b.nz + 2 * NG;
    return b;
}

// Per-device address of the allocation-free entry point's scratch array
// (hipGetSymbolAddress, once per device with that device current; the
// array's name is a host-side stand-in and is never passed to a kernel).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(double **out) {
    int dev = -1;
    NS3D_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(NS3D_E_UNSUPPORTED, "device ordinal %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NS3D_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return NS3D_OK;
}

int check_params(const ns3d_params *p) {
    if (!p) return fail(NS3D_E_ARG, "null params");
    if (p->interpolation == NS3D_INTERP_ZERO)
        return fail(NS3D_E_UNSUPPORTED,
                    "interpolation ZERO is not supported: its 2-point stencil sets one ghost layer while the "
                    "6th-order half-node difference reads three");
    if (p->interpolation != NS3D_INTERP_WENO5 && p->interpolation != NS3D_INTERP_MDCD_LINEAR &&
        p->interpolation != NS3D_INTERP_MDCD_HYBRID && p->interpolation != NS3D_INTERP_WENO5_JS)
        return fail(NS3D_E_ARG, "unknown interpolation %d", p->interpolation);
    if (p->riemann_solver < NS3D_RIEMANN_ROE || p->riemann_solver > NS3D_RIEMANN_AUSM)
        return fail(NS3D_E_ARG, "Unknown Riemann solver %d", p->riemann_solver);
    if (!(p->gamma > 1.0)) return fail(NS3D_E_ARG, "gamma must exceed 1 (got %g)", p->gamma);
    static const char *face[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    for (int f = 0; f < 6; ++f)
        if (p->bc[f] < NS3D_BC_PERIODIC || p->bc[f] > NS3D_BC_INFLOW)
            return fail(NS3D_E_ARG, "unknown boundary type %d on bc_%s", p->bc[f], face[f]);
    for (int a = 0; a < 3; ++a)
        if ((p->bc[2 * a] == NS3D_BC_PERIODIC) != (p->bc[2 * a + 1] == NS3D_BC_PERIODIC))
            return fail(NS3D_E_UNSUPPORTED,
                        "bc_%s and bc_%s: a Periodic face opposite a non-periodic one is not supported (the "
                        "reference leaves that face's ghosts at zero)",
                        face[2 * a], face[2 * a + 1]);
    return NS3D_OK;
}

bool any_physical(const ns3d_params &p) {
    for (int f = 0; f < 6; ++f)
        if (p.bc[f] != NS3D_BC_PERIODIC) return true;
    return false;
}

SchemeCoef<double> coef64(const ns3d_params &P) {
    SchemeCoef<double> c;
    const double diss = P.mdcd_diss, disp = P.mdcd_disp;
    c.gamma = P.gamma;
    c.p_ref = 0.0;
    c.H_ref = 0.0;
    c.E_ref = 0.0;
    // linear MDCD weights, interpolation.cpp:128-132
    c.mdcd_lin[0] = (3.0 * disp + 3.0 * diss) / 8.0;
    c.mdcd_lin[1] = (-18.0 * disp - 30 * diss - 1.0) / 16.0;
    c.mdcd_lin[2] = (12.0 * disp + 60.0 * diss + 9.0) / 16.0;
    c.mdcd_lin[3] = (12.0 * disp - 60.0 * diss + 9.0) / 16.0;
    c.mdcd_lin[4] = (-18.0 * disp + 30.0 * diss - 1.0) / 16.0;
    c.mdcd_lin[5] = (3.0 * disp - 3.0 * diss) / 8.0;
    // MDCD-WENO linear weights, interpolation.cpp:182-185
    c.mdcd_d[0] = 1.5 * (disp + diss);
    c.mdcd_d[1] = 0.5 - 1.5 * (disp - 3 * diss);
    c.mdcd_d[2] = 0.5 - 1.5 * (disp + 3 * diss);
    c.mdcd_d[3] = 1.5 * (disp - diss);
    return c;
}

SchemeCoef<float> coef32(const ns3d_params &P) {
    const SchemeCoef<double> d = coef64(P);
    SchemeCoef<float> c;
    c.gamma = (float)d.gamma;
    // perturbation form: p_ref = rho_ref T_ref Rgas with rho_ref = T_ref = 1
    // (the nondimensional reference state), E_ref = p_ref / (gamma - 1)
    const double pr = P.Rgas, er = pr / (P.gamma - 1.0);
    c.p_ref = (float)pr;
    c.H_ref = (float)(er + pr);
    c.E_ref = (float)er;
    for (int q = 0; q < 6; ++q) c.mdcd_lin[q] = (float)d.mdcd_lin[q];
    for (int q = 0; q < 4; ++q) c.mdcd_d[q] = (float)d.mdcd_d[q];
    return c;
}

// Reference state of the mixed method's perturbation form (zero when plain).
MixedRef mixed_ref(const ns3d_ctx *c) {
    if (c->precision != NS3D_PRECISION_MIXED) return MixedRef{0.0, 0.0};
    return MixedRef{c->P.Rgas / (c->P.gamma - 1.0), c->P.Rgas};
}

double **conserved(ns3d_ctx *c) { return &c->cell[NS3D_RHO]; }
double **recorded(ns3d_ctx *c) { return &c->cell[NS3D_RHO0]; }
double **prims(ns3d_ctx *c) { return &c->cell[NS3D_U]; }
double **rhs(ns3d_ctx *c) { return &c->cell[NS3D_RHS_RHO]; }
// Compact FP32 u, v, w, T - 1 handed between the stages of the mixed method:
// the second half of the z partial buffers, whose first half holds that
// method's FP32 partial sums (nint floats; a buffer holds field_stride(nint)
// doubles, at least twice the 64-rounded nint floats)
float *prims32(ns3d_ctx *c, int q) {
    return reinterpret_cast<float *>(c->rz[q]) + ((c->nint + 63) / 64) * 64;
}

int run_faces(ns3d_ctx *c) {
    FaceArgs<double> a64{};
    FaceArgs<float> a32{};
    for (int n = 0; n < 5; ++n) a64.U[n] = a32.U[n] = conserved(c)[n];
    for (int n = 0; n < 4; ++n) a64.Q[n] = a32.Q[n] = prims(c)[n];
    a64.b = a32.b = c->b;
    a64.c = coef64(c->P);
    a32.c = coef32(c->P);
    a64.ref = a32.ref = MixedRef{0.0, 0.0};
    for (int dir = 0; dir < 3; ++dir) {
        for (int n = 0; n < 5; ++n) a64.F[n] = a32.F[n] = c->face[dir][n];
        const int interp = c->P.interpolation;
        const bool chr = c->P.char_recon != 0;
        switch (c->P.riemann_solver) {
            case R_ROE: launch_face<R_ROE>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
            case R_RUSANOV: launch_face<R_RUSANOV>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
            case R_HLLC: launch_face<R_HLLC>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
            case R_HLL: launch_face<R_HLL>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
            case R_HLLC_P: launch_face<R_HLLC_P>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
            default: launch_face<R_AUSM>(dir, interp, chr, c->precision, a64, a32, c->stream); break;
        }
        NS3D_LAUNCHED("face flux");
    }
    return NS3D_OK;
}

ViscArgs visc_args(ns3d_ctx *c, bool with_grad, bool with_flux) {
    ViscArgs a{};
    const double *q[4] = {c->cell[NS3D_U], c->cell[NS3D_V], c->cell[NS3D_W], c->cell[NS3D_T]};
    for (int n = 0; n < 4; ++n) a.Q[n] = q[n];
    for (int n = 0; n < 12; ++n) a.G[n] = with_grad ? c->grad[n] : nullptr;
    for (int n = 0; n < NVISC; ++n) a.V[n] = with_flux ? c->visc[n] : nullptr;
    a.b = c->b;
    a.i60[0] = 1.0 / (60.0 * c->grid.dx);
    a.i60[1] = 1.0 / (60.0 * c->grid.dy);
    a.i60[2] = 1.0 / (60.0 * c->grid.dz);
    a.mu = c->P.mu;
    a.S_ref = c->P.S_ref;
    a.cp_pr = c->P.Cp / c->P.Pr;
    return a;
}

AsmArgs asm_args(ns3d_ctx *c, int stage, int mode, const double *d_dt, double dt) {
    AsmArgs a{};
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) a.F[d][n] = c->face[d][n];
    for (int n = 0; n < NVISC; ++n) a.V[n] = c->visc[n];
    for (int n = 0; n < 5; ++n) {
        a.U[n] = conserved(c)[n];
        a.U0[n] = recorded(c)[n];
        a.Q[n] = prims(c)[n];
        a.RHS[n] = rhs(c)[n];
    }
    a.b = c->b;
    a.ih[0] = 1.0 / c->grid.dx;
    a.ih[1] = 1.0 / c->grid.dy;
    a.ih[2] = 1.0 / c->grid.dz;
    for (int d = 0; d < 3; ++d) a.i60[d] = a.ih[d] / 60.0;
    a.gamma = c->P.gamma;
    a.Rgas = c->P.Rgas;
    a.dt = d_dt;
    a.dtv = dt;
    a.stage = stage;
    a.mode = mode;
    return a;
}

ConvArgs conv_args(ns3d_ctx *c) {
    ConvArgs a{};
    for (int n = 0; n < 5; ++n) {
        a.U[n] = conserved(c)[n];
        a.Q[n] = prims(c)[n];
    }
    a.n = (long)c->ncell;
    a.gamma = c->P.gamma;
    a.Rgas = c->P.Rgas;
    return a;
}

int bnd_boundary(ns3d_ctx *c);

int halo(ns3d_ctx *c, bool all) {
    // a bounded box keeps its ghost layers consistent after every stage; its
    // apply_boundary is the exchange plus the physical faces
    if (c->bounded) return all ? NS3D_OK : bnd_boundary(c);
    HaloArgs h{};
    h.b = c->b;
    if (all) {  // conserved and primitives after a fused stage
        for (int n = 0; n < 5; ++n) h.f[n] = conserved(c)[n];
        for (int n = 0; n < 5; ++n) h.f[5 + n] = prims(c)[n];
        h.nf = 10;
    } else {  // exchange_halos_physical: rho, u, v, w, p
        h.f[0] = c->cell[NS3D_RHO];
        for (int n = 0; n < 4; ++n) h.f[1 + n] = prims(c)[n];
        h.nf = 5;
    }
    {
        Timed t(c, NS3D_K_HALO);
        launch_halo(h, c->stream);
    }
    NS3D_LAUNCHED("halo");
    return NS3D_OK;
}

// State arrays of a fused stage other than the context's own: the stage reads
// Uin (its primitives in the context, or Qin when set), writes Uout (its
// primitives into the context's, or into Qout when set) and takes U0 as the
// recorded state (stage 1 with U0_READ, stages 2 and 3; a stage 3 may write
// Uout = U0: each thread reads U0 of a cell before it writes that cell).  Qin / Qout hold
// u, v, w, p, T in slots 0..4; slot 3 (p) is neither read nor written by
// stages 1 and 2 and may be null.
struct StageIO {
    const double *const *Uin;
    double *const *Uout;
    double *const *U0;
    const double *const *Qin = nullptr;
    double *const *Qout = nullptr;
};

// One fused stage (stage 0: RHS only): viscous fluxes, then the z, y and x
// pencils; stages 1..3 update the state in place, or from io->Uin into
// io->Uout.  The kernels read no ghost layer (periodic images instead); with
// ghosts the layers of the new state (the context's) are refilled for the
// callers that read them.  q32 (mixed method, set only by callers whose next
// step is another fused stage of the same state: stages 1 and 2 of a step,
// and stage 3 of a device time loop's steps other than the last): the x
// pencil hands the primitives to the next stage's viscous kernel in FP32.
int fused_stage(ns3d_ctx *c, int stage, const double *d_dt, double dt, int u0mode = U0_READ, bool ghosts = true,
                const StageIO *io = nullptr, bool q32 = false, bool fold_cfl = false) {
    const bool mixed = c->precision == NS3D_PRECISION_MIXED;
    {
        ViscArgs va = visc_args(c, false, true);
        if (io && io->Qin) {
            const int slot[4] = {0, 1, 2, 4};
            for (int q = 0; q < 4; ++q) va.Q[q] = io->Qin[slot[q]];
        }
        for (int q = 0; q < 4; ++q) va.Q32[q] = mixed && c->prims_fp32 ? prims32(c, q) : nullptr;
        Timed t(c, NS3D_K_VISCOUS);
        launch_visc_march(va, c->precision, c->stream);
    }
    NS3D_LAUNCHED("viscous");
    c->visc_fp32 = c->precision == NS3D_PRECISION_MIXED;
    PencilArgs a{};
    for (int n = 0; n < 5; ++n) {
        a.U[n] = io ? io->Uin[n] : conserved(c)[n];
        a.RHS[n] = rhs(c)[n];
        a.U0[n] = io ? io->U0[n] : recorded(c)[n];
        a.Unew[n] = io ? io->Uout[n] : conserved(c)[n];  // in place by default (the x pencil owns its line)
        a.Q[n] = io && io->Qout ? io->Qout[n] : prims(c)[n];
        a.Rz[n] = c->rz[n];
        a.Ry[n] = c->ry[n];
    }
    for (int n = 0; n < NVISC; ++n) a.V[n] = c->visc[n];
    a.b = c->b;
    a.ref = mixed_ref(c);
    a.record = u0mode;
    a.ihd[0] = 1.0 / c->grid.dx;
    a.ihd[1] = 1.0 / c->grid.dy;
    a.ihd[2] = 1.0 / c->grid.dz;
    a.gamma = c->P.gamma;
    a.Rgas = c->P.Rgas;
    a.dt = d_dt;
    a.dtv = dt;
    a.stage = stage;
    for (int q = 0; q < 4; ++q) a.Q32[q] = q32 && mixed && stage >= 1 ? prims32(c, q) : nullptr;
    // stage 3 of a device time loop's step that another step follows: the x
    // pencil folds the new state's CFL speeds into the slots the next step's
    // dt is taken from
    a.cfl_slots = fold_cfl && stage == 3 ? c->cfl_slots : nullptr;
    const SchemeCoef<double> c64 = coef64(c->P);
    const SchemeCoef<float> c32 = coef32(c->P);
    const int interp = c->P.interpolation;
    const bool chr = c->P.char_recon != 0;
    auto pencils = [&](int dir, hipStream_t st) {
        switch (c->P.riemann_solver) {
            case R_ROE: launch_pencils<R_ROE>(dir, interp, chr, c->precision, a, c64, c32, st); break;
            case R_RUSANOV: launch_pencils<R_RUSANOV>(dir, interp, chr, c->precision, a, c64, c32, st); break;
            case R_HLLC: launch_pencils<R_HLLC>(dir, interp, chr, c->precision, a, c64, c32, st); break;
            case R_HLL: launch_pencils<R_HLL>(dir, interp, chr, c->precision, a, c64, c32, st); break;
            case R_HLLC_P: launch_pencils<R_HLLC_P>(dir, interp, chr, c->precision, a, c64, c32, st); break;
            default: launch_pencils<R_AUSM>(dir, interp, chr, c->precision, a, c64, c32, st); break;
        }
    };
    // z on the side stream beside y on the context's stream (independent:
    // both read U and the viscous fluxes, each writes its own partials); the
    // x pencil, which reads both partials, waits for the join
    {
        hipStream_t zs = c->side ? c->side : c->stream;
        if (c->side) {
            NS3D_HIP(hipEventRecord(c->ev_fork, c->stream));
            NS3D_HIP(hipStreamWaitEvent(zs, c->ev_fork, 0));
        }
        Timed t(c, NS3D_K_PENCIL_Z, zs);
        pencils(2, zs);
    }
    {
        Timed t(c, NS3D_K_PENCIL_Y);
        pencils(1, c->stream);
    }
    if (c->side) {
        NS3D_HIP(hipEventRecord(c->ev_join, c->side));
        NS3D_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    {
        Timed t(c, NS3D_K_PENCIL_X);
        pencils(0, c->stream);
    }
    NS3D_LAUNCHED("pencils");
    if (stage >= 1) c->prims_fp32 = a.Q32[0] != nullptr;
    if (stage == 0 || !ghosts) return NS3D_OK;
    if (io && io->Uout[0] != conserved(c)[0]) return fail(NS3D_E_ARG, "internal: ghost refill of a foreign state");
    return halo(c, true);
}

// record: stage 1 also stores U0 (interior cells) from the state it reads.
// The intermediate stage states are not handed out, so only the last stage
// refills the ghost layers, and only with ghosts set: the device time loops
// run their steps without (no kernel of a step reads a ghost layer) and
// refill once before they return.  The mixed method hands stages 1 and 2's
// primitives to the next viscous kernel in FP32, and stage 3's too with
// q32_last (a device time loop's step that another step follows).
//
// through (a device time loop's step that another step follows, in place of
// record): U^n stays in the state arrays until stage 3 and is itself the
// stages' U0, so no U0 copy is stored.  Stage 1 reads U^n and writes U^(1)
// into the RHS arrays (which no kernel of the device loops reads otherwise),
// stage 2 updates them in place to U^(2), and stage 3 reads U^(2) there and
// writes U^(n+1) over U^n, each cell's U0 read before that cell is written.
// The results are those of the recorded step; the U0 arrays keep their
// previous contents and the RHS arrays hold U^(2) afterwards.
int record0(ns3d_ctx *c);
int bnd_rk3(ns3d_ctx *c, const double *d_dt, double dt);

int rk3(ns3d_ctx *c, const double *d_dt, double dt, bool record = false, bool ghosts = true,
        bool q32_last = false, bool through = false, bool fold_cfl = false) {
    int rc;
    if (c->bounded) {
        if ((record || through) && (rc = record0(c))) return rc;
        return bnd_rk3(c, d_dt, dt);
    }
    double *const *A = conserved(c);
    double *const *B = rhs(c);
    const StageIO io[3] = {{A, B, A}, {B, B, A}, {B, A, A}};
    for (int stage = 1; stage <= 3; ++stage) {
        const int u0mode = stage > 1 ? U0_READ : through ? U0_INPUT : record ? U0_RECORD : U0_READ;
        if ((rc = fused_stage(c, stage, d_dt, dt, u0mode, ghosts && stage == 3, through ? &io[stage - 1] : nullptr,
                              stage < 3 || q32_last, fold_cfl))) {
            c->prims_fp32 = false;
            return rc;
        }
    }
    return NS3D_OK;
}

int record0(ns3d_ctx *c) {
    for (int n = 0; n < 5; ++n)
        NS3D_HIP(hipMemcpyAsync(recorded(c)[n], conserved(c)[n], c->ncell * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
    return NS3D_OK;
}

// ------------------------------------------------------------ bounded box
// The reference's pass structure on a box with physical faces
// (range_kutta.cpp:11-92, boundary.cpp, ViscousFlux.cpp:15-255).
BndArgs bnd_args(ns3d_ctx *c) {
    BndArgs a{};
    a.P[0] = c->cell[NS3D_RHO];
    for (int n = 0; n < 4; ++n) a.P[1 + n] = prims(c)[n];  // u, v, w, p
    a.b = c->b;
    for (int d = 0; d < 3; ++d) a.per[d] = c->per[d];
    return a;
}

// apply_boundary (boundary.cpp:14-54): the face-ghost exchange of rho, u, v,
// w, p, then each physical face, xmin to zmax, over its full planes.
int bnd_boundary(ns3d_ctx *c) {
    BndArgs a = bnd_args(c);
    launch_bnd_exchange(a, c->stream);
    for (int f = 0; f < 6; ++f) {
        if (c->per[f / 2] || c->P.bc[f] == NS3D_BC_PERIODIC) continue;
        a.face = f;
        a.type = c->P.bc[f];
        launch_bnd_fill(a, c->stream);
    }
    NS3D_LAUNCHED("boundary");
    return NS3D_OK;
}

// compute_invis_flux (InviscousFlux.cpp:10-144): faces -1/2 .. n-1/2 of every
// interior line, reference face layout.
int bnd_faces(ns3d_ctx *c) {
    FaceArgs<double> a{};
    for (int n = 0; n < 5; ++n) a.U[n] = conserved(c)[n];
    for (int n = 0; n < 4; ++n) a.Q[n] = prims(c)[n];
    a.b = c->b;
    a.c = coef64(c->P);
    a.ref = MixedRef{0.0, 0.0};
    const int interp = c->P.interpolation;
    const bool chr = c->P.char_recon != 0;
    for (int dir = 0; dir < 3; ++dir) {
        for (int n = 0; n < 5; ++n) a.F[n] = c->fref[dir][n];
        switch (c->P.riemann_solver) {
            case R_ROE: launch_face_ref<R_ROE>(dir, interp, chr, a, c->stream); break;
            case R_RUSANOV: launch_face_ref<R_RUSANOV>(dir, interp, chr, a, c->stream); break;
            case R_HLLC: launch_face_ref<R_HLLC>(dir, interp, chr, a, c->stream); break;
            case R_HLL: launch_face_ref<R_HLL>(dir, interp, chr, a, c->stream); break;
            case R_HLLC_P: launch_face_ref<R_HLLC_P>(dir, interp, chr, a, c->stream); break;
            default: launch_face_ref<R_AUSM>(dir, interp, chr, a, c->stream); break;
        }
    }
    NS3D_LAUNCHED("face flux (bounded box)");
    return NS3D_OK;
}

// apply_boundary_halfnode_flux (boundary.cpp:207-300): the ghost faces.
int bnd_ghost_faces(ns3d_ctx *c) {
    for (int dir = 0; dir < 3; ++dir) {
        GhostFaceArgs g{};
        for (int n = 0; n < 5; ++n) g.F[n] = c->fref[dir][n];
        g.b = c->b;
        g.dir = dir;
        g.per = c->per[dir];
        g.type_lo = c->P.bc[2 * dir];
        g.type_hi = c->P.bc[2 * dir + 1];
        launch_ghost_faces(g, c->stream);
    }
    NS3D_LAUNCHED("ghost faces");
    return NS3D_OK;
}

// compute_gradients / compute_viscous_flux (ViscousFlux.cpp:109-208) with the
// near-face orders.
int bnd_visc(ns3d_ctx *c, bool grads, bool flux) {
    BndViscArgs a{};
    const double *q[4] = {c->cell[NS3D_U], c->cell[NS3D_V], c->cell[NS3D_W], c->cell[NS3D_T]};
    for (int n = 0; n < 4; ++n) a.Q[n] = q[n];
    for (int n = 0; n < 12; ++n) a.G[n] = grads ? c->grad[n] : nullptr;
    for (int n = 0; n < NVISC; ++n) a.V[n] = flux ? c->visc[n] : nullptr;
    a.b = c->b;
    for (int d = 0; d < 3; ++d) a.per[d] = c->per[d];
    a.h[0] = c->grid.dx;
    a.h[1] = c->grid.dy;
    a.h[2] = c->grid.dz;
    a.mu = c->P.mu;
    a.S_ref = c->P.S_ref;
    a.cp_pr = c->P.Cp / c->P.Pr;
    launch_bnd_visc(a, c->stream);
    NS3D_LAUNCHED("viscous (bounded box)");
    if (flux) c->visc_fp32 = false;
    return NS3D_OK;
}

// compute_invis_dflux / compute_vis_flux (InviscousFlux.cpp:1027-1056,
// ViscousFlux.cpp:210-255).
int bnd_assemble(ns3d_ctx *c, int mode) {
    BndRhsArgs a{};
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) a.F[d][n] = c->fref[d][n];
    for (int n = 0; n < NVISC; ++n) a.V[n] = c->visc[n];
    for (int n = 0; n < 5; ++n) a.RHS[n] = rhs(c)[n];
    a.b = c->b;
    for (int d = 0; d < 3; ++d) a.per[d] = c->per[d];
    a.h[0] = c->grid.dx;
    a.h[1] = c->grid.dy;
    a.h[2] = c->grid.dz;
    a.mode = mode;
    launch_bnd_rhs(a, c->stream);
    NS3D_LAUNCHED("rhs (bounded box)");
    return NS3D_OK;
}

// compute_rhs (range_kutta.cpp:11-43) of the interior cells.
int bnd_rhs(ns3d_ctx *c) {
    int rc;
    if ((rc = bnd_faces(c)) || (rc = bnd_ghost_faces(c)) || (rc = bnd_visc(c, false, true))) return rc;
    return bnd_assemble(c, ASM_INVISCID | ASM_VISCOUS);
}

// runge_kutta_3 (range_kutta.cpp:46-92): each stage on every cell, then
// conservedToPrimitive -> apply_boundary -> primitiveToConserved.
int bnd_rk3(ns3d_ctx *c, const double *d_dt, double dt) {
    int rc;
    for (int stage = 1; stage <= 3; ++stage) {
        if ((rc = bnd_rhs(c))) return rc;
        RkAllArgs r{};
        for (int n = 0; n < 5; ++n) {
            r.U[n] = conserved(c)[n];
            r.U0[n] = recorded(c)[n];
            r.R[n] = rhs(c)[n];
        }
        r.b = c->b;
        r.dt = d_dt;
        r.dtv = dt;
        r.stage = stage;
        launch_rk_all(r, c->stream);
        launch_c2p(conv_args(c), c->stream);
        NS3D_LAUNCHED("stage (bounded box)");
        if ((rc = bnd_boundary(c))) return rc;
        launch_p2c(conv_args(c), c->stream);
        NS3D_LAUNCHED("primitiveToConserved");
    }
    return NS3D_OK;
}

// dt = cfl * min(1e9, dx / max(|u| + a)) of the conserved state into the
// device scalar d_out (compute_timestep, timestep.cpp:5-23).
int cfl(ns3d_ctx *c, double *d_out) {
    CflArgs a{};
    for (int n = 0; n < 5; ++n) a.U[n] = conserved(c)[n];
    a.partial = c->partial;
    a.dt_out = d_out;
    a.b = c->b;
    a.gamma = c->P.gamma;
    a.dx = c->grid.dx;
    a.cfl = c->P.cfl;
    Timed t(c, NS3D_K_CFL);
    launch_cfl(a, c->stream);
    NS3D_LAUNCHED("cfl");
    return NS3D_OK;
}

// The same dt from the maxima the previous step's stage-3 x pencils folded
// into c->cfl_slots (bit-identical: the same speeds of the same stored
// state, and a maximum does not depend on the order), clearing the slots.
int cfl_folded(ns3d_ctx *c, double *d_out) {
    CflArgs a{};
    a.partial = reinterpret_cast<double *>(c->cfl_slots);
    a.dt_out = d_out;
    a.b = c->b;
    a.gamma = c->P.gamma;
    a.dx = c->grid.dx;
    a.cfl = c->P.cfl;
    Timed t(c, NS3D_K_CFL);
    launch_cfl_final_slots(a, c->stream);
    NS3D_LAUNCHED("cfl (folded)");
    return NS3D_OK;
}

// compact -> reference layout scatter of one periodic field (host side)
void scatter_cells(const ns3d_ctx *c, const std::vector<double> &compact, double *host, bool face_ghosts) {
    const Box &b = c->b;
    std::memset(host, 0, c->ncell * sizeof(double));
    auto at = [&](int i, int j, int k) { return compact[b.oidx(i, j, k)]; };
    for (int k = 0; k < b.nz; ++k)
        for (int j = 0; j < b.ny; ++j)
            for (int i = 0; i < b.nx; ++i) host[b.gidx(i, j, k)] = at(i, j, k);
    if (!face_ghosts) return;
    // a non-periodic axis keeps the zeros of the reference's exchange
    auto w = [](int i, int n) { return ((i % n) + n) % n; };
    if (c->per[0])
    for (int k = 0; k < b.nz; ++k)
        for (int j = 0; j < b.ny; ++j)
            for (int g = 1; g <= NG; ++g) {
                host[b.gidx(-g, j, k)] = at(w(-g, b.nx), j, k);
                host[b.gidx(b.nx - 1 + g, j, k)] = at(w(b.nx - 1 + g, b.nx), j, k);
            }
    if (c->per[1])
    for (int k = 0; k < b.nz; ++k)
        for (int i = 0; i < b.nx; ++i)
            for (int g = 1; g <= NG; ++g) {
                host[b.gidx(i, -g, k)] = at(i, w(-g, b.ny), k);
                host[b.gidx(i, b.ny - 1 + g, k)] = at(i, w(b.ny - 1 + g, b.ny), k);
            }
    if (c->per[2])
    for (int j = 0; j < b.ny; ++j)
        for (int i = 0; i < b.nx; ++i)
            for (int g = 1; g <= NG; ++g) {
                host[b.gidx(i, j, -g)] = at(i, j, w(-g, b.nz));
                host[b.gidx(i, j, b.nz - 1 + g)] = at(i, j, w(b.nz - 1 + g, b.nz));
            }
}

int fetch_compact(ns3d_ctx *c, const double *dptr, std::vector<double> &out) {
    out.resize(c->nint);
    NS3D_HIP(hipMemcpyAsync(out.data(), dptr, c->nint * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

size_t field_stride(size_t n) { return ((n + 63) / 64) * 64 + 320; }

int ensure_grad(ns3d_ctx *c) {
    if (c->grad_slab) return NS3D_OK;
    const size_t istride = field_stride(c->nint);
    if (hipMalloc(&c->grad_slab, 12 * istride * sizeof(double)) != hipSuccess) {
        c->grad_slab = nullptr;
        (void)hipGetLastError();
        return fail(NS3D_E_NOMEM, "gradient arrays: out of device memory");
    }
    for (int q = 0; q < 12; ++q) c->grad[q] = c->grad_slab + q * istride;
    NS3D_HIP(hipMemsetAsync(c->grad_slab, 0, 12 * istride * sizeof(double), c->stream));
    return NS3D_OK;
}

// ------------------------------------------------------------ derivatives
int ensure_adj(ns3d_ctx *c) {
    if (c->adj_slab) return NS3D_OK;
    const size_t cs = field_stride(c->ncell), is = field_stride(c->nint);
    constexpr size_t kGhosted = 5 + 4 + 5 + 4 + 5 + 5 + 5;  // xg, dq, s1, q2, a0, a1, pr
    constexpr size_t kCompact = 25 + 12 + 4;            // cw, gb, qb
    const size_t total = kGhosted * cs + kCompact * is;
    if (hipMalloc(&c->adj_slab, total * sizeof(double)) != hipSuccess) {
        c->adj_slab = nullptr;
        (void)hipGetLastError();
        return fail(NS3D_E_NOMEM, "tangent/adjoint workspace: out of device memory for %zu doubles", total);
    }
    double *p = c->adj_slab;
    for (int n = 0; n < 5; ++n, p += cs) c->xg[n] = p;
    for (int n = 0; n < 4; ++n, p += cs) c->dq[n] = p;
    for (int n = 0; n < 5; ++n, p += cs) c->s1[n] = p;
    for (int n = 0; n < 4; ++n, p += cs) c->q2[n] = p;
    for (int n = 0; n < 5; ++n, p += cs) c->a0[n] = p;
    for (int n = 0; n < 5; ++n, p += cs) c->a1[n] = p;
    for (int n = 0; n < 5; ++n, p += cs) c->pr[n] = p;
    for (int n = 0; n < 25; ++n, p += is) c->cw[n] = p;
    for (int n = 0; n < 12; ++n, p += is) c->gb[n] = p;
    for (int n = 0; n < 4; ++n, p += is) c->qb[n] = p;
    if ((size_t)(p - c->adj_slab) != total) {  // the carve must match the allocation
        (void)hipFree(c->adj_slab);
        c->adj_slab = nullptr;
        return fail(NS3D_E_HIP, "internal: tangent/adjoint workspace layout mismatch");
    }
    NS3D_HIP(hipMemsetAsync(c->adj_slab, 0, total * sizeof(double), c->stream));
    return NS3D_OK;
}

// The context's primitives from the state U (interior cells are the ones
// read; U is not written).
int prims_from(ns3d_ctx *c, const double *const *U) {
    ConvArgs a = conv_args(c);
    for (int n = 0; n < 5; ++n) a.U[n] = const_cast<double *>(U[n]);
    a.keep = 1;
    launch_c2p(a, c->stream);
    NS3D_LAUNCHED("c2p");
    return NS3D_OK;
}

// Primitives and ghost layers consistent with the conserved state.
int refresh_state(ns3d_ctx *c) {
    launch_c2p(conv_args(c), c->stream);
    NS3D_LAUNCHED("c2p");
    if (!c->bounded) return halo(c, true);
    int rc = bnd_boundary(c);
    if (rc) return rc;
    launch_p2c(conv_args(c), c->stream);
    NS3D_LAUNCHED("primitiveToConserved");
    return NS3D_OK;
}

int fill_ghosts(ns3d_ctx *c, double *const *f, int nf) {
    HaloArgs h{};
    h.b = c->b;
    for (int n = 0; n < nf; ++n) h.f[n] = f[n];
    h.nf = nf;
    launch_halo(h, c->stream);
    NS3D_LAUNCHED("halo");
    return NS3D_OK;
}

int copy_fields(ns3d_ctx *c, double *const *dst, const double *const *src, int nf) {
    for (int n = 0; n < nf; ++n)
        NS3D_HIP(hipMemcpyAsync(dst[n], src[n], c->ncell * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return NS3D_OK;
}

int zero_fields(ns3d_ctx *c, double *const *f, int nf) {
    for (int n = 0; n < nf; ++n) NS3D_HIP(hipMemsetAsync(f[n], 0, c->ncell * sizeof(double), c->stream));
    return NS3D_OK;
}

AdjArgs adj_args(ns3d_ctx *c, double *const *out) {
    AdjArgs a{};
    for (int n = 0; n < 5; ++n) {
        a.U[n] = conserved(c)[n];
        a.Q[n] = prims(c)[n];
        a.X[n] = c->xg[n];
        a.out[n] = out[n];
    }
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) a.FB[d][n] = c->cw[5 * d + n];
    for (int n = 0; n < 10; ++n) a.SB[n] = c->cw[15 + n];
    for (int n = 0; n < 12; ++n) a.Gb[n] = c->gb[n];
    for (int n = 0; n < 4; ++n) {
        a.qb[n] = c->qb[n];
        a.dQ[n] = c->dq[n];
    }
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) a.dF[d][n] = c->cw[5 * d + n];
    for (int n = 0; n < NVISC; ++n) a.dV[n] = c->cw[15 + n];
    a.b = c->b;
    a.c = coef64(c->P);
    const double h[3] = {c->grid.dx, c->grid.dy, c->grid.dz};
    for (int d = 0; d < 3; ++d) {
        a.ih[d] = 1.0 / h[d];
        a.i60[d] = 1.0 / (60.0 * h[d]);
    }
    a.gamma = c->P.gamma;
    a.Rgas = c->P.Rgas;
    a.mu = c->P.mu;
    a.S_ref = c->P.S_ref;
    a.cp_pr = c->P.Cp / c->P.Pr;
    return a;
}

// Tangent RHS (interior of out) at the current state, of the tangent X
// (ghosts filled).
int tangent_rhs_impl(ns3d_ctx *c, const double *const *X, double *const *out) {
    int rc;
    AdjArgs a = adj_args(c, out);
    for (int n = 0; n < 5; ++n) a.X[n] = X[n];
    launch_tl_prim(a, c->stream);
    NS3D_LAUNCHED("tangent primitives");
    if ((rc = fill_ghosts(c, c->dq, 4))) return rc;
    launch_tl_rhs(a, c->P.interpolation, c->P.riemann_solver, c->P.char_recon != 0, c->stream);
    NS3D_LAUNCHED("tangent rhs");
    return NS3D_OK;
}

// Adjoint RHS (interior of out) at the current state, of the adjoint X
// (interior read; stencils take periodic images).
// The adjoint RHS about the state U (the context's own by default), whose
// primitives the context holds.
// With a tail, the RHS is not stored: its value at each cell goes into the
// tail's combinations.
// U, Q: the state and its primitives (slots 0..4; p is not read) when not
// the context's.
int adjoint_rhs_impl(ns3d_ctx *c, const double *const *X, double *const *out, const double *const *U = nullptr,
                     const LinTail *tail = nullptr, const double *const *Q = nullptr) {
    AdjArgs a = adj_args(c, out);
    for (int n = 0; n < 5; ++n) {
        a.X[n] = X[n];
        if (U) a.U[n] = U[n];
        if (Q) a.Q[n] = Q[n];
    }
    LinTail t{};
    if (tail) t = *tail;
    launch_adj_rhs(a, c->P.interpolation, c->P.riemann_solver, c->stream, t);
    NS3D_LAUNCHED("adjoint rhs");
    return NS3D_OK;
}

// A caller's input into the workspace copy with filled ghosts.
int load_input(ns3d_ctx *c, const double *const *in) {
    int rc;
    if ((rc = copy_fields(c, c->xg, in, 5))) return rc;
    return fill_ghosts(c, c->xg, 5);
}

LinComb lc(ns3d_ctx *c, double *const *out, double alpha, double *const *A, double beta, double *const *B,
           double gamma, double *const *P) {
    LinComb l{};
    for (int n = 0; n < 5; ++n) {
        l.out[n] = out[n];
        l.A[n] = A ? A[n] : nullptr;
        l.B[n] = B ? B[n] : nullptr;
        l.P[n] = P ? P[n] : nullptr;
    }
    l.alpha = alpha;
    l.beta = beta;
    l.gamma = gamma;
    l.b = c->b;
    return l;
}

int lincomb(ns3d_ctx *c, double *const *out, double alpha, double *const *A, double beta, double *const *B,
            double gamma, double *const *P) {
    launch_lincomb(lc(c, out, alpha, A, beta, B, gamma, P), c->stream);
    NS3D_LAUNCHED("lincomb");
    return NS3D_OK;
}

int check_ptrs(const void *const *p, int n) {
    if (!p) return fail(NS3D_E_ARG, "null field array");
    for (int q = 0; q < n; ++q)
        if (!p[q]) return fail(NS3D_E_ARG, "null field pointer %d", q);
    return NS3D_OK;
}

int check_adjoint(ns3d_ctx *c) {
    if (!adjoint_supported(c->P.interpolation, c->P.riemann_solver, c->P.char_recon != 0))
        return fail(NS3D_E_UNSUPPORTED,
                    "the adjoint covers component-wise WCNS5 / WENO5-JS interpolation with any Riemann solver "
                    "(got interpolation %d, Riemann solver %d, char_recon %d)",
                    c->P.interpolation, c->P.riemann_solver, c->P.char_recon);
    return NS3D_OK;
}

// The reverse of one RK3 step: the context holds U^n, untouched; U1, U2:
// U^(1), U^(2) (interior cells are read; every adjoint kernel reads periodic
// images, not ghost layers).  Without Q1 / Q2 the context holds the
// primitives of U^(2) and the reverse forms those of U^(1) and U^n itself;
// with them (the primitives of U^(1) and U^(2) as the forward stages wrote
// them) the context holds the primitives of U^n.  adj: Ub^3 on entry, Ub^n on
// return (ghosts zero); the context's primitives are left those of U^n, so
// the context is at U^n as on entry.
int adjoint_reverse(ns3d_ctx *c, double dt, double *const *adj, const double *const *U1,
                    const double *const *U2, const double *const *Q1 = nullptr,
                    const double *const *Q2 = nullptr) {
    int rc;
    // Each adjoint RHS P hands its value to the stage combinations in its
    // last kernel (the P operand marks where it enters).
    LinTail t{};
    // stage 3 at U^(2): Ub2 = 2/3 (Ub3 + dt P), Ub0 = Ub3 / 3
    t.n = 2;
    t.c[0] = lc(c, c->a1, 2.0 / 3.0, adj, 0.0, nullptr, 2.0 / 3.0 * dt, c->pr);
    t.c[1] = lc(c, c->a0, 1.0 / 3.0, adj, 0.0, nullptr, 0.0, nullptr);
    if ((rc = adjoint_rhs_impl(c, adj, c->pr, U2, &t, Q2))) return rc;
    // stage 2 at U^(1): Ub0 += 3/4 Ub2, Ub1 = 1/4 (Ub2 + dt P)
    t.c[0] = lc(c, c->a0, 1.0, c->a0, 0.75, c->a1, 0.0, nullptr);
    t.c[1] = lc(c, c->a1, 0.25, c->a1, 0.0, nullptr, 0.25 * dt, c->pr);
    if (!Q1 && (rc = prims_from(c, U1))) return rc;
    if ((rc = adjoint_rhs_impl(c, c->a1, c->pr, U1, &t, Q1))) return rc;
    // stage 1 at U^n: Ub^n = Ub0 + Ub1 + dt P (the primitives of U^n over
    // every cell, ghosts included, as the context had them on entry)
    if (!Q1) {
        launch_c2p(conv_args(c), c->stream);
        NS3D_LAUNCHED("c2p");
    }
    t.n = 1;
    t.c[0] = lc(c, adj, 1.0, c->a0, 1.0, c->a1, dt, c->pr);
    if ((rc = adjoint_rhs_impl(c, c->a1, c->pr, nullptr, &t))) return rc;
    launch_zero_ghosts(lc(c, adj, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr), c->stream);
    NS3D_LAUNCHED("zero ghosts");
    return NS3D_OK;
}

}  // namespace

extern "C" {

const char *ns3d_last_error(void) { return g_err.c_str(); }

int ns3d_copy_field_to_device(ns3d_ctx *c, int field, double *dst) {
    NS3D_CTX(c);
    if (field < 0 || field >= NS3D_NUM_CELL_FIELDS || !dst) return fail(NS3D_E_ARG, "bad field or pointer");
    NS3D_HIP(hipMemcpyAsync(dst, c->cell[field], c->ncell * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return NS3D_OK;
}

int ns3d_copy_field_from_device(ns3d_ctx *c, int field, const double *src) {
    NS3D_CTX(c);
    if (field < 0 || field >= NS3D_NUM_CELL_FIELDS || !src) return fail(NS3D_E_ARG, "bad field or pointer");
    NS3D_HIP(hipMemcpyAsync(c->cell[field], src, c->ncell * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return NS3D_OK;
}

int ns3d_set_state(ns3d_ctx *c, const double *const *U) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_ptrs((const void *const *)U, 5))) return rc;
    if ((rc = copy_fields(c, conserved(c), U, 5))) return rc;
    return refresh_state(c);
}

int ns3d_get_state(ns3d_ctx *c, double *const *U) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_ptrs((const void *const *)U, 5))) return rc;
    return copy_fields(c, U, conserved(c), 5);
}

// The derivative and stored-stage entry points run on periodic boxes only.
int check_periodic_box(ns3d_ctx *c, const char *what) {
    if (c->bounded) return fail(NS3D_E_UNSUPPORTED, "%s: periodic boxes only (this box has physical faces)", what);
    return NS3D_OK;
}

int ns3d_tangent_rhs(ns3d_ctx *c, const double *const *dU, double *const *out) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "tangent_rhs"))) return rc;
    if ((rc = check_ptrs((const void *const *)dU, 5)) || (rc = check_ptrs((const void *const *)out, 5))) return rc;
    if ((rc = ensure_adj(c)) || (rc = refresh_state(c)) || (rc = load_input(c, dU))) return rc;
    if ((rc = zero_fields(c, out, 5))) return rc;  // zero ghosts
    return tangent_rhs_impl(c, c->xg, out);
}

int ns3d_adjoint_rhs(ns3d_ctx *c, const double *const *lam, double *const *out) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "adjoint_rhs"))) return rc;
    if ((rc = check_ptrs((const void *const *)lam, 5)) || (rc = check_ptrs((const void *const *)out, 5))) return rc;
    if ((rc = check_adjoint(c)) || (rc = ensure_adj(c)) || (rc = refresh_state(c)) || (rc = load_input(c, lam)))
        return rc;
    if ((rc = zero_fields(c, out, 5))) return rc;  // zero ghosts
    return adjoint_rhs_impl(c, c->xg, out);
}

int ns3d_tangent_step(ns3d_ctx *c, double dt, double *const *dU) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "tangent_step"))) return rc;
    if ((rc = check_ptrs((const void *const *)dU, 5))) return rc;
    if ((rc = ensure_adj(c)) || (rc = refresh_state(c)) || (rc = record0(c))) return rc;
    if ((rc = copy_fields(c, c->a0, dU, 5))) return rc;  // delta U^n
    static const double A[3] = {1.0, 0.75, 1.0 / 3.0}, B[3] = {0.0, 0.25, 2.0 / 3.0};
    for (int s = 0; s < 3; ++s) {
        // the tangent's ghosts are filled in place (interior values unchanged)
        if ((rc = fill_ghosts(c, dU, 5)) || (rc = tangent_rhs_impl(c, dU, c->pr))) return rc;
        // delta U^(s+1) = A delta U^n + B delta U^(s) + B dt P  (stage 1: delta U^n + dt P)
        const double g = s == 0 ? dt : B[s] * dt;
        if ((rc = lincomb(c, dU, A[s], c->a0, B[s], s == 0 ? nullptr : dU, g, c->pr))) return rc;
        if ((rc = fused_stage(c, s + 1, nullptr, dt))) return rc;
    }
    return NS3D_OK;
}

int ns3d_adjoint_step(ns3d_ctx *c, double dt, double *const *adj) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "adjoint_step"))) return rc;
    if ((rc = check_ptrs((const void *const *)adj, 5))) return rc;
    if ((rc = check_adjoint(c)) || (rc = ensure_adj(c))) return rc;
    launch_c2p(conv_args(c), c->stream);  // the primitives of U^n
    NS3D_LAUNCHED("c2p");
    // forward stages 1 and 2 from U^n, which stays in the context with its
    // primitives: U^(1) into s1 with its u, v, w, T into dq (the tangent's,
    // free during an adjoint), U^(2) into xg (free during a step's adjoint)
    // with its u, v, w, T into q2; the reverse takes those primitives as they
    // are.  The ghost layers of the stage states are not formed (nothing
    // reads them).
    double *q1[5] = {c->dq[0], c->dq[1], c->dq[2], nullptr, c->dq[3]};
    double *q2[5] = {c->q2[0], c->q2[1], c->q2[2], nullptr, c->q2[3]};
    const StageIO st1{conserved(c), c->s1, conserved(c), nullptr, q1}, st2{c->s1, c->xg, conserved(c), q1, q2};
    if ((rc = fused_stage(c, 1, nullptr, dt, U0_INPUT, false, &st1)) ||
        (rc = fused_stage(c, 2, nullptr, dt, U0_READ, false, &st2)))
        return rc;
    return adjoint_reverse(c, dt, adj, c->s1, c->xg, q1, q2);
}

int ns3d_step_stages(ns3d_ctx *c, double dt, double *const *U1, double *const *U2) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "step_stages"))) return rc;
    if ((rc = check_ptrs((const void *const *)U1, 5)) || (rc = check_ptrs((const void *const *)U2, 5))) return rc;
    // the stages written straight into the caller's arrays (interior cells),
    // U0 recorded by the stage-1 x pencil
    const StageIO st1{conserved(c), U1, recorded(c)}, st2{U1, U2, recorded(c)}, st3{U2, conserved(c), recorded(c)};
    if ((rc = fused_stage(c, 1, nullptr, dt, U0_RECORD, false, &st1, true)) ||
        (rc = fused_stage(c, 2, nullptr, dt, U0_READ, false, &st2, true))) {
        c->prims_fp32 = false;
        return rc;
    }
    return fused_stage(c, 3, nullptr, dt, U0_READ, true, &st3);
}

int ns3d_adjoint_step_stages(ns3d_ctx *c, double dt, double *const *adj, const double *const *U1,
                             const double *const *U2) {
    NS3D_CTX(c);
    int rc;
    if ((rc = check_periodic_box(c, "adjoint_step_stages"))) return rc;
    if ((rc = check_ptrs((const void *const *)adj, 5)) || (rc = check_ptrs((const void *const *)U1, 5)) ||
        (rc = check_ptrs((const void *const *)U2, 5)))
        return rc;
    if ((rc = check_adjoint(c)) || (rc = ensure_adj(c))) return rc;
    // the primitives of U^(2) from the caller's array; U^n stays in the context
    if ((rc = prims_from(c, U2))) return rc;
    return adjoint_reverse(c, dt, adj, U1, U2);
}

int ns3d_create(ns3d_ctx **out, const ns3d_grid *grid, const ns3d_params *params, int precision) {
    if (!out || !grid) return fail(NS3D_E_ARG, "ns3d_create: null argument");
    *out = nullptr;
    int rc = check_params(params);
    if (rc) return rc;
    if ((rc = check_box(grid))) return rc;
    if (precision != NS3D_PRECISION_FP64 && precision != NS3D_PRECISION_MIXED)
        return fail(NS3D_E_ARG, "unknown precision %d", precision);
    if (any_physical(*params) && precision != NS3D_PRECISION_FP64)
        return fail(NS3D_E_UNSUPPORTED, "the mixed method runs on periodic boxes only");
    // set-up of the allocation-free entry point on the current device
    double *scratch_dev = nullptr;
    if ((rc = noalloc_scratch(&scratch_dev))) return rc;
    ns3d_ctx *c = new ns3d_ctx();
    c->grid = *grid;
    c->P = *params;
    c->precision = precision;
    c->stream = nullptr;
    Box &b = c->b;
    b = box_of(*grid);
    c->ncell = (size_t)b.sx * b.sy * b.sz;
    c->nint = (size_t)b.nx * b.ny * b.nz;
    const size_t npart = (size_t)ke_partial_count(b);
    // Field strides: rounded to 512 B and staggered by 2.5 KB, so that arrays
    // whose sizes are powers of two (128 MiB at 256^3) do not all start on
    // the same HBM channel when a kernel touches the same index of many.
    const size_t cstride = field_stride(c->ncell), istride = field_stride(c->nint);
    const size_t total = NS3D_NUM_CELL_FIELDS * cstride + (25 + NVISC) * istride + npart + 16 + CFL_SLOTS;
    if (hipMalloc(&c->slab, total * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return fail(NS3D_E_NOMEM, "out of device memory for %zu doubles", total);
    }
    double *p = c->slab;
    for (int f = 0; f < NS3D_NUM_CELL_FIELDS; ++f, p += cstride) c->cell[f] = p;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n, p += istride) c->face[d][n] = p;
    for (int n = 0; n < NVISC; ++n, p += istride) c->visc[n] = p;
    for (int n = 0; n < 5; ++n, p += istride) c->rz[n] = p;
    for (int n = 0; n < 5; ++n, p += istride) c->ry[n] = p;
    c->partial = p;
    p += npart;
    c->scal = p;
    p += 16;
    c->cfl_slots = reinterpret_cast<unsigned long long *>(p);
    c->grad_slab = nullptr;
    c->adj_slab = nullptr;
    c->obs_work = nullptr;
    c->obs_cap = 0;
    c->fft_plan = nullptr;
    c->timing = false;
    c->visc_fp32 = false;
    c->prims_fp32 = false;
    c->side = nullptr;
    c->ev_fork = c->ev_join = nullptr;
    c->bounded = any_physical(*params);
    for (int a = 0; a < 3; ++a) c->per[a] = params->bc[2 * a] == NS3D_BC_PERIODIC;
    c->fref_slab = nullptr;
    if (c->bounded) {
        // reference-layout face arrays, 5 per direction, each smaller than a cell field
        const size_t n = 15 * cstride;
        if (hipMalloc(&c->fref_slab, n * sizeof(double)) != hipSuccess ||
            hipMemset(c->fref_slab, 0, n * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            if (c->fref_slab) (void)hipFree(c->fref_slab);
            (void)hipFree(c->slab);
            delete c;
            return fail(NS3D_E_NOMEM, "face arrays of the bounded box: out of device memory for %zu doubles", n);
        }
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < 5; ++v) c->fref[d][v] = c->fref_slab + (5 * d + v) * cstride;
    }
    if (hipMemset(c->slab, 0, total * sizeof(double)) != hipSuccess) {
        (void)hipFree(c->slab);
        delete c;
        return fail(NS3D_E_HIP, "hipMemset failed");
    }
    launch_fill(c->cell[NS3D_RHO], (long)c->ncell, 1.0, nullptr);  // Field3D::allocate: rho = 1
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(c->slab);
        delete c;
        return fail(NS3D_E_HIP, "initialisation failed");
    }
    // the fused stage's second stream (default priority), when asked for;
    // without it the kernels run in order on `stream`
    const char *concurrent = std::getenv("NS3D_CONCURRENT_PENCILS");
    if (concurrent && concurrent[0] == '1') {
        const bool ok = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) == hipSuccess &&
                        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess &&
                        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            destroy_side(c);
        }
    }
    *out = c;
    return NS3D_OK;
}

int ns3d_set_kernel_timing(ns3d_ctx *c, int enable) {
    NS3D_CTX(c);
    c->timing = enable != 0;
    return NS3D_OK;
}

int ns3d_kernel_times(ns3d_ctx *c, double *ms, int *launches) {
    NS3D_CTX(c);
    if (!ms || !launches) return fail(NS3D_E_ARG, "null output");
    for (int k = 0; k < NS3D_NUM_KERNEL_CLASSES; ++k) {
        ms[k] = 0.0;
        launches[k] = 0;
    }
    NS3D_HIP(hipStreamSynchronize(c->stream));
    for (auto &e : c->events) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e.second.first, e.second.second) == hipSuccess) {
            ms[e.first] += t;
            launches[e.first] += 1;
        }
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    c->events.clear();
    return NS3D_OK;
}

int ns3d_destroy(ns3d_ctx *c) {
    if (!c) return NS3D_OK;
    (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->events) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    if (c->grad_slab) (void)hipFree(c->grad_slab);
    if (c->adj_slab) (void)hipFree(c->adj_slab);
    if (c->obs_work) (void)hipFree(c->obs_work);
    destroy_fft_plan(c->fft_plan);
    if (c->fref_slab) (void)hipFree(c->fref_slab);
    destroy_side(c);
    (void)hipFree(c->slab);
    delete c;
    return NS3D_OK;
}

int ns3d_set_stream(ns3d_ctx *c, void *stream) {
    NS3D_CTX(c);
    c->stream = (hipStream_t)stream;
    return NS3D_OK;
}

int ns3d_set_params(ns3d_ctx *c, const ns3d_params *params) {
    NS3D_CTX(c);
    int rc = check_params(params);
    if (rc) return rc;
    for (int f = 0; f < 6; ++f)
        if (params->bc[f] != c->P.bc[f])
            return fail(NS3D_E_ARG, "boundary types are fixed when the context is created");
    c->P = *params;
    return NS3D_OK;
}

size_t ns3d_cell_count(const ns3d_ctx *c) { return c ? c->ncell : 0; }

int ns3d_synchronize(ns3d_ctx *c) {
    NS3D_CTX(c);
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_upload(ns3d_ctx *c, int field, const double *host) {
    NS3D_CTX(c);
    if (field < 0 || field >= NS3D_NUM_CELL_FIELDS || !host) return fail(NS3D_E_ARG, "ns3d_upload: bad field or pointer");
    NS3D_HIP(hipMemcpyAsync(c->cell[field], host, c->ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_download(ns3d_ctx *c, int field, double *host) {
    NS3D_CTX(c);
    if (field < 0 || field >= NS3D_NUM_CELL_FIELDS || !host)
        return fail(NS3D_E_ARG, "ns3d_download: bad field or pointer");
    NS3D_HIP(hipMemcpyAsync(host, c->cell[field], c->ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_field_device_ptr(ns3d_ctx *c, int field, double **dptr) {
    NS3D_CTX(c);
    if (field < 0 || field >= NS3D_NUM_CELL_FIELDS || !dptr) return fail(NS3D_E_ARG, "bad field or pointer");
    *dptr = c->cell[field];
    return NS3D_OK;
}

int ns3d_conserved_to_primitive(ns3d_ctx *c) {
    NS3D_CTX(c);
    launch_c2p(conv_args(c), c->stream);
    NS3D_LAUNCHED("conservedToPrimitive");
    return NS3D_OK;
}

int ns3d_primitive_to_conserved(ns3d_ctx *c) {
    NS3D_CTX(c);
    launch_p2c(conv_args(c), c->stream);
    NS3D_LAUNCHED("primitiveToConserved");
    return NS3D_OK;
}

int ns3d_record_conserved_to0(ns3d_ctx *c) {
    NS3D_CTX(c);
    return record0(c);
}

int ns3d_apply_boundary(ns3d_ctx *c) {
    NS3D_CTX(c);
    return halo(c, false);
}

int ns3d_apply_boundary_halfnode_flux(ns3d_ctx *c) {
    NS3D_CTX(c);
    return c->bounded ? bnd_ghost_faces(c) : NS3D_OK;
}

int ns3d_exchange_halos_viscous_flux(ns3d_ctx *c) {
    NS3D_CTX(c);
    return NS3D_OK;
}

int ns3d_compute_invis_flux(ns3d_ctx *c) {
    NS3D_CTX(c);
    return c->bounded ? bnd_faces(c) : run_faces(c);
}

int ns3d_compute_invis_dflux(ns3d_ctx *c) {
    NS3D_CTX(c);
    if (c->bounded) return bnd_assemble(c, ASM_INVISCID | ASM_ACCUMULATE);
    launch_assemble(asm_args(c, 0, ASM_INVISCID | ASM_ACCUMULATE, nullptr, 0.0), c->stream);
    NS3D_LAUNCHED("invis_dflux");
    return NS3D_OK;
}

size_t ns3d_face_count(const ns3d_ctx *c, int dir) {
    if (!c || dir < 0 || dir > 2) return 0;
    const Box &b = c->b;
    return dir == 0 ? (size_t)(b.sx - 1) * b.sy * b.sz
                    : (dir == 1 ? (size_t)b.sx * (b.sy - 1) * b.sz : (size_t)b.sx * b.sy * (b.sz - 1));
}

int ns3d_download_face_flux(ns3d_ctx *c, int dir, int comp, double *host) {
    NS3D_CTX(c);
    if (dir < 0 || dir > 2 || comp < 0 || comp > 4 || !host) return fail(NS3D_E_ARG, "bad face/component");
    if (c->bounded) {  // already in the reference's face layout
        NS3D_HIP(hipMemcpyAsync(host, c->fref[dir][comp], ns3d_face_count(c, dir) * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
        NS3D_HIP(hipStreamSynchronize(c->stream));
        return NS3D_OK;
    }
    std::vector<double> f;
    int rc = fetch_compact(c, c->face[dir][comp], f);
    if (rc) return rc;
    const Box &b = c->b;
    std::memset(host, 0, ns3d_face_count(c, dir) * sizeof(double));
    const int n = dir == 0 ? b.nx : (dir == 1 ? b.ny : b.nz);
    auto w = [](int i, int m) { return ((i % m) + m) % m; };
    // reference face index q (between ghosted cells q and q+1) = compact face q - NG
    for (int k = 0; k < b.nz; ++k)
        for (int j = 0; j < b.ny; ++j)
            for (int i = 0; i < b.nx; ++i) {
                if (dir == 0 && i == 0) {
                    for (int q = 0; q < b.sx - 1; ++q)
                        host[((long)(k + NG) * b.sy + (j + NG)) * (b.sx - 1) + q] = f[b.oidx(w(q - NG, n), j, k)];
                } else if (dir == 1 && j == 0) {
                    for (int q = 0; q < b.sy - 1; ++q)
                        host[((long)(k + NG) * (b.sy - 1) + q) * b.sx + (i + NG)] = f[b.oidx(i, w(q - NG, n), k)];
                } else if (dir == 2 && k == 0) {
                    for (int q = 0; q < b.sz - 1; ++q)
                        host[((long)q * b.sy + (j + NG)) * b.sx + (i + NG)] = f[b.oidx(i, j, w(q - NG, n))];
                }
            }
    return NS3D_OK;
}

int ns3d_compute_gradients(ns3d_ctx *c) {
    NS3D_CTX(c);
    int rc = ensure_grad(c);
    if (rc) return rc;
    if (c->bounded) return bnd_visc(c, true, false);
    launch_visc(visc_args(c, true, false), c->stream);
    NS3D_LAUNCHED("gradients");
    return NS3D_OK;
}

int ns3d_compute_viscous_flux(ns3d_ctx *c) {
    NS3D_CTX(c);
    int rc = ensure_grad(c);
    if (rc) return rc;
    if (c->bounded) return bnd_visc(c, true, true);
    launch_visc_from_grad(visc_args(c, true, true), c->stream);
    NS3D_LAUNCHED("viscous flux");
    c->visc_fp32 = false;
    return NS3D_OK;
}

int ns3d_compute_vis_flux(ns3d_ctx *c) {
    NS3D_CTX(c);
    if (c->visc_fp32)
        return fail(NS3D_E_ARG, "the viscous fluxes were last written by a mixed-precision stage: call "
                                "compute_viscous_flux first");
    if (c->bounded) return bnd_assemble(c, ASM_VISCOUS | ASM_ACCUMULATE);
    launch_assemble(asm_args(c, 0, ASM_VISCOUS | ASM_ACCUMULATE, nullptr, 0.0), c->stream);
    NS3D_LAUNCHED("vis_flux");
    return NS3D_OK;
}

int ns3d_download_gradient(ns3d_ctx *c, int q, double *host) {
    NS3D_CTX(c);
    if (q < 0 || q >= 12 || !host) return fail(NS3D_E_ARG, "bad gradient index");
    int rc = ensure_grad(c);
    if (rc) return rc;
    std::vector<double> f;
    if ((rc = fetch_compact(c, c->grad[q], f))) return rc;
    scatter_cells(c, f, host, false);
    return NS3D_OK;
}

int ns3d_download_viscous_flux(ns3d_ctx *c, int dir, int comp, double *host) {
    NS3D_CTX(c);
    if (dir < 0 || dir > 2 || comp < 0 || comp > 4 || !host) return fail(NS3D_E_ARG, "bad viscous flux index");
    if (comp == 0) {
        std::memset(host, 0, c->ncell * sizeof(double));
        return NS3D_OK;
    }
    if (c->visc_fp32)
        return fail(NS3D_E_ARG, "the viscous fluxes were last written by a mixed-precision stage: call "
                                "compute_viscous_flux first");
    static const int slot[3][4] = {{VXX, VXY, VXZ, VEX}, {VXY, VYY, VYZ, VEY}, {VXZ, VYZ, VZZ, VEZ}};
    std::vector<double> f;
    int rc = fetch_compact(c, c->visc[slot[dir][comp - 1]], f);
    if (rc) return rc;
    scatter_cells(c, f, host, true);
    return NS3D_OK;
}

int ns3d_compute_rhs(ns3d_ctx *c) {
    NS3D_CTX(c);
    // the RHS of ghost cells is zero, as after the reference's std::fill
    for (int n = 0; n < 5; ++n) NS3D_HIP(hipMemsetAsync(rhs(c)[n], 0, c->ncell * sizeof(double), c->stream));
    return c->bounded ? bnd_rhs(c) : fused_stage(c, 0, nullptr, 0.0);
}

int ns3d_runge_kutta_3(ns3d_ctx *c, double dt) {
    NS3D_CTX(c);
    return rk3(c, nullptr, dt);
}

int ns3d_runge_kutta_3_dev(ns3d_ctx *c, const double *d_dt) {
    NS3D_CTX(c);
    if (!d_dt) return fail(NS3D_E_ARG, "null device dt");
    return rk3(c, d_dt, 0.0);
}

int ns3d_compute_timestep(ns3d_ctx *c, double *dt_out) {
    NS3D_CTX(c);
    if (!dt_out) return fail(NS3D_E_ARG, "null dt_out");
    int rc = cfl(c, c->scal);
    if (rc) return rc;
    NS3D_HIP(hipMemcpyAsync(dt_out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_compute_timestep_dev(ns3d_ctx *c, double *d_dt) {
    NS3D_CTX(c);
    if (!d_dt) return fail(NS3D_E_ARG, "null device dt");
    return cfl(c, d_dt);
}

int ns3d_compute_timestep_noalloc(const ns3d_grid *grid, const ns3d_params *params, const double *d_rho,
                                  const double *d_rhou, const double *d_E, double *d_dt, void *stream) {
    if (!grid || !params) return fail(NS3D_E_ARG, "ns3d_compute_timestep_noalloc: null grid or params");
    if (!d_rho || !d_rhou || !d_E || !d_dt)
        return fail(NS3D_E_ARG, "ns3d_compute_timestep_noalloc: null device pointer");
    int rc = check_box(grid);
    if (rc) return rc;
    if (!(params->gamma > 1.0)) return fail(NS3D_E_ARG, "gamma must exceed 1 (got %g)", params->gamma);
    CflArgs a{};
    a.U[0] = d_rho;
    a.U[1] = d_rhou;
    a.U[4] = d_E;
    a.dt_out = d_dt;
    a.b = box_of(*grid);
    a.gamma = params->gamma;
    a.dx = grid->dx;
    a.cfl = params->cfl;
    const long blocks = cfl_block_count(a.b);
    if (blocks > NS3D_NOALLOC_SCRATCH_DOUBLES)
        return fail(NS3D_E_ARG, "ns3d_compute_timestep_noalloc: %ld block maxima exceed the %d-slot scratch", blocks,
                    (int)NS3D_NOALLOC_SCRATCH_DOUBLES);
    if ((rc = noalloc_scratch(&a.partial))) return rc;
    launch_cfl(a, (hipStream_t)stream);
    NS3D_LAUNCHED("cfl (no allocation)");
    return NS3D_OK;
}

int ns3d_time_advance(ns3d_ctx *c, int max_steps, double dt_fixed, double total_time, double output_timefreq,
                      double *dt_out, int *steps_taken) {
    NS3D_CTX(c);
    double t = 0.0, out_t = output_timefreq;
    int taken = 0, rc;
    for (int step = 1; step <= max_steps; ++step) {
        double dt = dt_fixed;
        if ((rc = record0(c))) return rc;
        const double eps = 1e-12;
        if (std::abs(total_time - t) <= eps) break;
        if (dt_fixed < 0.0 && (rc = ns3d_compute_timestep(c, &dt))) return rc;
        if (t >= out_t - eps) out_t += output_timefreq;
        if (t + dt > total_time) dt = total_time - t;
        if (t + dt > out_t) {
            dt = out_t - t;
            out_t += output_timefreq;
        }
        if (dt <= eps) {
            fprintf(stderr, "dt too small or zero; aborting to avoid stall at Time=%g\n", t);
            break;
        }
        t += dt;
        if ((rc = rk3(c, nullptr, dt, false, false))) return rc;
        if (dt_out) dt_out[taken] = dt;
        ++taken;
        if (t >= total_time) break;
    }
    if (taken > 0 && (rc = halo(c, true))) return rc;
    NS3D_HIP(hipStreamSynchronize(c->stream));
    if (steps_taken) *steps_taken = taken;
    return NS3D_OK;
}

int ns3d_advance_steps(ns3d_ctx *c, int nsteps, double *d_dt_hist) {
    NS3D_CTX(c);
    int rc;
    bool folded = false;
    for (int s = 0; s < nsteps; ++s) {
        double *d_dt = d_dt_hist ? d_dt_hist + s : c->scal;
        // recordConservedTo0 fused into the first stage of the last step;
        // the steps before it keep U^n in place as their U0 and store no
        // copy.  Ghost layers refilled after the last step only, which alone
        // hands out FP64 primitives (the mixed method's steps before it pass
        // theirs on in FP32; the CFL reduction reads the conserved state)
        const bool last = s == nsteps - 1;
        // steps after the first take dt from the CFL maxima that the previous
        // step's stage-3 x pencils folded in (periodic boxes), instead of a
        // separate pass over the state
        if ((rc = folded ? cfl_folded(c, d_dt) : cfl(c, d_dt))) return rc;
        const bool fold = !last && !c->bounded;
        if ((rc = rk3(c, d_dt, 0.0, last, last, !last, !last, fold))) return rc;
        folded = fold;
    }
    return NS3D_OK;
}

int ns3d_total_kinetic_energy(ns3d_ctx *c, double *out) {
    NS3D_CTX(c);
    if (!out) return fail(NS3D_E_ARG, "null output");
    KeArgs a{};
    a.rho = c->cell[NS3D_RHO];
    a.u = c->cell[NS3D_U];
    a.v = c->cell[NS3D_V];
    a.w = c->cell[NS3D_W];
    a.partial = c->partial;
    a.b = c->b;
    a.vol = c->grid.dx * c->grid.dy * c->grid.dz;
    launch_ke(a, c->scal + 1, c->stream);
    NS3D_LAUNCHED("kinetic energy");
    NS3D_HIP(hipMemcpyAsync(out, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

}  // extern "C"

namespace {

int ensure_obs(ns3d_ctx *c, long doubles) {
    if (c->obs_cap >= doubles) return NS3D_OK;
    if (c->obs_work) (void)hipFree(c->obs_work);
    c->obs_work = nullptr;
    c->obs_cap = 0;
    if (hipMalloc(&c->obs_work, (size_t)doubles * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NS3D_E_NOMEM, "observables: out of device memory for %ld doubles", doubles);
    }
    c->obs_cap = doubles;
    return NS3D_OK;
}

int spectrum(ns3d_ctx *c, int nshell, double *host) {
    int rc;
    if ((rc = ensure_obs(c, spectrum_work_doubles(c->b, nshell)))) return rc;
    const double *vel[3] = {c->cell[NS3D_U], c->cell[NS3D_V], c->cell[NS3D_W]};
    if (launch_spectrum(c->b, vel, nshell, c->obs_work, c->fft_plan, c->stream))
        return fail(NS3D_E_HIP, "hipFFT failed for the %d x %d x %d transform", c->b.nz, c->b.ny, c->b.nx);
    NS3D_LAUNCHED("energy spectrum");
    NS3D_HIP(hipMemcpyAsync(host, spectrum_result(c->b, nshell, c->obs_work), (size_t)nshell * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

}  // namespace

int ns3d::set_error(int code, const char *msg) { return fail(code, "%s", msg); }
SchemeCoef<double> ns3d::scheme_coef64(const ns3d_params &P) { return coef64(P); }

extern "C" {

int ns3d_energy_spectrum(ns3d_ctx *c, double *Ek, int nshell) {
    NS3D_CTX(c);
    if (!Ek || nshell < 1) return fail(NS3D_E_ARG, "energy spectrum: bad output");
    return spectrum(c, nshell, Ek);
}

int ns3d_turbulence_statistics(ns3d_ctx *c, ns3d_turbulence *out) {
    NS3D_CTX(c);
    if (!out) return fail(NS3D_E_ARG, "null output");
    const int nshell = 2 * c->b.nx + 1;
    std::vector<double> ek(nshell);
    int rc;
    if ((rc = spectrum(c, nshell, ek.data()))) return rc;
    if ((rc = ensure_obs(c, std::max(c->obs_cap, moment_work_doubles())))) return rc;
    MomentArgs m{};
    m.rho = c->cell[NS3D_RHO];
    m.u = c->cell[NS3D_U];
    m.v = c->cell[NS3D_V];
    m.w = c->cell[NS3D_W];
    m.T = c->cell[NS3D_T];
    m.b = c->b;
    m.mu = c->P.mu;
    m.S_ref = c->P.S_ref;
    m.partial = c->obs_work;
    m.out = c->obs_work + moment_work_doubles() - 8;
    launch_moments(m, c->stream);
    NS3D_LAUNCHED("moments");
    double s[8];
    NS3D_HIP(hipMemcpyAsync(s, m.out, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    NS3D_HIP(hipStreamSynchronize(c->stream));
    const double n = (double)c->nint;
    const double nu = (s[1] / n) / (s[0] / n);
    out->kinetic_energy = 0.5 * (s[5] + s[6] + s[7]) / n;
    out->u_rms = std::sqrt((s[5] + s[6] + s[7]) / (3.0 * n));
    double tke = 0.0;
    for (int q = 0; q < 3; ++q) tke += s[5 + q] / n - (s[2 + q] / n) * (s[2 + q] / n);
    out->tke = 0.5 * tke;
    out->nu = nu;
    double eps = 0.0;
    for (int k = 0; k < nshell; ++k) eps += 2.0 * nu * (double)k * (double)k * ek[k];
    out->dissipation = eps;
    return NS3D_OK;
}

int ns3d_interpolate_batch(const ns3d_params *params, const double *d_st, int n, double flag, double *d_out,
                           void *stream) {
    int rc = check_params(params);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_st || !d_out))) return fail(NS3D_E_ARG, "bad batch");
    if (n == 0) return NS3D_OK;
    launch_interp_batch(params->interpolation, coef64(*params), d_st, n, flag, d_out, (hipStream_t)stream);
    NS3D_LAUNCHED("interpolate batch");
    return NS3D_OK;
}

int ns3d_riemann_batch(const ns3d_params *params, const double *d_UL, const double *d_UR, int n, int dim,
                       double *d_F, void *stream) {
    int rc = check_params(params);
    if (rc) return rc;
    if (n < 0 || dim < 0 || dim > 2 || (n > 0 && (!d_UL || !d_UR || !d_F))) return fail(NS3D_E_ARG, "bad batch");
    if (n == 0) return NS3D_OK;
    launch_riemann_batch(params->riemann_solver, dim, params->gamma, d_UL, d_UR, n, d_F, (hipStream_t)stream);
    NS3D_LAUNCHED("riemann batch");
    return NS3D_OK;
}

}  // extern "C"
