/*
 * ns3d_hip.h -- C ABI of the MI355X (gfx950) RHS + SSP-RK3 library for one
 * periodic box of the WCNS compressible Navier-Stokes solver.
 *
 * Every entry point below replaces one function of the reference's call
 * surface (ns3d_func.h / field_structures.h of wylx-2/ns3d_wcns); the
 * reference function is cited next to each declaration.  Signatures use plain
 * pointers, sizes and POD structs only.
 *
 * Model
 *   - A context owns device-resident SoA fields in the reference's ghosted
 *     layout: sx = nx + 2*ng cells per x line (ng = 3), x fastest,
 *     idx3 = (k*sy + j)*sx + i (field_structures.h:64-67).
 *   - Host arrays handed to upload/download use that same layout, so a
 *     Field3D's std::vector<double>::data() can be passed directly.
 *   - All work is enqueued on the context's HIP stream (ns3d_set_stream);
 *     functions that return values to the host synchronise that stream.
 *   - Status: 0 on success, a negative NS3D_E_* code on failure; the message
 *     is available from ns3d_last_error() (per thread).  The reference
 *     reports errors on std::cerr and returns early (InviscousFlux.cpp:19-31);
 *     here the caller gets the code and the message instead.
 *   - Boundary conditions (ns3d_params.bc): periodic on all six sides runs
 *     the fused pencil path; a box with physical faces (Wall, Symmetry,
 *     Outflow, Inflow; boundary.cpp:58-300, one-sided viscous orders of
 *     ViscousFlux.cpp:15-50) runs the reference's pass structure in FP64
 *     (stepwise kernels with reference-layout face arrays).  The mixed
 *     method and the derivative entry points need a periodic box.
 */
#ifndef NS3D_HIP_H
#define NS3D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define NS3D_OK 0
#define NS3D_E_ARG (-1)         /* bad argument (null pointer, size) */
#define NS3D_E_UNSUPPORTED (-2) /* configuration outside this library */
#define NS3D_E_HIP (-3)         /* HIP runtime error */
#define NS3D_E_NOMEM (-4)       /* device allocation failed */

/* ---- enums: same values as SolverParams' enums (field_structures.h:120-146) ---- */
#define NS3D_INTERP_WENO5 0       /* interpolation = WENO5: WCNS5 weights 1/16,5/8,5/16 (interpolation.cpp:86-115) */
#define NS3D_INTERP_ZERO 1        /* rejected: reads past the ghost layers in the reference */
#define NS3D_INTERP_MDCD_LINEAR 2 /* interpolation.cpp:118-133 */
#define NS3D_INTERP_MDCD_HYBRID 3 /* interpolation.cpp:135-209 */
#define NS3D_INTERP_WENO5_JS 4    /* weno5_reconstruction, Jiang-Shu weights (reconstruction.cpp:69-100) */

#define NS3D_RIEMANN_ROE 0
#define NS3D_RIEMANN_RUSANOV 1
#define NS3D_RIEMANN_HLLC 2
#define NS3D_RIEMANN_HLL 3
#define NS3D_RIEMANN_HLLC_P 4
#define NS3D_RIEMANN_AUSM 5

/* Boundary types, same values as SolverParams::BCType (field_structures.h:159).
 * An axis is periodic when both of its faces are Periodic (initialization.cpp:
 * 160-168); a Periodic face opposite a non-periodic one is rejected. */
#define NS3D_BC_PERIODIC 0
#define NS3D_BC_WALL 1     /* mirror, normal velocity negated (boundary.cpp:58-134) */
#define NS3D_BC_SYMMETRY 2 /* same as Wall (boundary.cpp:137-141) */
#define NS3D_BC_OUTFLOW 3  /* copy of the nearest interior cell (boundary.cpp:144-176) */
#define NS3D_BC_INFLOW 4   /* rho = 1, u = 1, v = w = 0, p = 1 (boundary.cpp:179-198) */

#define NS3D_PRECISION_FP64 0  /* plain FP64 method */
#define NS3D_PRECISION_MIXED 1 /* FP32 interpolation/flux/viscous stencils, FP64 state and accumulation */

/* Numerical fields of SolverParams (field_structures.h:86-186).  Derived
 * values (mu, Cv, Cp, Rgas) must already be set, as
 * read_solver_params_from_file does (initialization.cpp:194-197). */
typedef struct ns3d_params {
    double gamma, Pr, Ma, Re, cfl;
    double mu, S_ref, Cv, Cp, Rgas;
    double mdcd_diss, mdcd_disp;
    int interpolation;  /* NS3D_INTERP_* */
    int riemann_solver; /* NS3D_RIEMANN_* */
    int char_recon;     /* 0 component-wise, 1 characteristic-wise */
    int bc[6];          /* NS3D_BC_* of the xmin, xmax, ymin, ymax, zmin, zmax faces */
    int pad_;
} ns3d_params;

/* GridDesc + LocalDesc of one periodic rank (field_structures.h:33-60). */
typedef struct ns3d_grid {
    int nx, ny, nz; /* interior cells; each >= 6 */
    int ng;         /* ghost layers; must be 3 (initialization.cpp:175-191) */
    double dx, dy, dz;
} ns3d_grid;

typedef struct ns3d_ctx ns3d_ctx;

/* Field ids for upload/download and device-pointer access. */
enum ns3d_field {
    NS3D_RHO = 0, NS3D_RHOU, NS3D_RHOV, NS3D_RHOW, NS3D_E,         /* conserved */
    NS3D_RHO0, NS3D_RHOU0, NS3D_RHOV0, NS3D_RHOW0, NS3D_E0,        /* recordConservedTo0 copy */
    NS3D_U, NS3D_V, NS3D_W, NS3D_P, NS3D_T,                        /* primitives */
    NS3D_RHS_RHO, NS3D_RHS_RHOU, NS3D_RHS_RHOV, NS3D_RHS_RHOW, NS3D_RHS_E,
    NS3D_NUM_CELL_FIELDS
};

/* ---- lifetime ---- */
/* Field3D::allocate (field_structures.h:250-373): rho = 1, all else 0.
 * precision NS3D_PRECISION_MIXED selects the mixed method for the fused RHS
 * and RK3 path (compute_rhs, runge_kutta_3, time_advance, advance_steps):
 * face interpolation, characteristic projection, Riemann flux and viscous
 * stencils in FP32 on E - E_ref and p - p_ref (E_ref = Rgas/(gamma-1),
 * p_ref = Rgas: the nondimensional reference state rho = T = 1), with the
 * state, the flux-difference sums, the RK3 accumulation, the pressure
 * recovery and the CFL reduction in FP64.  The stepwise entry points
 * (compute_invis_flux, compute_gradients, ...) always run FP64. */
int ns3d_create(ns3d_ctx **out, const ns3d_grid *grid, const ns3d_params *params, int precision);
int ns3d_destroy(ns3d_ctx *ctx);
/* stream: a hipStream_t passed as void* (NULL = the null stream). */
int ns3d_set_stream(ns3d_ctx *ctx, void *stream);
int ns3d_set_params(ns3d_ctx *ctx, const ns3d_params *params);
const char *ns3d_last_error(void);
/* Cells per field including ghosts (sx*sy*sz). */
size_t ns3d_cell_count(const ns3d_ctx *ctx);
int ns3d_synchronize(ns3d_ctx *ctx);

/* ---- data movement (ghosted host arrays of ns3d_cell_count doubles) ---- */
int ns3d_upload(ns3d_ctx *ctx, int field, const double *host);
int ns3d_download(ns3d_ctx *ctx, int field, double *host);
/* Device address of a field (ghosted layout), for zero-copy use. */
int ns3d_field_device_ptr(ns3d_ctx *ctx, int field, double **dptr);

/* ---- Field3D members (field_structures.h:444-497) ---- */
int ns3d_conserved_to_primitive(ns3d_ctx *ctx); /* Field3D::conservedToPrimitive :444-465 */
int ns3d_primitive_to_conserved(ns3d_ctx *ctx); /* Field3D::primitiveToConserved :468-488 */
int ns3d_record_conserved_to0(ns3d_ctx *ctx);   /* Field3D::recordConservedTo0 :491-497 */

/* ---- boundary (boundary.cpp, field_structures.h:511-1056) ---- */
/* apply_boundary (boundary.cpp:14-54 -> exchange_halos_physical,
 * field_structures.h:936-994): periodic copy of rho,u,v,w,p into the three
 * ghost layers of each face (tangential range interior only); on a box with
 * physical faces the exchange zeroes the face ghosts of a non-periodic axis
 * (its MPI_PROC_NULL receive) and each physical face, xmin to zmax, is then
 * filled over its full planes (apply_wall_bc / apply_symmetry_bc /
 * apply_outflow_bc / apply_inflow_bc, boundary.cpp:58-205). */
int ns3d_apply_boundary(ns3d_ctx *ctx);
/* apply_boundary_halfnode_flux (boundary.cpp:207-300) and
 * exchange_halos_viscous_flux (field_structures.h:877-932): on a periodic
 * axis the copied values are the periodic images the consuming kernels read
 * by wrapped indexing, so the periodic box moves no data.  On a box with
 * physical faces the ghost faces are set: periodic images, zero on Wall /
 * Symmetry / Inflow faces, the boundary face on Outflow faces. */
int ns3d_apply_boundary_halfnode_flux(ns3d_ctx *ctx);
int ns3d_exchange_halos_viscous_flux(ns3d_ctx *ctx);

/* ---- inviscid flux (InviscousFlux.cpp) ---- */
/* compute_invis_flux (InviscousFlux.cpp:10-144) with
 * WCNS_Riemann_InviscidFlux (:146-268): the numerical flux at every face of
 * every interior line, into the library's face arrays. */
int ns3d_compute_invis_flux(ns3d_ctx *ctx);
/* compute_invis_dflux (InviscousFlux.cpp:1027-1056): rhs -= 6th-order
 * half-node difference of the face fluxes, x then y then z. */
int ns3d_compute_invis_dflux(ns3d_ctx *ctx);
/* Face fluxes in the reference's face layout (idx_fx/idx_fy/idx_fz,
 * field_structures.h:73-82) for dir 0/1/2 and component 0..4 (mass, momx,
 * momy, momz, E); host array of the face-array size, interior lines filled,
 * including the ng-1 periodic ghost faces per side. */
int ns3d_download_face_flux(ns3d_ctx *ctx, int dir, int comp, double *host);
size_t ns3d_face_count(const ns3d_ctx *ctx, int dir);

/* ---- viscous flux (ViscousFlux.cpp) ---- */
int ns3d_compute_gradients(ns3d_ctx *ctx);      /* compute_gradients :109-148 */
int ns3d_compute_viscous_flux(ns3d_ctx *ctx);   /* compute_viscous_flux :151-208 */
int ns3d_compute_vis_flux(ns3d_ctx *ctx);       /* compute_vis_flux :210-255 */
/* gradient q = 0..11 (du_dx du_dy du_dz dv_dx ... dT_dz); viscous flux
 * dir 0/1/2 (Fv = x, Hv = y, Gv = z) and comp 1..4.  Ghosted cell layout,
 * interior filled (viscous fluxes also in the face ghost layers). */
int ns3d_download_gradient(ns3d_ctx *ctx, int q, double *host);
int ns3d_download_viscous_flux(ns3d_ctx *ctx, int dir, int comp, double *host);

/* ---- RHS, time step, RK3 (range_kutta.cpp, timestep.cpp) ---- */
/* compute_rhs (range_kutta.cpp:11-43): RHS fields = inviscid + viscous. */
int ns3d_compute_rhs(ns3d_ctx *ctx);
/* runge_kutta_3 (range_kutta.cpp:46-92): three Shu-Osher stages from the
 * recorded U0, each followed by C2P -> apply_boundary -> P2C. */
int ns3d_runge_kutta_3(ns3d_ctx *ctx, double dt);
/* Same, reading dt from device memory (no host synchronisation). */
int ns3d_runge_kutta_3_dev(ns3d_ctx *ctx, const double *d_dt);
/* compute_timestep (timestep.cpp:5-23): cfl * min over the interior of
 * dx / (|u| + a), x-velocity and dx only.  Synchronises. */
int ns3d_compute_timestep(ns3d_ctx *ctx, double *dt_out);
/* Same, result left in device memory at d_dt. */
int ns3d_compute_timestep_dev(ns3d_ctx *ctx, double *d_dt);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  ns3d_compute_timestep_dev without a context: the same kernels
 * and the same dt, from the caller's ghosted device fields rho, rho u and E
 * of the box `grid` (interior read), into the device double d_dt, enqueued on
 * `stream` (a hipStream_t as void*; NULL = the null stream).  It allocates
 * nothing: the reduction's block maxima live in a fixed device array of
 * NS3D_NOALLOC_SCRATCH_DOUBLES doubles (64 KiB) per device, held until the
 * process ends.  Any box ns3d_create accepts fits it (the first pass uses at
 * most 8192 blocks); other boxes are rejected with NS3D_E_ARG /
 * NS3D_E_UNSUPPORTED.  Every call on one device shares that array: calls
 * that overlap on one device (concurrent streams or host threads) race. */
#define NS3D_NOALLOC_SCRATCH_DOUBLES 8192
int ns3d_compute_timestep_noalloc(const ns3d_grid *grid, const ns3d_params *params, const double *d_rho,
                                  const double *d_rhou, const double *d_E, double *d_dt, void *stream);

/* time_advance (maincycle.cpp:7-113) without monitors and output: per step
 * record U0, dt from the CFL (or dt_fixed when > 0), clip dt to
 * output_timefreq / total_time, stop when dt <= 1e-12, runge_kutta_3.
 * dt_out (may be NULL) receives each step's dt; *steps_taken the count.
 * Synchronises once per step to run the reference's host-side dt logic.
 * The ghost layers of the state and primitives are refilled once, before
 * the call returns (no kernel of a step reads them). */
int ns3d_time_advance(ns3d_ctx *ctx, int max_steps, double dt_fixed, double total_time,
                      double output_timefreq, double *dt_out, int *steps_taken);
/* nsteps of (record U0, CFL dt, RK3) entirely on the device; d_dt_hist (device,
 * nsteps doubles, may be NULL) receives each step's dt.  No clipping: the
 * caller's total time and output interval must lie beyond the run.  Ghost
 * layers as for time_advance.  Afterwards U0 holds the state at the start of
 * the last step, as after time_advance; the steps before it keep that state
 * in place as their U0 and pass their stage states through the RHS arrays,
 * whose contents are unspecified afterwards. */
int ns3d_advance_steps(ns3d_ctx *ctx, int nsteps, double *d_dt_hist);

/* ---- observables (postprocess.cpp:103-128) ---- */
/* compute_total_energy: sum over the interior of 0.5*rho*|u|^2*dx*dy*dz
 * from the primitive arrays.  Synchronises. */
int ns3d_total_kinetic_energy(ns3d_ctx *ctx, double *out);

/* ---- measurement (no counterpart in the reference) ---- */
/* Algorithmic floating-point operations per cell of the configured scheme:
 * one face flux (interpolation + Riemann solver, counted exactly on the device
 * numerics), the half-node differences of three directions, the viscous
 * fluxes (gradients, Sutherland mu, stresses, heat flux) and their
 * divergence, the stage update with the C2P/P2C round trip and the CFL
 * reduction; per_stage = 3 face + diff + visc_flux + visc_div + update and
 * per_step = 3 per_stage + cfl.  +,-,*,/,sqrt,fabs,fmin,fmax count 1 each. */
typedef struct ns3d_flops {
    double face, diff, visc_flux, visc_div, update, cfl, per_stage, per_step;
} ns3d_flops;
int ns3d_algorithmic_flops(const ns3d_params *params, ns3d_flops *out);

/* Kernel classes of one fused stage, for ns3d_kernel_times. */
enum ns3d_kernel_class {
    NS3D_K_VISCOUS = 0, NS3D_K_PENCIL_Z, NS3D_K_PENCIL_Y, NS3D_K_PENCIL_X, NS3D_K_HALO, NS3D_K_CFL,
    NS3D_NUM_KERNEL_CLASSES
};
/* enable != 0: record HIP events on the context's stream around every launch
 * of the classes above until disabled; ns3d_kernel_times then synchronises and
 * returns, per class, the summed event time (ms) and the launch count since
 * the last call, and releases the events. */
int ns3d_set_kernel_timing(ns3d_ctx *ctx, int enable);
int ns3d_kernel_times(ns3d_ctx *ctx, double *ms, int *launches);

/* ---- scalar kernels over batches (device pointers), for known-answer tests ---- */
/* interpolate_select (interpolation.cpp:13-77): stencils n x 6, flag +1 / -1. */
int ns3d_interpolate_batch(const ns3d_params *params, const double *d_stencils, int n, double flag,
                           double *d_out, void *stream);
/* X_Riemann_solver (InviscousFlux.cpp:270-846): UL, UR n x 5 conserved,
 * normal along axis dim; F n x 5. */
int ns3d_riemann_batch(const ns3d_params *params, const double *d_UL, const double *d_UR, int n, int dim,
                       double *d_F, void *stream);

/* ---- single stencils on the host (the scalar functions of ns3d_func.h:73-131) ----
 * The kernels' own numerics (ns3d_device.h) compiled for the host, for
 * callers that evaluate one stencil or one face at a time; no device is
 * used.  Divisions and square roots are the host's IEEE operations, where
 * the kernels refine the hardware estimates: the two agree to a few ulp.
 *   scalar_interpolate: interpolate_select (interpolation.cpp:13-77), the
 *     scheme params->interpolation on stencil[0..n-1]; n must be 6 (2 for
 *     ZERO, zero_interpolate :80-83), NS3D_E_ARG otherwise, with the
 *     reference's message; flag < 0 reverses the stencil.  WENO5 is
 *     weno5_interpolate (:86-115), MDCD_LINEAR / MDCD_HYBRID
 *     mdcd_linear_interpolate / mdcd_hybrid_interpolate (:118-209),
 *     WENO5_JS weno5_reconstruction (reconstruction.cpp:69-100).
 *   scalar_riemann: X_Riemann_solver (InviscousFlux.cpp:270-846) of
 *     params->riemann_solver, conserved UL, UR -> F (5 each), with the unit
 *     face normal along axis dim (0, 1, 2: the normals that
 *     WCNS_Riemann_InviscidFlux passes, :156-160).
 *   scalar_face_flux: WCNS_Riemann_InviscidFlux (InviscousFlux.cpp:146-268):
 *     Ut (conserved rho, rho u, rho v, rho w, E) and ut (primitive rho, u,
 *     v, w, p) of cells i-2..i+3, 5 x 6 row-major; ut is read only for
 *     characteristic-wise interpolation (cells i and i+1). */
int ns3d_scalar_interpolate(const ns3d_params *params, const double *stencil, int n, double flag, double *out);
int ns3d_scalar_riemann(const ns3d_params *params, const double *UL, const double *UR, int dim, double *F);
int ns3d_scalar_face_flux(const ns3d_params *params, const double *Ut, const double *ut, int dim, double *F);

/* ---- isotropic-turbulence observables on the device ----
 * energy_spectrum: E(k) of the velocity fields (compute_energy_spectrum_rank0,
 * EnergySpectrum.cpp:63-156): forward 3D FFT normalised by 1/N, 0.5 (|U|^2 +
 * |V|^2 + |W|^2) summed over shells lround(|k|), k = 0..nshell-1 (the
 * reference writes nshell = 2 nx + 1).
 * turbulence_statistics (compute_turbulence_statistics,
 * ana_isotropic.cpp:238-305): mean kinetic energy 0.5 <|u|^2>, TKE (half the
 * summed component variances), u_rms, nu = <mu(T)> / <rho> and the spectral
 * dissipation 2 nu sum k^2 E(k). */
typedef struct ns3d_turbulence {
    double kinetic_energy, tke, u_rms, nu, dissipation;
} ns3d_turbulence;
int ns3d_energy_spectrum(ns3d_ctx *ctx, double *Ek, int nshell);
int ns3d_turbulence_statistics(ns3d_ctx *ctx, ns3d_turbulence *out);

/* ---- device-to-device field copies (no counterpart in the reference) ----
 * For harnesses that keep states in their own device buffers (sx*sy*sz
 * doubles, the ghosted layout), e.g. checkpoints of the adjoint. */
int ns3d_copy_field_to_device(ns3d_ctx *ctx, int field, double *d_dst);
int ns3d_copy_field_from_device(ns3d_ctx *ctx, int field, const double *d_src);
/* set_state: the conserved state from five device fields (interior read),
 * then the primitives (conservedToPrimitive) and the periodic ghost layers
 * of both, as before a step; get_state: the conserved state, ghosts included. */
int ns3d_set_state(ns3d_ctx *ctx, const double *const *d_U);
int ns3d_get_state(ns3d_ctx *ctx, double *const *d_U);

/* ---- tangent-linear and adjoint model (no counterpart in the reference) ----
 * Derivatives of compute_rhs (range_kutta.cpp:11-43) and of one
 * runge_kutta_3 step with a fixed dt (range_kutta.cpp:46-92; the C2P/P2C
 * round trip is taken as the identity) at the context's current conserved
 * state.  Arrays are five device pointers (rho, rho u, rho v, rho w, E) in
 * the ghosted layout; only interior cells are read.  The RHS outputs and
 * the returned adjoint have zero ghosts; tangent_step leaves the periodic
 * images in the ghosts of its array.  The non-smooth parts (min/max wave speeds, |.|, branches) use a
 * fixed subgradient, the same in the tangent and the adjoint.
 *   tangent_rhs:  out = (dRHS/dU) dU           any scheme combination
 *   adjoint_rhs:  out = (dRHS/dU)^T lam        component-wise WCNS5,
 *                                              WENO5-JS or MDCD, any solver
 *   tangent_step: advances the state one step and dU to its tangent
 *   adjoint_step: adj = (dU^{n+1}/dU^n)^T adj; the state stays U^n
 * The adjoint entry points return NS3D_E_UNSUPPORTED for
 * characteristic-wise interpolation. */
int ns3d_tangent_rhs(ns3d_ctx *ctx, const double *const *d_dU, double *const *d_out);
int ns3d_adjoint_rhs(ns3d_ctx *ctx, const double *const *d_lam, double *const *d_out);
int ns3d_tangent_step(ns3d_ctx *ctx, double dt, double *const *d_dU);
int ns3d_adjoint_step(ns3d_ctx *ctx, double dt, double *const *d_adj);
/* The same step pair with the RK3 stage states kept by the caller (five
 * ghosted device fields each), so that the adjoint need not recompute them:
 *   step_stages:         recordConservedTo0 + one runge_kutta_3 step with dt,
 *                        storing U^(1) and U^(2) (interior cells; their ghost
 *                        layers are not written);
 *   adjoint_step_stages: adjoint_step of the step from the current state U^n
 *                        whose stages step_stages stored; the state stays U^n. */
int ns3d_step_stages(ns3d_ctx *ctx, double dt, double *const *d_U1, double *const *d_U2);
int ns3d_adjoint_step_stages(ns3d_ctx *ctx, double dt, double *const *d_adj, const double *const *d_U1,
                             const double *const *d_U2);

/* ---- host-side derivative checks (no device work) ----
 * The device derivative numerics on one stencil, evaluated on the host:
 * face: x-normal 6-cell conserved stencil st[var*6 + m], cells i-2..i+3,
 * giving F (5) and its tangent dF along dst, or the reverse stb = J^T Fb;
 * viscous: (u, v, w, T) q, gradients g[3*q + d], the nine flux components
 * (slots xx, yy, zz, xy, xz, yz, energy x, y, z) and their tangent, or the
 * reverse (gb, qb); c2p: Ub = (d(u, v, w, T)/dU)^T qb. */
int ns3d_check_face_tangent(const ns3d_params *params, const double *st, const double *dst, double *F,
                            double *dF);
int ns3d_check_face_adjoint(const ns3d_params *params, const double *st, const double *Fb, double *stb);
int ns3d_check_visc_tangent(const ns3d_params *params, const double *q, const double *g, const double *dq,
                            const double *dg, double *V, double *dV);
int ns3d_check_visc_adjoint(const ns3d_params *params, const double *q, const double *g, const double *Vb,
                            double *gb, double *qb);
int ns3d_check_c2p_adjoint(const ns3d_params *params, const double *U, const double *qb, double *Ub);
/* Reverse of both reconstructions of one 5-cell stencil f (cells c-2..c+2;
 * left state of face c+1/2 and right state of face c-1/2) with shared
 * smoothness indicators (fb_pair), and the same from two separate
 * reconstructions (fb_single). */
int ns3d_check_weno_pair_adjoint(const ns3d_params *params, const double *f, double ybL, double ybR,
                                 double *fb_pair, double *fb_single);

#ifdef __cplusplus
}
#endif

#endif /* NS3D_HIP_H */
