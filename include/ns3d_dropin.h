/* ns3d_dropin.h -- single-GPU drop-in for the reference's hot-path call
 * surface (C++).
 *
 * The types and free functions below carry the names, members and argument
 * lists of the reference's field_structures.h / ns3d_func.h for the RHS +
 * RK3 path (GridDesc, LocalDesc, SolverParams, Field3D, runge_kutta_3,
 * compute_rhs, compute_timestep, ...), so that a time-advance loop written
 * against them (maincycle.cpp:7-113) compiles and runs unchanged on one GPU
 * of one periodic box.  The implementation (libns3d_dropin.so) keeps a
 * device mirror of each Field3D in libns3d_hip (include/ns3d_hip.h).
 *
 * Coherence.  By default every hot-path function uploads the host fields it
 * reads and downloads the fields the reference writes (faithful mode: the
 * host vectors are the state between calls).  ns3d_set_resident(F, true)
 * keeps the state on the device instead: runge_kutta_3, compute_timestep,
 * compute_rhs and the Field3D members act on the mirror, and the host
 * vectors are refreshed by ns3d_sync_host(F) (before diagnostics or output)
 * and pushed by ns3d_sync_device(F).
 *
 * One rank: CartDecomp describes the single rank (periodic along an axis
 * whose two faces are Periodic) and HaloRequests has nothing to wait for;
 * the six boundary types of SolverParams select the device's boundary
 * fills.  Errors are printed to std::cerr with an "ns3d:" prefix and the
 * function returns, as the reference's functions do
 * (InviscousFlux.cpp:19-31); interpolate_select throws std::runtime_error on
 * a stencil of the wrong size, as the reference's does
 * (interpolation.cpp:18-24).
 *
 * The scalar functions of ns3d_func.h:73-131 (interpolate_select,
 * weno5_interpolate, the MDCD interpolations, weno5_reconstruction, the six
 * X_Riemann_solver and WCNS_Riemann_InviscidFlux) evaluate one stencil or
 * one face on the host with the numerics the kernels run
 * (ns3d_scalar_* in include/ns3d_hip.h); the solvers take the unit normals
 * along x, y or z that the reference's flux loops pass.
 *
 * include/compat/ holds field_structures.h and ns3d_func.h, which include
 * this header under the reference's file names, so that a translation unit
 * written against them (maincycle.cpp) compiles unchanged.
 */
#ifndef NS3D_DROPIN_H
#define NS3D_DROPIN_H

#include <array>
#include <string>
#include <vector>

struct GridDesc {  // field_structures.h:33-38
    int global_nx = 0, global_ny = 0, global_nz = 0;
    double x0 = 0., y0 = 0., z0 = 0.;
    double Lx = 1., Ly = 1., Lz = 1.;
    double dx = 1., dy = 1., dz = 1.;
};

struct CartDecomp {  // one rank of a periodic 1x1x1 decomposition
    int rank = 0;
    int size = 1;
    int dims[3] = {1, 1, 1};
    int coords[3] = {0, 0, 0};
    int periods[3] = {1, 1, 1};
};

struct LocalDesc {  // field_structures.h:50-60
    int nx = 0, ny = 0, nz = 0;
    int ngx = 1, ngy = 1, ngz = 1;
    int sx = 0, sy = 0, sz = 0;
    int ox = 0, oy = 0, oz = 0;
    int nbr_xm = 0, nbr_xp = 0, nbr_ym = 0, nbr_yp = 0, nbr_zm = 0, nbr_zp = 0;
};

inline int idx3(int i, int j, int k, const LocalDesc &L) noexcept { return (k * L.sy + j) * L.sx + i; }
inline int idx_fx(int i, int j, int k, const LocalDesc &L) noexcept { return (k * L.sy + j) * (L.sx - 1) + i; }
inline int idx_fy(int i, int j, int k, const LocalDesc &L) noexcept { return (k * (L.sy - 1) + j) * L.sx + i; }
inline int idx_fz(int i, int j, int k, const LocalDesc &L) noexcept { return (k * L.sy + j) * L.sx + i; }

struct SolverParams {  // numerical and control fields of field_structures.h:86-186
    double gamma = 1.4;
    double Pr = 0.71;
    double Ma = 0.4;
    double Re = 500.0;
    double cfl = 0.1;
    double dt_fixed = -1.0;
    double mu = 1.0 / Re;
    double S_ref = 110.4 / 288.15;
    double Cv = 1.0 / (gamma * (gamma - 1.0) * Ma * Ma);
    double Cp = Cv * gamma;
    double Rgas = 1.0 / (Ma * Ma * gamma);
    double get_mu(double T) const;  // Sutherland
    bool use_periodic = false;      // carried, not read (as in the reference)
    // carried, not read: the flux-vector-splitting, reconstruction and
    // viscous-scheme selectors of the reference's unused FD paths
    enum class FVS_Type { StegerWarming, LaxFriedrichs, Rusanov, VanLeer };
    FVS_Type fvs_type = FVS_Type::StegerWarming;
    enum class Reconstruction { WENO5, WCNS, LINEAR, MDCD };
    enum class ViscousScheme { C6th, C4th };
    // WENO5_JS (weno5_reconstruction's Jiang-Shu weights) is this library's
    // addition to the reference's four
    enum class Interpolation { WENO5, ZERO, MDCD_LINEAR, MDCD_HYBRID, WENO5_JS };
    enum class RiemannSolver { Roe, Rusanov, HLLC, HLL, HLLC_p, AUSM };
    enum class BCType { Periodic, Wall, Symmetry, Outflow, Inflow };
    RiemannSolver riemann_solver = RiemannSolver::Roe;
    Reconstruction recon = Reconstruction::WENO5;
    Interpolation interpolation = Interpolation::WENO5;
    ViscousScheme vis_scheme = ViscousScheme::C6th;
    double mdcd_diss = 0.01;
    double mdcd_disp = 0.0463783;
    int stencil = 6;
    int ghost_layers = 3;
    bool char_recon = false;
    BCType bc_xmin = BCType::Periodic, bc_xmax = BCType::Periodic;
    BCType bc_ymin = BCType::Periodic, bc_ymax = BCType::Periodic;
    BCType bc_zmin = BCType::Periodic, bc_zmax = BCType::Periodic;
    int max_steps = 1;
    int monitor_Stepfreq = 1;
    double output_Timefreq = 1.0;
    double TotalTime = 1.0;
    bool monitor_res = true;
    bool monitor_energy = true;
    bool post_basicfield = true;         // field.h5 at the output times (maincycle.cpp:94-98)
    bool isotropic_analyse = true;       // parsed; the spectral post-processing is not part of this library
    bool restart = false;                // initial state from restart_file (main.cpp:69-84)
    std::string restart_file = "field.h5";
};

struct HaloRequests {};  // nothing outstanding on one GPU

struct ns3d_mirror;  // device state of one Field3D (ns3d_dropin.cpp)

// The data members of field_structures.h:196-245 (Field3D adds the device
// mirror and the member functions).
struct ns3d_field_data {
    LocalDesc L;
    std::vector<double> rho, rhou, rhov, rhow, E;
    std::vector<double> rho0, rhou0, rhov0, rhow0, E0;
    std::vector<double> res_rho, res_rhou, res_rhov, res_rhow, res_E;  // U - U0 (updateResiduals)
    double global_res_rho = 0.0, global_res_rhou = 0.0, global_res_rhov = 0.0, global_res_rhow = 0.0,
           global_res_E = 0.0;  // compute_diagnostics
    double global_Etot = 0.0;   // compute_total_energy
    std::vector<double> u, v, w, p, T;
    std::vector<double> rhs_rho, rhs_rhou, rhs_rhov, rhs_rhow, rhs_E;
    // cell-centre flux arrays of the reference's FD path (allocated, not written)
    std::vector<double> Fflux_mass, Fflux_momx, Fflux_momy, Fflux_momz, Fflux_E;
    std::vector<double> Hflux_mass, Hflux_momx, Hflux_momy, Hflux_momz, Hflux_E;
    std::vector<double> Gflux_mass, Gflux_momx, Gflux_momy, Gflux_momz, Gflux_E;
    // written by compute_viscous_flux and compute_gradients
    std::vector<double> Fvflux_mass, Fvflux_momx, Fvflux_momy, Fvflux_momz, Fvflux_E;
    std::vector<double> Hvflux_mass, Hvflux_momx, Hvflux_momy, Hvflux_momz, Hvflux_E;
    std::vector<double> Gvflux_mass, Gvflux_momx, Gvflux_momy, Gvflux_momz, Gvflux_E;
    std::vector<double> du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz, dT_dx, dT_dy, dT_dz;
    // face arrays (idx_fx / idx_fy / idx_fz), written by compute_invis_flux
    // and apply_boundary_halfnode_flux
    std::vector<double> flux_fx_mass, flux_fx_momx, flux_fx_momy, flux_fx_momz, flux_fx_E;
    std::vector<double> flux_fy_mass, flux_fy_momx, flux_fy_momy, flux_fy_momz, flux_fy_E;
    std::vector<double> flux_fz_mass, flux_fz_momx, flux_fz_momy, flux_fz_momz, flux_fz_E;
    // half-node interpolated values of the reference's FD path (allocated, not written)
    std::vector<double> interp_fx_mass, interp_fx_momx, interp_fx_momy, interp_fx_momz, interp_fx_E;
    std::vector<double> interp_fy_mass, interp_fy_momx, interp_fy_momy, interp_fy_momz, interp_fy_E;
    std::vector<double> interp_fz_mass, interp_fz_momx, interp_fz_momy, interp_fz_momz, interp_fz_E;
};

struct Field3D : ns3d_field_data {  // field_structures.h:196-509
    Field3D() = default;
    Field3D(const Field3D &o) : ns3d_field_data(o) {}  // host copy; the copy has no device mirror
    Field3D &operator=(const Field3D &) = delete;
    ~Field3D();

    // sx = nx + 2 ngx, ...; every array of the reference's allocate at its
    // size: rho = 1, everything else 0 (field_structures.h:247-372)
    void allocate(const LocalDesc &Ld);
    int I(int i, int j, int k) const noexcept { return idx3(i, j, k, L); }
    void conservedToPrimitive(const SolverParams &par);
    void primitiveToConserved(const SolverParams &par);
    void recordConservedTo0();
    // res = U - U0 over every cell (field_structures.h:500-508); a resident
    // field first refreshes its host conserved and recorded fields
    void updateResiduals();

    // element accessors (field_structures.h:374-437)
    double &Rho(int i, int j, int k) noexcept { return rho[I(i, j, k)]; }
    double &RhoU(int i, int j, int k) noexcept { return rhou[I(i, j, k)]; }
    double &RhoV(int i, int j, int k) noexcept { return rhov[I(i, j, k)]; }
    double &RhoW(int i, int j, int k) noexcept { return rhow[I(i, j, k)]; }
    double &Eint(int i, int j, int k) noexcept { return E[I(i, j, k)]; }
    double &U(int i, int j, int k) noexcept { return u[I(i, j, k)]; }
    double &V(int i, int j, int k) noexcept { return v[I(i, j, k)]; }
    double &W(int i, int j, int k) noexcept { return w[I(i, j, k)]; }
    double &P(int i, int j, int k) noexcept { return p[I(i, j, k)]; }
    double &Temp(int i, int j, int k) noexcept { return T[I(i, j, k)]; }
    double &RHS_Rho(int i, int j, int k) noexcept { return rhs_rho[I(i, j, k)]; }
    double &RHS_RhoU(int i, int j, int k) noexcept { return rhs_rhou[I(i, j, k)]; }
    double &RHS_RhoV(int i, int j, int k) noexcept { return rhs_rhov[I(i, j, k)]; }
    double &RHS_RhoW(int i, int j, int k) noexcept { return rhs_rhow[I(i, j, k)]; }
    double &RHS_E(int i, int j, int k) noexcept { return rhs_E[I(i, j, k)]; }
    double &FX_mass(int f, int j, int k) noexcept { return flux_fx_mass[idx_fx(f, j, k, L)]; }
    double &FX_momx(int f, int j, int k) noexcept { return flux_fx_momx[idx_fx(f, j, k, L)]; }
    double &FX_momy(int f, int j, int k) noexcept { return flux_fx_momy[idx_fx(f, j, k, L)]; }
    double &FX_momz(int f, int j, int k) noexcept { return flux_fx_momz[idx_fx(f, j, k, L)]; }
    double &FX_E(int f, int j, int k) noexcept { return flux_fx_E[idx_fx(f, j, k, L)]; }
    double &FY_mass(int i, int f, int k) noexcept { return flux_fy_mass[idx_fy(i, f, k, L)]; }
    double &FY_momx(int i, int f, int k) noexcept { return flux_fy_momx[idx_fy(i, f, k, L)]; }
    double &FY_momy(int i, int f, int k) noexcept { return flux_fy_momy[idx_fy(i, f, k, L)]; }
    double &FY_momz(int i, int f, int k) noexcept { return flux_fy_momz[idx_fy(i, f, k, L)]; }
    double &FY_E(int i, int f, int k) noexcept { return flux_fy_E[idx_fy(i, f, k, L)]; }
    double &FZ_mass(int i, int j, int f) noexcept { return flux_fz_mass[idx_fz(i, j, f, L)]; }
    double &FZ_momx(int i, int j, int f) noexcept { return flux_fz_momx[idx_fz(i, j, f, L)]; }
    double &FZ_momy(int i, int j, int f) noexcept { return flux_fz_momy[idx_fz(i, j, f, L)]; }
    double &FZ_momz(int i, int j, int f) noexcept { return flux_fz_momz[idx_fz(i, j, f, L)]; }
    double &FZ_E(int i, int j, int f) noexcept { return flux_fz_E[idx_fz(i, j, f, L)]; }

    ns3d_mirror *dev = nullptr;
};

// solver.in reader with the reference's keys and derived values
// (initialization.cpp:36-200); CartDecomp gets the periodicity.
bool read_solver_params_from_file(const std::string &fname, SolverParams &P, GridDesc &G, CartDecomp &C);

// The one-rank local description of G (ghost layers from P).
LocalDesc make_local_desc(const GridDesc &G, const SolverParams &P);

// Hot path (ns3d_func.h:60-140).
void apply_boundary(Field3D &F, GridDesc &G, CartDecomp &C, const SolverParams &P);
void apply_boundary_halfnode_flux(Field3D &F, const GridDesc &G, CartDecomp &C, const SolverParams &P);
void compute_invis_flux(Field3D &F, const SolverParams &P);
void compute_invis_dflux(Field3D &F, const SolverParams &P, const GridDesc &G);
void compute_gradients(Field3D &F, const GridDesc &G);
void compute_viscous_flux(Field3D &F, const SolverParams &P);
void exchange_halos_viscous_flux(Field3D &F, CartDecomp &C, LocalDesc &L, HaloRequests &reqs);
void compute_vis_flux(Field3D &F, const GridDesc &G);
void compute_rhs(Field3D &F, CartDecomp &C, GridDesc &G, SolverParams &P, HaloRequests &reqs);
void runge_kutta_3(Field3D &F, CartDecomp &C, GridDesc &G, SolverParams &P, HaloRequests &reqs, double dt);
double compute_timestep(Field3D &F, const GridDesc &G, const SolverParams &P);
// sum 0.5 rho |u|^2 dV over the interior into F.global_Etot (postprocess.cpp:103-128)
void compute_total_energy(Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &P);

// Scalar functions (ns3d_func.h:73-131): one stencil or one face on the host.
void WCNS_Riemann_InviscidFlux(std::vector<double> &Fface, const std::vector<std::vector<double>> &Ut,
                               const std::vector<std::vector<double>> &ut, const SolverParams &P, int dim);
void Roe_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL, const std::vector<double> &UR,
                        double nx, double ny, double nz, double gamma);
void Rusanov_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL,
                            const std::vector<double> &UR, double nx, double ny, double nz, double gamma);
void HLLC_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL, const std::vector<double> &UR,
                         double nx, double ny, double nz, double gamma);
void HLLC_p_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL,
                           const std::vector<double> &UR, double nx, double ny, double nz, double gamma);
void HLL_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL, const std::vector<double> &UR,
                        double nx, double ny, double nz, double gamma);
void AUSM_Riemann_solver(std::vector<double> &Fface, const std::vector<double> &UL, const std::vector<double> &UR,
                         double nx, double ny, double nz, double gamma);
double interpolate_select(const std::vector<double> &vstencil, double flag, const SolverParams P);
double weno5_interpolate(const std::array<double, 6> &stencil);
double zero_interpolate(const std::array<double, 2> &stencil);
double mdcd_linear_interpolate(const std::array<double, 6> &stencil, SolverParams P);
double mdcd_hybrid_interpolate(const std::array<double, 6> &stencil, SolverParams P);
double weno5_reconstruction(const std::array<double, 6> &stencil);
// reconstruction.cpp:13-124, not on the reference's path (no caller): the
// P.recon selector with the reference's size checks (std::runtime_error),
// reversal for flag < 0 and fallbacks; 2-point linear (stencil[0]); the
// 6-point MDCD form as the reference evaluates it, with C++ integer
// division in its constants (1/2, 1/12, 7/12 -> 0; 3/2 -> 1; 5/2 -> 2); the
// 6-point (1, -8, 37, 37, -8, 1)/60 and 4-point (-1, 7, 7, -1)/12 central
// interpolations.
double reconstruct_select(const std::vector<double> &vstencil, double flag, const SolverParams P);
double linear_reconstruction(const std::array<double, 2> &stencil);
double mdcd_reconstruction(const std::array<double, 6> &stencil, SolverParams P);
double c6th_reconstruction(const std::array<double, 6> &stencil);
double c4th_reconstruction(const std::array<double, 4> &stencil);

// Monitoring and post-processing (postprocess.cpp, output.cpp:208-242,
// EnergySpectrum.cpp, ana_isotropic.cpp:196-340).
// updateResiduals, then global_res_X = sqrt(sum over the interior of
// res_X^2 dx^3) / max |X| (postprocess.cpp:7-101).
void compute_diagnostics(Field3D &F, const SolverParams &P, const GridDesc &G);
// output/<filename>: "step res_rho res_rhou res_rhov res_rhow res_E Etot",
// the title and variable lines first and the file truncated at step 1.
void write_residuals_tecplot(const Field3D &F, int step, const std::string &filename = "residuals.dat");
// E(k), k = 0 .. 2 global_nx, of the velocity field (forward FFT over the
// box normalised by 1/N, 0.5 (|U|^2 + |V|^2 + |W|^2) summed over shells
// lround(|k|)), one "k E" line each at 16 significant digits; computed on
// the device.
void compute_energy_spectrum(const Field3D &F, const GridDesc &G, const CartDecomp &C,
                             const std::string &filename = "Energy-spectrum.dat");
// compute_turbulence_statistics (ana_isotropic.cpp:196-340): kinetic energy,
// TKE, u_rms, dissipation 2 nu sum k^2 E(k), Taylor, Kolmogorov and
// Taylor-Li scales, Re_lambda and Mach_t, appended to
// output/turbulence_stats.dat (header when the file is new) and printed.
void compute_turbulence_statistics(Field3D &F, const GridDesc &G, const SolverParams &P, const CartDecomp &C,
                                   const double current_time);
// output/output_time_<t, 5 decimals>_spectrum.dat, then the statistics
// (postprocess.cpp:133-145).
void isotropic_post_process(Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &P,
                            const double current_time);

// field.h5 output and restart with serial HDF5 on the one rank: attributes
// global_nx/ny/nz (int), time, dx, dy, dz (double); datasets X Y Z rho u v w
// E p T of shape (global_nz, global_ny, global_nx), IEEE F64LE, x fastest
// (output.cpp:19-179, initialization.cpp:270-420).  The HDF5 library is
// loaded when first needed (NS3D_HDF5_LIBRARY, else libhdf5.so.103 from the
// loader path or /opt/conda/lib); without it these print the reason and
// return false, as the reference's serial-HDF5 build does.  A device-resident
// Field3D is synchronised to the host before it is written and pushed to the
// device after it is read.
bool write_hdf5_field(const Field3D &F, const GridDesc &G, const CartDecomp &C, const std::string &filepath,
                      double time);
// output/time_<int(time * 1e5)>/field.h5 (output.cpp:181-203)
void write_tecplot_field(const Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &P,
                         double time = 0.0);
// interior rho, u, v, w, E, p, T from the file; rho u = rho * u, ... (initialization.cpp:270-420)
bool initialize_from_hdf5(Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &P,
                          const std::string &filename);

// Built-in initial conditions of the reference's driver (primitives rho, u, v,
// w, p; the caller applies apply_boundary and primitiveToConserved, as
// main.cpp:98-99 does).
// initialize_riemann_2d (initialization.cpp:516-568): four quadrants about
// x = y = 0.5, on every cell including the ghost layers; the reference's
// default run (solver.in, main.cpp:85).
void initialize_riemann_2d(Field3D &F, const GridDesc &G, const SolverParams &P);
// initialize_sod_shock_tube (initialization.cpp:570-599): Sod's tube along
// x, interior cells.
void initialize_sod_shock_tube(Field3D &F, const GridDesc &G, const SolverParams &P);

// Mirror control (no counterpart in the reference).
bool ns3d_is_resident(const Field3D &F);
void ns3d_set_resident(Field3D &F, bool resident);
void ns3d_sync_host(Field3D &F);    // device state -> host vectors (conserved, primitives, RHS)
void ns3d_sync_device(Field3D &F);  // host conserved fields -> device (primitives and ghosts refreshed)

#endif /* NS3D_DROPIN_H */
