// ns3d_dropin.cpp -- the reference's hot-path call surface (include/ns3d_dropin.h)
// over the libns3d_hip C-ABI (include/ns3d_hip.h).
//
// Each Field3D owns a device mirror: a libns3d_hip context created on first
// use with the grid spacing of the call that creates it (recreated when a
// later call brings a different spacing) and the parameters of every call.
// Faithful mode uploads the host fields a function reads and downloads the
// fields the reference function writes; resident mode leaves the state on
// the device until ns3d_sync_host.
#include "../../include/ns3d_dropin.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "../../include/ns3d_hip.h"

struct ns3d_mirror {
    ns3d_ctx *ctx = nullptr;
    double dx = 0, dy = 0, dz = 0;
    bool grid_known = false;
    bool resident = false;
    int bc[6] = {0, 0, 0, 0, 0, 0};  // boundary types of the context
};

namespace {

bool ok(int rc, const char *what) {
    if (rc == NS3D_OK) return true;
    std::cerr << "ns3d: " << what << ": " << ns3d_last_error() << "\n";
    return false;
}

ns3d_params to_c(const SolverParams &P) {
    ns3d_params c{};
    c.gamma = P.gamma;
    c.Pr = P.Pr;
    c.Ma = P.Ma;
    c.Re = P.Re;
    c.cfl = P.cfl;
    c.mu = P.mu;
    c.S_ref = P.S_ref;
    c.Cv = P.Cv;
    c.Cp = P.Cp;
    c.Rgas = P.Rgas;
    c.mdcd_diss = P.mdcd_diss;
    c.mdcd_disp = P.mdcd_disp;
    c.interpolation = static_cast<int>(P.interpolation);
    c.riemann_solver = static_cast<int>(P.riemann_solver);
    c.char_recon = P.char_recon ? 1 : 0;
    const SolverParams::BCType bc[6] = {P.bc_xmin, P.bc_xmax, P.bc_ymin, P.bc_ymax, P.bc_zmin, P.bc_zmax};
    for (int f = 0; f < 6; ++f) c.bc[f] = static_cast<int>(bc[f]);
    return c;
}

// The field's mirror with the given spacing and boundary types (nullptr:
// keep the current ones or, for a new mirror, use unit spacing and periodic
// faces until a call supplies them).
ns3d_ctx *mirror(Field3D &F, const SolverParams *P, const GridDesc *G) {
    if (!F.dev) F.dev = new ns3d_mirror();
    ns3d_mirror &m = *F.dev;
    const ns3d_params pc = to_c(P ? *P : SolverParams());
    bool rebc = false;
    if (P)
        for (int f = 0; f < 6; ++f) rebc = rebc || pc.bc[f] != m.bc[f];
    const bool regrid = (G && (!m.grid_known || m.dx != G->dx || m.dy != G->dy || m.dz != G->dz)) || rebc;
    if (m.ctx && regrid) {
        if (m.resident) {
            std::cerr << "ns3d: a resident field was used with a different grid spacing; its state is reloaded "
                         "from the host vectors\n";
        }
        ns3d_destroy(m.ctx);
        m.ctx = nullptr;
    }
    if (!m.ctx) {
        if (G) {
            m.dx = G->dx;
            m.dy = G->dy;
            m.dz = G->dz;
            m.grid_known = true;
        } else if (!m.grid_known) {
            m.dx = m.dy = m.dz = 1.0;
        }
        ns3d_grid g{F.L.nx, F.L.ny, F.L.nz, F.L.ngx, m.dx, m.dy, m.dz};
        ns3d_params cp = pc;
        if (!P)
            for (int f = 0; f < 6; ++f) cp.bc[f] = m.bc[f];
        if (!ok(ns3d_create(&m.ctx, &g, &cp, NS3D_PRECISION_FP64), "create")) return nullptr;
        for (int f = 0; f < 6; ++f) m.bc[f] = cp.bc[f];
        if (m.resident) {  // a new context starts from the host state
            m.resident = false;
            ns3d_sync_device(F);
            m.resident = true;
        }
    }
    if (P) {
        const ns3d_params cp = to_c(*P);
        if (!ok(ns3d_set_params(m.ctx, &cp), "set_params")) return nullptr;
    }
    return m.ctx;
}

bool resident(const Field3D &F) { return F.dev && F.dev->resident; }

std::vector<double> *conserved(Field3D &F, int n) {
    std::vector<double> *f[5] = {&F.rho, &F.rhou, &F.rhov, &F.rhow, &F.E};
    return f[n];
}
std::vector<double> *recorded(Field3D &F, int n) {
    std::vector<double> *f[5] = {&F.rho0, &F.rhou0, &F.rhov0, &F.rhow0, &F.E0};
    return f[n];
}
std::vector<double> *prims(Field3D &F, int n) {
    std::vector<double> *f[5] = {&F.u, &F.v, &F.w, &F.p, &F.T};
    return f[n];
}
std::vector<double> *rhs(Field3D &F, int n) {
    std::vector<double> *f[5] = {&F.rhs_rho, &F.rhs_rhou, &F.rhs_rhov, &F.rhs_rhow, &F.rhs_E};
    return f[n];
}

bool put(ns3d_ctx *c, int field, std::vector<double> &h) { return ok(ns3d_upload(c, field, h.data()), "upload"); }
bool get(ns3d_ctx *c, int field, std::vector<double> &h) { return ok(ns3d_download(c, field, h.data()), "download"); }

// faithful mode: the host state (conserved, primitives; optionally the
// recorded copy and the RHS) to the device
bool push(Field3D &F, ns3d_ctx *c, bool with_record, bool with_rhs) {
    if (resident(F)) return true;
    for (int n = 0; n < 5; ++n) {
        if (!put(c, NS3D_RHO + n, *conserved(F, n)) || !put(c, NS3D_U + n, *prims(F, n))) return false;
        if (with_record && !put(c, NS3D_RHO0 + n, *recorded(F, n))) return false;
        if (with_rhs && !put(c, NS3D_RHS_RHO + n, *rhs(F, n))) return false;
    }
    return true;
}

bool pull(Field3D &F, ns3d_ctx *c, bool state, bool with_rhs) {
    if (resident(F)) return true;
    for (int n = 0; n < 5; ++n) {
        if (state && (!get(c, NS3D_RHO + n, *conserved(F, n)) || !get(c, NS3D_U + n, *prims(F, n)))) return false;
        if (with_rhs && !get(c, NS3D_RHS_RHO + n, *rhs(F, n))) return false;
    }
    return true;
}

size_t cells(const LocalDesc &L) { return (size_t)L.sx * L.sy * L.sz; }

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    return s;
}

std::string trim(const std::string &s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

}  // namespace

// ------------------------------------------------------------ SolverParams
double SolverParams::get_mu(double T) const { return mu * std::pow(T, 1.5) * (1.0 + S_ref) / (T + S_ref); }

// ------------------------------------------------------------ Field3D
Field3D::~Field3D() {
    if (dev) {
        if (dev->ctx) ns3d_destroy(dev->ctx);
        delete dev;
    }
}

void Field3D::allocate(const LocalDesc &Ld) {
    L = Ld;
    L.sx = L.nx + 2 * L.ngx;
    L.sy = L.ny + 2 * L.ngy;
    L.sz = L.nz + 2 * L.ngz;
    const size_t n = cells(L);
    rho.assign(n, 1.0);
    for (auto *f : {&rhou, &rhov, &rhow, &E, &rho0, &rhou0, &rhov0, &rhow0, &E0, &res_rho, &res_rhou, &res_rhov,
                    &res_rhow, &res_E, &u, &v, &w, &p, &T, &rhs_rho, &rhs_rhou, &rhs_rhov, &rhs_rhow, &rhs_E,
                    &Fflux_mass, &Fflux_momx, &Fflux_momy, &Fflux_momz, &Fflux_E, &Hflux_mass, &Hflux_momx,
                    &Hflux_momy, &Hflux_momz, &Hflux_E, &Gflux_mass, &Gflux_momx, &Gflux_momy, &Gflux_momz,
                    &Gflux_E, &Fvflux_mass, &Fvflux_momx, &Fvflux_momy, &Fvflux_momz, &Fvflux_E, &Hvflux_mass,
                    &Hvflux_momx, &Hvflux_momy, &Hvflux_momz, &Hvflux_E, &Gvflux_mass, &Gvflux_momx, &Gvflux_momy,
                    &Gvflux_momz, &Gvflux_E, &du_dx, &du_dy, &du_dz, &dv_dx, &dv_dy, &dv_dz, &dw_dx, &dw_dy,
                    &dw_dz, &dT_dx, &dT_dy, &dT_dz})
        f->assign(n, 0.0);
    const size_t nf[3] = {(size_t)(L.sx - 1) * L.sy * L.sz, (size_t)L.sx * (L.sy - 1) * L.sz,
                          (size_t)L.sx * L.sy * (L.sz - 1)};
    std::vector<double> *faces[3][10] = {
        {&flux_fx_mass, &flux_fx_momx, &flux_fx_momy, &flux_fx_momz, &flux_fx_E, &interp_fx_mass, &interp_fx_momx,
         &interp_fx_momy, &interp_fx_momz, &interp_fx_E},
        {&flux_fy_mass, &flux_fy_momx, &flux_fy_momy, &flux_fy_momz, &flux_fy_E, &interp_fy_mass, &interp_fy_momx,
         &interp_fy_momy, &interp_fy_momz, &interp_fy_E},
        {&flux_fz_mass, &flux_fz_momx, &flux_fz_momy, &flux_fz_momz, &flux_fz_E, &interp_fz_mass, &interp_fz_momx,
         &interp_fz_momy, &interp_fz_momz, &interp_fz_E}};
    for (int d = 0; d < 3; ++d)
        for (auto *f : faces[d]) f->assign(nf[d], 0.0);
    if (dev && dev->ctx) {
        ns3d_destroy(dev->ctx);
        dev->ctx = nullptr;
    }
}

void Field3D::conservedToPrimitive(const SolverParams &par) {
    if (resident(*this)) {
        ns3d_ctx *c = mirror(*this, &par, nullptr);
        if (c) ok(ns3d_conserved_to_primitive(c), "conservedToPrimitive");
        return;
    }
    const double g = par.gamma;
    const size_t n = cells(L);
    for (size_t id = 0; id < n; ++id) {
        double r = rho[id];
        if (r <= 0.0) {
            r = 1e-12;
            rho[id] = r;
        }
        u[id] = rhou[id] / r;
        v[id] = rhov[id] / r;
        w[id] = rhow[id] / r;
        const double kin = 0.5 * (u[id] * u[id] + v[id] * v[id] + w[id] * w[id]);
        p[id] = (g - 1.0) * r * (E[id] / r - kin);
        T[id] = p[id] / (r * par.Rgas);
    }
}

void Field3D::primitiveToConserved(const SolverParams &par) {
    if (resident(*this)) {
        ns3d_ctx *c = mirror(*this, &par, nullptr);
        if (c) ok(ns3d_primitive_to_conserved(c), "primitiveToConserved");
        return;
    }
    const double g = par.gamma;
    const size_t n = cells(L);
    for (size_t id = 0; id < n; ++id) {
        double r = rho[id];
        if (r <= 0.0) {
            r = 1e-12;
            rho[id] = r;
        }
        rhou[id] = r * u[id];
        rhov[id] = r * v[id];
        rhow[id] = r * w[id];
        const double kin = 0.5 * (u[id] * u[id] + v[id] * v[id] + w[id] * w[id]);
        E[id] = r * (p[id] / ((g - 1.0) * r) + kin);
        T[id] = p[id] / (r * par.Rgas);
    }
}

void Field3D::recordConservedTo0() {
    if (resident(*this)) {
        ns3d_ctx *c = mirror(*this, nullptr, nullptr);
        if (c) ok(ns3d_record_conserved_to0(c), "recordConservedTo0");
        return;
    }
    rho0 = rho;
    rhou0 = rhou;
    rhov0 = rhov;
    rhow0 = rhow;
    E0 = E;
}

void Field3D::updateResiduals() {
    if (resident(*this) && dev->ctx)
        for (int n = 0; n < 5; ++n)
            if (!get(dev->ctx, NS3D_RHO + n, *conserved(*this, n)) || !get(dev->ctx, NS3D_RHO0 + n, *recorded(*this, n)))
                return;
    std::vector<double> *res[5] = {&res_rho, &res_rhou, &res_rhov, &res_rhow, &res_E};
    for (int n = 0; n < 5; ++n) {
        const std::vector<double> &a = *conserved(*this, n), &b = *recorded(*this, n);
        std::vector<double> &r = *res[n];
        r.resize(a.size());
        for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] - b[i];
    }
}

// ------------------------------------------------------------ set-up
bool read_solver_params_from_file(const std::string &fname, SolverParams &P, GridDesc &G, CartDecomp &C) {
    std::ifstream fin(fname);
    if (!fin) {
        std::cerr << "Cannot open " << fname << "\n";
        return false;
    }
    // defaults first (initialization.cpp:48-52): every field of P, the grid
    // size and origin of G
    P = SolverParams();
    G.global_nx = G.global_ny = G.global_nz = 16;
    G.x0 = G.y0 = G.z0 = 0.0;
    using BC = SolverParams::BCType;
    auto flag = [](const std::string &v) { return v == "true" || v == "ture" || v == "yes" || v == "1" || v == "on"; };
    auto bc_of = [](const std::string &v) {
        if (v == "inflow") return BC::Inflow;
        if (v == "wall") return BC::Wall;
        if (v == "symmetry") return BC::Symmetry;
        if (v == "outflow") return BC::Outflow;
        return BC::Periodic;
    };
    std::string line;
    while (std::getline(fin, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string k = lower(trim(line.substr(0, eq))), val = trim(line.substr(eq + 1)), v = lower(val);
        try {
            if (k == "gamma") P.gamma = std::stod(val);
            else if (k == "pr") P.Pr = std::stod(val);
            else if (k == "ma") P.Ma = std::stod(val);
            else if (k == "re") P.Re = std::stod(val);
            else if (k == "cfl") P.cfl = std::stod(val);
            else if (k == "dt_fixed") P.dt_fixed = std::stod(val);
            else if (k == "interpolation") {
                if (v == "weno5") P.interpolation = SolverParams::Interpolation::WENO5;
                else if (v == "zero") P.interpolation = SolverParams::Interpolation::ZERO;
                else if (v == "mdcd_linear") P.interpolation = SolverParams::Interpolation::MDCD_LINEAR;
                else if (v == "mdcd_hybrid") P.interpolation = SolverParams::Interpolation::MDCD_HYBRID;
                else if (v == "weno5_js") P.interpolation = SolverParams::Interpolation::WENO5_JS;
            } else if (k == "char_recon") P.char_recon = (v == "yes" || v == "true");
            else if (k == "riemann_solver") {
                using RS = SolverParams::RiemannSolver;
                if (v == "roe") P.riemann_solver = RS::Roe;
                else if (v == "rusanov") P.riemann_solver = RS::Rusanov;
                else if (v == "hllc") P.riemann_solver = RS::HLLC;
                else if (v == "hll") P.riemann_solver = RS::HLL;
                else if (v == "hllc_p") P.riemann_solver = RS::HLLC_p;
                else if (v == "ausm") P.riemann_solver = RS::AUSM;
            } else if (k == "mdcd_diss") P.mdcd_diss = std::stod(val);
            else if (k == "mdcd_disp") P.mdcd_disp = std::stod(val);
            else if (k == "global_nx") G.global_nx = std::stoi(val);
            else if (k == "global_ny") G.global_ny = std::stoi(val);
            else if (k == "global_nz") G.global_nz = std::stoi(val);
            else if (k == "x0") G.x0 = std::stod(val);
            else if (k == "y0") G.y0 = std::stod(val);
            else if (k == "z0") G.z0 = std::stod(val);
            else if (k == "lx") G.Lx = std::stod(val);
            else if (k == "ly") G.Ly = std::stod(val);
            else if (k == "lz") G.Lz = std::stod(val);
            else if (k == "max_steps") P.max_steps = std::stoi(val);
            else if (k == "monitor_stepfreq") P.monitor_Stepfreq = std::stoi(val);
            else if (k == "output_timefreq") P.output_Timefreq = std::stod(val);
            else if (k == "totaltime") P.TotalTime = std::stod(val);
            else if (k == "bc_xmin") P.bc_xmin = bc_of(v);
            else if (k == "bc_xmax") P.bc_xmax = bc_of(v);
            else if (k == "bc_ymin") P.bc_ymin = bc_of(v);
            else if (k == "bc_ymax") P.bc_ymax = bc_of(v);
            else if (k == "bc_zmin") P.bc_zmin = bc_of(v);
            else if (k == "bc_zmax") P.bc_zmax = bc_of(v);
            else if (k == "restart") P.restart = flag(v);
            else if (k == "restart_file") P.restart_file = val;
            else if (k == "post_basicfield") P.post_basicfield = flag(v);
            else if (k == "isotropicanalyse") P.isotropic_analyse = flag(v);
            else if (k == "monitor_res") P.monitor_res = flag(v);
            else if (k == "monitor_energy") P.monitor_energy = flag(v);
        } catch (const std::exception &) {
            std::cerr << "ns3d: bad value for " << k << ": " << val << "\n";
            return false;
        }
    }
    C.periods[0] = P.bc_xmin == BC::Periodic && P.bc_xmax == BC::Periodic;
    C.periods[1] = P.bc_ymin == BC::Periodic && P.bc_ymax == BC::Periodic;
    C.periods[2] = P.bc_zmin == BC::Periodic && P.bc_zmax == BC::Periodic;
    // derived values (initialization.cpp:161-197)
    G.dx = G.Lx / G.global_nx;
    G.dy = G.Ly / G.global_ny;
    G.dz = G.Lz / G.global_nz;
    if (P.interpolation == SolverParams::Interpolation::ZERO) {
        P.ghost_layers = 1;
        P.stencil = 2;
    } else {
        P.ghost_layers = 3;
        P.stencil = 6;
    }
    P.Cv = 1.0 / (P.gamma * (P.gamma - 1.0) * P.Ma * P.Ma);
    P.Cp = P.Cv * P.gamma;
    P.Rgas = 1.0 / (P.Ma * P.Ma * P.gamma);
    P.mu = 1.0 / P.Re;
    return true;
}

LocalDesc make_local_desc(const GridDesc &G, const SolverParams &P) {
    LocalDesc L;
    L.nx = G.global_nx;
    L.ny = G.global_ny;
    L.nz = G.global_nz;
    L.ngx = L.ngy = L.ngz = P.ghost_layers;
    L.sx = L.nx + 2 * L.ngx;
    L.sy = L.ny + 2 * L.ngy;
    L.sz = L.nz + 2 * L.ngz;
    return L;
}

// ------------------------------------------------------------ hot path
void apply_boundary(Field3D &F, GridDesc &G, CartDecomp &, const SolverParams &P) {
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !push(F, c, false, false) || !ok(ns3d_apply_boundary(c), "apply_boundary")) return;
    if (resident(F)) return;
    // exchange_halos_physical writes the ghosts of rho, u, v, w, p
    get(c, NS3D_RHO, F.rho);
    for (int n = 0; n < 4; ++n) get(c, NS3D_U + n, *prims(F, n));
}

namespace {

// the device's face arrays into the host's (faithful mode)
void pull_faces(Field3D &F, ns3d_ctx *c) {
    if (resident(F)) return;
    std::vector<double> *ff[3][5] = {
        {&F.flux_fx_mass, &F.flux_fx_momx, &F.flux_fx_momy, &F.flux_fx_momz, &F.flux_fx_E},
        {&F.flux_fy_mass, &F.flux_fy_momx, &F.flux_fy_momy, &F.flux_fy_momz, &F.flux_fy_E},
        {&F.flux_fz_mass, &F.flux_fz_momx, &F.flux_fz_momy, &F.flux_fz_momz, &F.flux_fz_E}};
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) {
            ff[d][n]->resize(ns3d_face_count(c, d));
            if (!ok(ns3d_download_face_flux(c, d, n, ff[d][n]->data()), "face flux")) return;
        }
}

}  // namespace

void apply_boundary_halfnode_flux(Field3D &F, const GridDesc &G, CartDecomp &, const SolverParams &P) {
    // on a periodic axis the faces compute_invis_flux wrote already are the
    // periodic images the reference exchanges; physical faces set their
    // ghost faces on the device
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !ok(ns3d_apply_boundary_halfnode_flux(c), "apply_boundary_halfnode_flux")) return;
    pull_faces(F, c);
}

void exchange_halos_viscous_flux(Field3D &F, CartDecomp &, LocalDesc &, HaloRequests &) {
    // the viscous-flux consumers read periodic images by wrapped indexing and
    // the one-sided stencils of physical faces, so no data moves
    ns3d_ctx *c = mirror(F, nullptr, nullptr);
    if (c) ok(ns3d_exchange_halos_viscous_flux(c), "exchange_halos_viscous_flux");
}

void compute_invis_flux(Field3D &F, const SolverParams &P) {
    ns3d_ctx *c = mirror(F, &P, nullptr);
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_invis_flux(c), "compute_invis_flux")) return;
    pull_faces(F, c);
}

void compute_invis_dflux(Field3D &F, const SolverParams &P, const GridDesc &G) {
    // the device keeps the face fluxes of the last compute_invis_flux
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !push(F, c, false, true) || !ok(ns3d_compute_invis_dflux(c), "compute_invis_dflux")) return;
    pull(F, c, false, true);
}

void compute_gradients(Field3D &F, const GridDesc &G) {
    ns3d_ctx *c = mirror(F, nullptr, &G);
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_gradients(c), "compute_gradients")) return;
    std::vector<double> *g[12] = {&F.du_dx, &F.du_dy, &F.du_dz, &F.dv_dx, &F.dv_dy, &F.dv_dz,
                                  &F.dw_dx, &F.dw_dy, &F.dw_dz, &F.dT_dx, &F.dT_dy, &F.dT_dz};
    for (int q = 0; q < 12; ++q) {
        g[q]->assign(cells(F.L), 0.0);
        if (!ok(ns3d_download_gradient(c, q, g[q]->data()), "gradient")) return;
    }
}

void compute_viscous_flux(Field3D &F, const SolverParams &P) {
    // from the gradients of the last compute_gradients on the device
    ns3d_ctx *c = mirror(F, &P, nullptr);
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_viscous_flux(c), "compute_viscous_flux")) return;
    std::vector<double> *vf[3][5] = {
        {&F.Fvflux_mass, &F.Fvflux_momx, &F.Fvflux_momy, &F.Fvflux_momz, &F.Fvflux_E},
        {&F.Hvflux_mass, &F.Hvflux_momx, &F.Hvflux_momy, &F.Hvflux_momz, &F.Hvflux_E},
        {&F.Gvflux_mass, &F.Gvflux_momx, &F.Gvflux_momy, &F.Gvflux_momz, &F.Gvflux_E}};
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) {
            vf[d][n]->assign(cells(F.L), 0.0);
            if (!ok(ns3d_download_viscous_flux(c, d, n, vf[d][n]->data()), "viscous flux")) return;
        }
}

void compute_vis_flux(Field3D &F, const GridDesc &G) {
    ns3d_ctx *c = mirror(F, nullptr, &G);
    if (!c || !push(F, c, false, true) || !ok(ns3d_compute_vis_flux(c), "compute_vis_flux")) return;
    pull(F, c, false, true);
}

void compute_rhs(Field3D &F, CartDecomp &, GridDesc &G, SolverParams &P, HaloRequests &) {
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_rhs(c), "compute_rhs")) return;
    pull(F, c, false, true);
}

void runge_kutta_3(Field3D &F, CartDecomp &, GridDesc &G, SolverParams &P, HaloRequests &, double dt) {
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !push(F, c, true, false) || !ok(ns3d_runge_kutta_3(c, dt), "runge_kutta_3")) return;
    pull(F, c, true, false);
}

double compute_timestep(Field3D &F, const GridDesc &G, const SolverParams &P) {
    ns3d_ctx *c = mirror(F, &P, &G);
    double dt = 0.0;
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_timestep(c, &dt), "compute_timestep")) return 0.0;
    return dt;
}

void compute_total_energy(Field3D &F, const GridDesc &G, const CartDecomp &, const SolverParams &P) {
    ns3d_ctx *c = mirror(F, &P, &G);
    double e = 0.0;
    if (!c || !push(F, c, false, false) || !ok(ns3d_total_kinetic_energy(c, &e), "compute_total_energy")) return;
    F.global_Etot = e;
}

// ------------------------------------------------------------ mirror control
void initialize_riemann_2d(Field3D &F, const GridDesc &G, const SolverParams &P) {
    (void)P;
    const LocalDesc &L = F.L;
    for (int k = 0; k < L.sz; ++k)
        for (int j = 0; j < L.sy; ++j) {
            const double y = (L.oy + j - L.ngy + 0.5) * G.dy;
            for (int i = 0; i < L.sx; ++i) {
                const double x = (L.ox + i - L.ngx + 0.5) * G.dx;
                // quadrant states (initialization.cpp:536-556)
                const bool right = x >= 0.5, top = y >= 0.5;
                const double rho = right == top ? 1.1 : 0.5065;
                const double u = right ? 0.0 : 0.8939;
                const double v = top ? 0.0 : 0.8939;
                const double p = right == top ? 1.1 : 0.35;
                const int id = F.I(i, j, k);
                F.rho[id] = rho;
                F.u[id] = u;
                F.v[id] = v;
                F.w[id] = 0.0;
                F.p[id] = p;
            }
        }
}

void initialize_sod_shock_tube(Field3D &F, const GridDesc &G, const SolverParams &P) {
    (void)P;
    const LocalDesc &L = F.L;
    const double x_mid = 0.5 * G.global_nx * G.dx;
    for (int k = L.ngz; k < L.ngz + L.nz; ++k)
        for (int j = L.ngy; j < L.ngy + L.ny; ++j)
            for (int i = L.ngx; i < L.ngx + L.nx; ++i) {
                const int id = F.I(i, j, k);
                const double x = (L.ox + i - L.ngx + 0.5) * G.dx;
                const bool left = x < x_mid;
                F.rho[id] = left ? 1.0 : 0.125;
                F.u[id] = 0.0;
                F.v[id] = 0.0;
                F.w[id] = 0.0;
                F.p[id] = left ? 1.0 : 0.1;
            }
}

bool ns3d_is_resident(const Field3D &F) { return resident(F); }

void ns3d_set_resident(Field3D &F, bool on) {
    if (!F.dev) F.dev = new ns3d_mirror();
    if (on == F.dev->resident) return;
    if (on) {
        F.dev->resident = false;
        ns3d_sync_device(F);
        F.dev->resident = true;
    } else {
        ns3d_sync_host(F);
        F.dev->resident = false;
    }
}

void ns3d_sync_host(Field3D &F) {
    if (!F.dev || !F.dev->ctx) return;
    ns3d_ctx *c = F.dev->ctx;
    for (int n = 0; n < 5; ++n)
        if (!get(c, NS3D_RHO + n, *conserved(F, n)) || !get(c, NS3D_RHO0 + n, *recorded(F, n)) ||
            !get(c, NS3D_U + n, *prims(F, n)) || !get(c, NS3D_RHS_RHO + n, *rhs(F, n)))
            return;
}

void ns3d_sync_device(Field3D &F) {
    ns3d_ctx *c = mirror(F, nullptr, nullptr);
    if (!c) return;
    for (int n = 0; n < 5; ++n)
        if (!put(c, NS3D_RHO + n, *conserved(F, n)) || !put(c, NS3D_RHO0 + n, *recorded(F, n)) ||
            !put(c, NS3D_U + n, *prims(F, n)) || !put(c, NS3D_RHS_RHO + n, *rhs(F, n)))
            return;
}

// ------------------------------------------------------------ scalar functions
// One stencil or one face on the host (ns3d_func.h:73-131), through
// ns3d_scalar_* of libns3d_hip.
namespace {

SolverParams with_interp(SolverParams P, SolverParams::Interpolation i) {
    P.interpolation = i;
    return P;
}

double interp_one(const double *st, int n, double flag, const SolverParams &P) {
    const ns3d_params c = to_c(P);
    double out = 0.0;
    if (ns3d_scalar_interpolate(&c, st, n, flag, &out) != NS3D_OK)
        throw std::runtime_error(ns3d_last_error());  // a stencil of the wrong size (interpolation.cpp:18-24)
    return out;
}

// the axis of a unit coordinate normal, or -1
int axis_of(double nx, double ny, double nz) {
    if (nx == 1.0 && ny == 0.0 && nz == 0.0) return 0;
    if (nx == 0.0 && ny == 1.0 && nz == 0.0) return 1;
    if (nx == 0.0 && ny == 0.0 && nz == 1.0) return 2;
    return -1;
}

void riemann_one(SolverParams::RiemannSolver rs, const char *name, std::vector<double> &F,
                 const std::vector<double> &UL, const std::vector<double> &UR, double nx, double ny, double nz,
                 double gamma) {
    const int dim = axis_of(nx, ny, nz);
    if (dim < 0) {
        std::cerr << "ns3d: " << name << ": the face normal (" << nx << ", " << ny << ", " << nz
                  << ") is not (1,0,0), (0,1,0) or (0,0,1)\n";
        return;
    }
    if (UL.size() < 5 || UR.size() < 5) {
        std::cerr << "ns3d: " << name << ": UL and UR need 5 conserved values\n";
        return;
    }
    SolverParams P;
    P.gamma = gamma;
    P.riemann_solver = rs;
    const ns3d_params c = to_c(P);
    if (F.size() < 5) F.resize(5);
    ok(ns3d_scalar_riemann(&c, UL.data(), UR.data(), dim, F.data()), name);
}

}  // namespace

double interpolate_select(const std::vector<double> &vstencil, double flag, const SolverParams P) {
    return interp_one(vstencil.data(), (int)vstencil.size(), flag, P);
}

double weno5_interpolate(const std::array<double, 6> &s) {
    return interp_one(s.data(), 6, 1.0, with_interp(SolverParams(), SolverParams::Interpolation::WENO5));
}

double zero_interpolate(const std::array<double, 2> &s) { return s[0]; }

double mdcd_linear_interpolate(const std::array<double, 6> &s, SolverParams P) {
    return interp_one(s.data(), 6, 1.0, with_interp(P, SolverParams::Interpolation::MDCD_LINEAR));
}

double mdcd_hybrid_interpolate(const std::array<double, 6> &s, SolverParams P) {
    return interp_one(s.data(), 6, 1.0, with_interp(P, SolverParams::Interpolation::MDCD_HYBRID));
}

double weno5_reconstruction(const std::array<double, 6> &s) {
    return interp_one(s.data(), 6, 1.0, with_interp(SolverParams(), SolverParams::Interpolation::WENO5_JS));
}

// reconstruction.cpp:13-124 (host; nothing on the path calls these)
double reconstruct_select(const std::vector<double> &v, double flag, const SolverParams P) {
    const int n = (int)v.size();
    auto require = [&](int want, const char *name) {
        if (n != want)
            throw std::runtime_error(std::string("reconstruct_select: reconstruction '") + name +
                                     " requires stencil size " + std::to_string(want) + ", got " +
                                     std::to_string(n));
    };
    using R = SolverParams::Reconstruction;
    if (P.recon == R::WENO5 || P.recon == R::MDCD) {
        require(6, P.recon == R::WENO5 ? "WENO5" : "MDCD");
        std::array<double, 6> a;
        for (int i = 0; i < 6; ++i) a[i] = v[flag < 0.0 ? 5 - i : i];
        return P.recon == R::WENO5 ? weno5_reconstruction(a) : mdcd_reconstruction(a, P);
    }
    if (P.recon == R::LINEAR) {
        require(2, "LINEAR");
        return linear_reconstruction({v[flag < 0.0 ? 1 : 0], v[flag < 0.0 ? 0 : 1]});
    }
    // WCNS and anything else: the mean of the two cells beside the middle
    if (n >= 2) {
        const int mid = n / 2;
        return flag >= 0.0 ? 0.5 * (v[mid] + v[mid - 1]) : 0.5 * (v[mid] + v[mid + 1]);
    }
    return v.empty() ? 0.0 : v[0];
}

double linear_reconstruction(const std::array<double, 2> &s) { return s[0]; }

double mdcd_reconstruction(const std::array<double, 6> &s, SolverParams P) {
    // the reference's constants are integer quotients: 1/2, 1/12 and 7/12
    // are 0, 3/2 is 1 and 5/2 is 2, so the form is
    // (-disp - 2 diss) s1 + (disp + 5 diss) s2 + (disp - 5 diss) s3 + (-disp + 2 diss) s4
    const double diss = P.mdcd_diss, disp = P.mdcd_disp;
    return (0 * disp + 0 * diss) * s[0] + (-1 * disp - 2 * diss - 0) * s[1] + (0 + disp + 5 * diss) * s[2] +
           (0 + disp - 5 * diss) * s[3] + (-1 * disp + 2 * diss - 0) * s[4] + (0 * disp - 0 * diss) * s[5];
}

double c6th_reconstruction(const std::array<double, 6> &s) {
    return (1.0 / 60.0) * (s[0] - 8 * s[1] + 37 * s[2] + 37 * s[3] - 8 * s[4] + s[5]);
}

double c4th_reconstruction(const std::array<double, 4> &s) { return (1.0 / 12.0) * (-s[0] + 7 * s[1] + 7 * s[2] - s[3]); }

void Roe_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                        double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::Roe, "Roe_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}
void Rusanov_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                            double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::Rusanov, "Rusanov_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}
void HLLC_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                         double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::HLLC, "HLLC_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}
void HLLC_p_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                           double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::HLLC_p, "HLLC_p_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}
void HLL_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                        double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::HLL, "HLL_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}
void AUSM_Riemann_solver(std::vector<double> &F, const std::vector<double> &UL, const std::vector<double> &UR,
                         double nx, double ny, double nz, double gamma) {
    riemann_one(SolverParams::RiemannSolver::AUSM, "AUSM_Riemann_solver", F, UL, UR, nx, ny, nz, gamma);
}

void WCNS_Riemann_InviscidFlux(std::vector<double> &Fface, const std::vector<std::vector<double>> &Ut,
                               const std::vector<std::vector<double>> &ut, const SolverParams &P, int dim) {
    // 5 x 6 row-major stencils (cells i-2..i+3); ut is read for
    // characteristic-wise interpolation only
    double st[30], pt[30] = {};
    if (Ut.size() < 5 || (P.char_recon && ut.size() < 5)) {
        std::cerr << "ns3d: WCNS_Riemann_InviscidFlux: Ut and ut need 5 rows\n";
        return;
    }
    for (int n = 0; n < 5; ++n) {
        if (Ut[n].size() < 6 || (P.char_recon && ut[n].size() < 6)) {
            std::cerr << "ns3d: WCNS_Riemann_InviscidFlux: stencils of 6 cells are needed\n";
            return;
        }
        for (int m = 0; m < 6; ++m) {
            st[6 * n + m] = Ut[n][m];
            if (P.char_recon) pt[6 * n + m] = ut[n][m];
        }
    }
    const ns3d_params c = to_c(P);
    if (Fface.size() < 5) Fface.resize(5);
    ok(ns3d_scalar_face_flux(&c, st, pt, dim, Fface.data()), "WCNS_Riemann_InviscidFlux");
}

// ------------------------------------------------------------ monitoring and post-processing
void compute_diagnostics(Field3D &F, const SolverParams &, const GridDesc &G) {
    const LocalDesc &L = F.L;
    const double dx = G.dx, dx3 = dx * dx * dx;  // dx only, as the reference
    F.updateResiduals();
    const std::vector<double> *U[5] = {&F.rho, &F.rhou, &F.rhov, &F.rhow, &F.E};
    const std::vector<double> *R[5] = {&F.res_rho, &F.res_rhou, &F.res_rhov, &F.res_rhow, &F.res_E};
    double sq[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
    for (int k = L.ngz; k < L.ngz + L.nz; ++k)
        for (int j = L.ngy; j < L.ngy + L.ny; ++j)
            for (int i = L.ngx; i < L.ngx + L.nx; ++i) {
                const int id = F.I(i, j, k);
                for (int n = 0; n < 5; ++n) {
                    mx[n] = std::max(mx[n], std::abs((*U[n])[id]));
                    const double r = (*R[n])[id];
                    sq[n] += r * r * dx3;
                }
            }
    double *out[5] = {&F.global_res_rho, &F.global_res_rhou, &F.global_res_rhov, &F.global_res_rhow, &F.global_res_E};
    for (int n = 0; n < 5; ++n) *out[n] = std::sqrt(sq[n]) / mx[n];
}

void write_residuals_tecplot(const Field3D &F, int step, const std::string &filename) {
    std::error_code ec;
    std::filesystem::create_directories("output", ec);
    if (ec) std::cerr << "Warning: could not create output directory 'output': " << ec.message() << "\n";
    const std::filesystem::path path = std::filesystem::path("output") / filename;
    std::ofstream ofs(path.string(), step == 1 ? std::ofstream::out : std::ofstream::out | std::ofstream::app);
    if (!ofs) {
        std::cerr << "Failed to open residuals file " << path.string() << "\n";
        return;
    }
    if (step == 1) {
        ofs << "TITLE = \"Residuals vs Step\"\n";
        ofs << "VARIABLES = \"Step\" \"Res_rho\" \"Res_rhou\" \"Res_rhov\" \"Res_rhow\" \"Res_E\" \"Etot\"\n";
    }
    ofs << std::scientific << std::setprecision(8);
    ofs << step << " " << F.global_res_rho << " " << F.global_res_rhou << " " << F.global_res_rhov << " "
        << F.global_res_rhow << " " << F.global_res_E << " " << F.global_Etot << "\n";
}

namespace {

// E(k), k = 0 .. 2 global_nx, on the device
bool spectrum(Field3D &F, const GridDesc &G, std::vector<double> &Ek) {
    ns3d_ctx *c = mirror(F, nullptr, &G);
    if (!c || !push(F, c, false, false)) return false;
    Ek.assign(2 * (size_t)G.global_nx + 1, 0.0);
    return ok(ns3d_energy_spectrum(c, Ek.data(), (int)Ek.size()), "compute_energy_spectrum");
}

std::string time_prefix(double t) {
    std::stringstream ss;
    ss << "output/output_time_" << std::fixed << std::setprecision(5) << t;
    return ss.str();
}

}  // namespace

void compute_energy_spectrum(const Field3D &Fc, const GridDesc &G, const CartDecomp &, const std::string &filename) {
    Field3D &F = const_cast<Field3D &>(Fc);  // the device mirror is a cache of the field
    std::vector<double> Ek;
    if (!spectrum(F, G, Ek)) return;
    std::ofstream fout(filename);
    fout << std::scientific << std::setprecision(16);
    for (size_t k = 0; k < Ek.size(); ++k) fout << k << " " << Ek[k] << "\n";
}

void compute_turbulence_statistics(Field3D &F, const GridDesc &G, const SolverParams &P, const CartDecomp &,
                                   const double current_time) {
    // kinetic energy, TKE, u_rms, nu = <mu(T)> / <rho> and the dissipation
    // 2 nu sum k^2 E(k) on the device; <(du/dx)^2> from compute_gradients
    // and the mean sound speed sqrt(gamma Rgas T) over the interior
    ns3d_ctx *c = mirror(F, &P, &G);
    if (!c || !push(F, c, false, false)) return;
    ns3d_turbulence st{};
    if (!ok(ns3d_turbulence_statistics(c, &st), "compute_turbulence_statistics") ||
        !ok(ns3d_compute_gradients(c), "compute_gradients"))
        return;
    const size_t n = cells(F.L);
    std::vector<double> dudx(n), T(n);
    if (!ok(ns3d_download_gradient(c, 0, dudx.data()), "gradient") || !get(c, NS3D_T, T)) return;
    const LocalDesc &L = F.L;
    double dudx2 = 0.0, sound = 0.0;
    for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < L.ny; ++j)
            for (int i = 0; i < L.nx; ++i) {
                const int id = F.I(i + L.ngx, j + L.ngy, k + L.ngz);
                dudx2 += dudx[id] * dudx[id];
                sound += std::sqrt(P.gamma * P.Rgas * T[id]);
            }
    const double N = (double)G.global_nx * G.global_ny * G.global_nz;
    const double nu = st.nu, eps = st.dissipation;
    const double taylor_Li = st.u_rms / std::sqrt(dudx2 / N);
    const double mach_t = std::sqrt(st.kinetic_energy * 2) / (sound / N);
    const double re_lambda = st.u_rms * taylor_Li / nu;
    const double eta = std::pow(nu * nu * nu / eps, 0.25), u_eta = std::pow(nu * eps, 0.25),
                 tau_eta = std::pow(nu / eps, 0.5);
    const double taylor = std::sqrt(15.0 * nu * st.u_rms * st.u_rms / eps);
    std::cout << "[Post] nu = " << nu << ", u_rms = " << st.u_rms << ", dudx = " << dudx2 / N
              << ", Taylor_Li = " << taylor_Li << ", Mach_t = " << mach_t << "\n";
    std::error_code ec;
    std::filesystem::create_directories("output", ec);
    const std::filesystem::path fpath("output/turbulence_stats.dat");
    const bool exists = std::filesystem::exists(fpath);
    std::ofstream fout(fpath, std::ios::app);
    if (!fout) {
        std::cout << "[Error] Cannot open output/turbulence_stats.dat for writing.\n";
        return;
    }
    if (!exists)
        fout << "current_time KineticEnergy TKE u_rms Dissipation Taylor Kol_scale Kol_vel Kol_time Taylor_Li "
                "Re_lambda Mach_t\n";
    fout << std::scientific << std::setprecision(8) << current_time << " " << st.kinetic_energy << " " << st.tke
         << " " << st.u_rms << " " << eps << " " << taylor << " " << eta << " " << u_eta << " " << tau_eta << " "
         << taylor_Li << " " << re_lambda << " " << mach_t << "\n";
}

void isotropic_post_process(Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &P,
                            const double current_time) {
    std::error_code ec;
    std::filesystem::create_directories("output", ec);
    compute_energy_spectrum(F, G, C, time_prefix(current_time) + "_spectrum.dat");
    compute_turbulence_statistics(F, G, P, C, current_time);
}
