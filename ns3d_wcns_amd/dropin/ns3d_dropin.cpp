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
#include <fstream>
#include <iostream>

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
Field3D::Field3D(const Field3D &o)
    : L(o.L), rho(o.rho), rhou(o.rhou), rhov(o.rhov), rhow(o.rhow), E(o.E), rho0(o.rho0), rhou0(o.rhou0),
      rhov0(o.rhov0), rhow0(o.rhow0), E0(o.E0), u(o.u), v(o.v), w(o.w), p(o.p), T(o.T), rhs_rho(o.rhs_rho),
      rhs_rhou(o.rhs_rhou), rhs_rhov(o.rhs_rhov), rhs_rhow(o.rhs_rhow), rhs_E(o.rhs_E), global_Etot(o.global_Etot),
      dev(nullptr) {}

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
    for (auto *f : {&rhou, &rhov, &rhow, &E, &rho0, &rhou0, &rhov0, &rhow0, &E0, &u, &v, &w, &p, &T, &rhs_rho,
                    &rhs_rhou, &rhs_rhov, &rhs_rhow, &rhs_E})
        f->assign(n, 0.0);
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

void apply_boundary_halfnode_flux(Field3D &, const GridDesc &, CartDecomp &, const SolverParams &) {
    // on a periodic box the faces computed by compute_invis_flux already are
    // the periodic images the reference exchanges
}

void exchange_halos_viscous_flux(Field3D &, CartDecomp &, LocalDesc &, HaloRequests &) {
    // likewise for the viscous fluxes
}

void compute_invis_flux(Field3D &F, const SolverParams &P) {
    ns3d_ctx *c = mirror(F, &P, nullptr);
    if (!c || !push(F, c, false, false) || !ok(ns3d_compute_invis_flux(c), "compute_invis_flux")) return;
    std::vector<double> *ff[3][5] = {
        {&F.flux_fx_mass, &F.flux_fx_momx, &F.flux_fx_momy, &F.flux_fx_momz, &F.flux_fx_E},
        {&F.flux_fy_mass, &F.flux_fy_momx, &F.flux_fy_momy, &F.flux_fy_momz, &F.flux_fy_E},
        {&F.flux_fz_mass, &F.flux_fz_momx, &F.flux_fz_momy, &F.flux_fz_momz, &F.flux_fz_E}};
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 5; ++n) {
            ff[d][n]->assign(ns3d_face_count(c, d), 0.0);
            if (!ok(ns3d_download_face_flux(c, d, n, ff[d][n]->data()), "face flux")) return;
        }
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

void compute_total_energy(Field3D &F, const GridDesc &G, CartDecomp &, const SolverParams &P) {
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
