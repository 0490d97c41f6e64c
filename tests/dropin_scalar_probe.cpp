// dropin_scalar_probe.cpp -- calls the drop-in's scalar functions with the
// reference's signatures (include/ns3d_dropin.h) on cases read from stdin and
// prints the results, one value per line at 17 significant digits, for
// tests/test_dropin_scalar.py.
//
// stdin, one case per line:
//   interp <interp> <flag> <diss> <disp> <n> s0 .. s(n-1)     interpolate_select
//   named <which> <diss> <disp> s0 .. s5                      weno5_interpolate (0), mdcd_linear (1),
//                                                             mdcd_hybrid (2), weno5_reconstruction (3),
//                                                             zero_interpolate (4, s0 s1 only used)
//   riemann <rs> <nx> <ny> <nz> <gamma> UL0..UL4 UR0..UR4      X_Riemann_solver
//   face <interp> <rs> <char> <dim> <gamma> Ut[5][6] ut[5][6]   WCNS_Riemann_InviscidFlux
//   diag <nx> <ny> <nz> <ng> <dx> <step> <seed>                 compute_diagnostics + write_residuals_tecplot
//   recon <kind> <recon> <flag> <diss> <disp> <n> s0 .. s(n-1)  reconstruct_select (kind 0, P.recon = recon),
//                                                             linear_ (1), mdcd_ (2), c6th_ (3), c4th_ (4)
//                                                             reconstruction on s0 ..
//   post <n> <time>                                             isotropic_post_process of the analytic field
//                                                               post_field() on a periodic n^3 box of side
//                                                               2 pi (device; files under ./output)
// An exception prints "throw <what>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../include/ns3d_dropin.h"

namespace {

void put(double v) { std::printf("%.17g\n", v); }

using RS = SolverParams::RiemannSolver;
using IT = SolverParams::Interpolation;

void riemann(int rs, std::vector<double> &F, const std::vector<double> &L, const std::vector<double> &R, double nx,
             double ny, double nz, double g) {
    switch (rs) {
        case 0: Roe_Riemann_solver(F, L, R, nx, ny, nz, g); break;
        case 1: Rusanov_Riemann_solver(F, L, R, nx, ny, nz, g); break;
        case 2: HLLC_Riemann_solver(F, L, R, nx, ny, nz, g); break;
        case 3: HLL_Riemann_solver(F, L, R, nx, ny, nz, g); break;
        case 4: HLLC_p_Riemann_solver(F, L, R, nx, ny, nz, g); break;
        default: AUSM_Riemann_solver(F, L, R, nx, ny, nz, g); break;
    }
}

// a deterministic pseudo-random value in [0, 1)
double lcg(unsigned long long &s) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) / 9007199254740992.0;
}

// rho, u, v, w, p at (x, y, z); p0 = 1 / (gamma Ma^2)
void post_field(double x, double y, double z, double p0, double *q) {
    q[0] = 1.0 + 0.1 * std::sin(x) * std::cos(y);
    q[1] = std::sin(x) * std::cos(y) * std::cos(z) + 0.1 * std::sin(2.0 * z);
    q[2] = -std::cos(x) * std::sin(y) * std::cos(z);
    q[3] = 0.05 * std::cos(3.0 * x);
    q[4] = p0 + 0.1 * std::cos(x + y);
}

}  // namespace

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string kind;
        in >> kind;
        try {
            if (kind == "interp") {
                int it, n;
                double flag;
                SolverParams P;
                in >> it >> flag >> P.mdcd_diss >> P.mdcd_disp >> n;
                P.interpolation = static_cast<IT>(it);
                std::vector<double> s(n);
                for (auto &x : s) in >> x;
                put(interpolate_select(s, flag, P));
            } else if (kind == "named") {
                int which;
                SolverParams P;
                std::array<double, 6> s{};
                in >> which >> P.mdcd_diss >> P.mdcd_disp;
                for (auto &x : s) in >> x;
                if (which == 0) put(weno5_interpolate(s));
                else if (which == 1) put(mdcd_linear_interpolate(s, P));
                else if (which == 2) put(mdcd_hybrid_interpolate(s, P));
                else if (which == 3) put(weno5_reconstruction(s));
                else put(zero_interpolate({s[0], s[1]}));
            } else if (kind == "riemann") {
                int rs;
                double nx, ny, nz, g;
                std::vector<double> L(5), R(5), F(5, 0.0);
                in >> rs >> nx >> ny >> nz >> g;
                for (auto &x : L) in >> x;
                for (auto &x : R) in >> x;
                riemann(rs, F, L, R, nx, ny, nz, g);
                for (double f : F) put(f);
            } else if (kind == "face") {
                int it, rs, chr, dim;
                SolverParams P;
                in >> it >> rs >> chr >> dim >> P.gamma;
                P.interpolation = static_cast<IT>(it);
                P.riemann_solver = static_cast<RS>(rs);
                P.char_recon = chr != 0;
                std::vector<std::vector<double>> Ut(5, std::vector<double>(6)), ut(5, std::vector<double>(6));
                for (auto &row : Ut)
                    for (auto &x : row) in >> x;
                for (auto &row : ut)
                    for (auto &x : row) in >> x;
                std::vector<double> F(5, 0.0);
                WCNS_Riemann_InviscidFlux(F, Ut, ut, P, dim);
                for (double f : F) put(f);
            } else if (kind == "recon") {
                int which, rc, n;
                double flag;
                SolverParams P;
                in >> which >> rc >> flag >> P.mdcd_diss >> P.mdcd_disp >> n;
                P.recon = static_cast<SolverParams::Reconstruction>(rc);
                std::vector<double> v(n);
                for (auto &x : v) in >> x;
                v.resize(std::max(n, 6), 0.0);
                if (which == 0) put(reconstruct_select(std::vector<double>(v.begin(), v.begin() + n), flag, P));
                else if (which == 1) put(linear_reconstruction({v[0], v[1]}));
                else if (which == 2) put(mdcd_reconstruction({v[0], v[1], v[2], v[3], v[4], v[5]}, P));
                else if (which == 3) put(c6th_reconstruction({v[0], v[1], v[2], v[3], v[4], v[5]}));
                else put(c4th_reconstruction({v[0], v[1], v[2], v[3]}));
            } else if (kind == "post") {
                int n;
                double t;
                in >> n >> t;
                GridDesc G;
                G.global_nx = G.global_ny = G.global_nz = n;
                G.Lx = G.Ly = G.Lz = 2.0 * M_PI;
                G.dx = G.dy = G.dz = 2.0 * M_PI / n;
                SolverParams P;
                CartDecomp C;
                Field3D F;
                F.allocate(make_local_desc(G, P));
                const LocalDesc &L = F.L;
                const double p0 = 1.0 / (P.gamma * P.Ma * P.Ma);
                for (int k = 0; k < n; ++k)
                    for (int j = 0; j < n; ++j)
                        for (int i = 0; i < n; ++i) {
                            double q[5];
                            post_field((i + 0.5) * G.dx, (j + 0.5) * G.dy, (k + 0.5) * G.dz, p0, q);
                            const int id = F.I(i + L.ngx, j + L.ngy, k + L.ngz);
                            F.rho[id] = q[0];
                            F.u[id] = q[1];
                            F.v[id] = q[2];
                            F.w[id] = q[3];
                            F.p[id] = q[4];
                        }
                apply_boundary(F, G, C, P);
                F.primitiveToConserved(P);
                isotropic_post_process(F, G, C, P, t);
                put(F.T[F.I(L.ngx, L.ngy, L.ngz)]);
            } else if (kind == "diag") {
                int nx, ny, nz, ng, step;
                double dx;
                unsigned long long seed;
                in >> nx >> ny >> nz >> ng >> dx >> step >> seed;
                LocalDesc L;
                L.nx = nx;
                L.ny = ny;
                L.nz = nz;
                L.ngx = L.ngy = L.ngz = ng;
                Field3D F;
                F.allocate(L);
                std::vector<double> *U[5] = {&F.rho, &F.rhou, &F.rhov, &F.rhow, &F.E};
                std::vector<double> *U0[5] = {&F.rho0, &F.rhou0, &F.rhov0, &F.rhow0, &F.E0};
                for (int q = 0; q < 5; ++q)
                    for (size_t i = 0; i < U[q]->size(); ++i) {
                        (*U0[q])[i] = (q == 0 || q == 4 ? 1.0 : 0.0) + lcg(seed) - 0.5;
                        (*U[q])[i] = (*U0[q])[i] + 1e-3 * (lcg(seed) - 0.5);
                    }
                GridDesc G;
                G.dx = dx;
                SolverParams P;
                compute_diagnostics(F, P, G);
                F.global_Etot = 0.25;
                write_residuals_tecplot(F, step, "residuals_probe.dat");
                put(F.global_res_rho);
                put(F.global_res_rhou);
                put(F.global_res_rhov);
                put(F.global_res_rhow);
                put(F.global_res_E);
                put((double)F.res_rho.size());
                put(F.res_E[0] - (F.E[0] - F.E0[0]));
            }
        } catch (const std::exception &e) {
            std::printf("throw %s\n", e.what());
        }
        std::fflush(stdout);
    }
    return 0;
}
