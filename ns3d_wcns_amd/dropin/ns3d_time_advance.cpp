// ns3d_time_advance.cpp -- a time-advance program written against the
// reference's call surface (include/ns3d_dropin.h): solver.in, the initial
// primitives, then the time_advance loop of maincycle.cpp:7-113 (dt from
// compute_timestep or dt_fixed, clipped to TotalTime and output_Timefreq,
// stop on dt <= 1e-12; runge_kutta_3 per step; the kinetic-energy monitor)
// without file output.
//
//   ns3d_time_advance solver.in initial.bin final.bin [resident]
//   ns3d_time_advance solver.in --print-params    (parsed and derived values)
//
// initial.bin: int32 nx, ny, nz, then rho, u, v, w, p over the interior
// (nz, ny, nx), doubles.  final.bin: int32 steps, the dt of each step, the
// final conserved fields rho, rho u, rho v, rho w, E (ghosted, sz*sy*sx
// each) and the monitored kinetic energy after every step.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../include/ns3d_dropin.h"

namespace {

int time_advance(Field3D &F, CartDecomp &C, GridDesc &G, SolverParams &P, std::vector<double> &dts,
                 std::vector<double> &energy) {
    const double eps = 1e-12;
    double t = 0.0, out_t = P.output_Timefreq;
    HaloRequests reqs;
    for (int step = 1; step <= P.max_steps; ++step) {
        F.recordConservedTo0();
        if (std::abs(P.TotalTime - t) <= eps) break;
        double dt = P.dt_fixed;
        if (P.dt_fixed < 0.0) dt = compute_timestep(F, G, P);
        if (t >= out_t - eps) out_t += P.output_Timefreq;
        if (t + dt > P.TotalTime) dt = P.TotalTime - t;
        if (t + dt > out_t) {
            dt = out_t - t;
            out_t += P.output_Timefreq;
        }
        if (dt <= eps) {
            std::cerr << "dt too small or zero; aborting to avoid stall at Time=" << t << "\n";
            break;
        }
        t += dt;
        runge_kutta_3(F, C, G, P, reqs, dt);
        dts.push_back(dt);
        if (P.monitor_energy && (step % P.monitor_Stepfreq == 0 || step == 1)) {
            compute_total_energy(F, G, C, P);
            energy.push_back(F.global_Etot);
        }
        if (t >= P.TotalTime) break;
    }
    return (int)dts.size();
}

}  // namespace

int main(int argc, char **argv) {
    const bool print_only = argc == 3 && std::strcmp(argv[2], "--print-params") == 0;
    if (argc < 4 && !print_only) {
        std::cerr << "usage: " << argv[0] << " solver.in initial.bin final.bin [resident]\n"
                  << "       " << argv[0] << " solver.in --print-params\n";
        return 2;
    }
    SolverParams P;
    GridDesc G;
    CartDecomp C;
    if (!read_solver_params_from_file(argv[1], P, G, C)) return 1;
    if (print_only) {
        std::printf("gamma=%.17g\nPr=%.17g\nMa=%.17g\nRe=%.17g\ncfl=%.17g\ndt_fixed=%.17g\nmu=%.17g\n", P.gamma,
                    P.Pr, P.Ma, P.Re, P.cfl, P.dt_fixed, P.mu);
        std::printf("Cv=%.17g\nCp=%.17g\nRgas=%.17g\nmdcd_diss=%.17g\nmdcd_disp=%.17g\n", P.Cv, P.Cp, P.Rgas,
                    P.mdcd_diss, P.mdcd_disp);
        std::printf("interpolation=%d\nriemann_solver=%d\nchar_recon=%d\nghost_layers=%d\nstencil=%d\n",
                    (int)P.interpolation, (int)P.riemann_solver, (int)P.char_recon, P.ghost_layers, P.stencil);
        std::printf("max_steps=%d\nmonitor_Stepfreq=%d\noutput_Timefreq=%.17g\nTotalTime=%.17g\n", P.max_steps,
                    P.monitor_Stepfreq, P.output_Timefreq, P.TotalTime);
        std::printf("global_nx=%d\nglobal_ny=%d\nglobal_nz=%d\ndx=%.17g\ndy=%.17g\ndz=%.17g\n", G.global_nx,
                    G.global_ny, G.global_nz, G.dx, G.dy, G.dz);
        std::printf("periods=%d,%d,%d\n", C.periods[0], C.periods[1], C.periods[2]);
        return 0;
    }
    FILE *fi = std::fopen(argv[2], "rb");
    if (!fi) {
        std::cerr << "cannot open " << argv[2] << "\n";
        return 1;
    }
    int32_t dims[3];
    if (std::fread(dims, sizeof(int32_t), 3, fi) != 3 || dims[0] != G.global_nx || dims[1] != G.global_ny ||
        dims[2] != G.global_nz) {
        std::cerr << "initial state does not match the grid of " << argv[1] << "\n";
        return 1;
    }
    Field3D F;
    F.allocate(make_local_desc(G, P));
    const LocalDesc &L = F.L;
    const size_t nint = (size_t)L.nx * L.ny * L.nz;
    std::vector<double> buf(nint);
    std::vector<double> *dst[5] = {&F.rho, &F.u, &F.v, &F.w, &F.p};
    for (auto *d : dst) {
        if (std::fread(buf.data(), sizeof(double), nint, fi) != nint) {
            std::cerr << "short initial state\n";
            return 1;
        }
        size_t q = 0;
        for (int k = 0; k < L.nz; ++k)
            for (int j = 0; j < L.ny; ++j)
                for (int i = 0; i < L.nx; ++i) (*d)[F.I(i + L.ngx, j + L.ngy, k + L.ngz)] = buf[q++];
    }
    std::fclose(fi);
    // main.cpp:98-99: ghost layers, then the conserved fields
    apply_boundary(F, G, C, P);
    F.primitiveToConserved(P);
    const bool res = argc > 4 && std::strcmp(argv[4], "resident") == 0;
    if (res) ns3d_set_resident(F, true);
    std::vector<double> dts, energy;
    const int steps = time_advance(F, C, G, P, dts, energy);
    if (res) ns3d_sync_host(F);
    FILE *fo = std::fopen(argv[3], "wb");
    if (!fo) {
        std::cerr << "cannot write " << argv[3] << "\n";
        return 1;
    }
    const int32_t s = steps;
    std::fwrite(&s, sizeof(int32_t), 1, fo);
    std::fwrite(dts.data(), sizeof(double), dts.size(), fo);
    for (auto *f : {&F.rho, &F.rhou, &F.rhov, &F.rhow, &F.E}) std::fwrite(f->data(), sizeof(double), f->size(), fo);
    std::fwrite(energy.data(), sizeof(double), energy.size(), fo);
    std::fclose(fo);
    std::cout << steps << " steps, final kinetic energy " << (energy.empty() ? 0.0 : energy.back()) << "\n";
    return 0;
}
