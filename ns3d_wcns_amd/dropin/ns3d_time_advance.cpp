// ns3d_time_advance.cpp -- a time-advance program written against the
// reference's call surface (include/ns3d_dropin.h): solver.in, the initial
// state (primitives from a binary file, or field.h5 when solver.in sets
// restart = true; main.cpp:69-99), then the time_advance loop of
// maincycle.cpp:7-113 (dt from compute_timestep or dt_fixed, clipped to
// TotalTime and the output times, stop on dt <= 1e-12; runge_kutta_3 per
// step; the kinetic-energy monitor; field.h5 at the output times).
//
//   ns3d_time_advance solver.in initial.bin|riemann2d|sod|- final.bin [resident] [--write-fields]
//   ns3d_time_advance solver.in --print-params    (parsed and derived values)
//   ns3d_time_advance solver.in --convert initial.bin|in.h5 out.h5 [time]
//
// initial.bin: int32 nx, ny, nz, then rho, u, v, w, p over the interior
// (nz, ny, nx), doubles; "riemann2d" / "sod": the reference's built-in
// initialize_riemann_2d (its default run, main.cpp:85) / initialize_sod_shock_tube;
// "-" with restart = true.  Every boundary type of solver.in is honoured.  final.bin: int32 steps,
// the dt of each step, the final conserved fields rho, rho u, rho v, rho w, E
// (ghosted, sz*sy*sx each) and the monitored kinetic energy after every step.
// --write-fields writes output/time_<int(t * 1e5)>/field.h5 in the working
// directory at t = 0, at the output times (post_basicfield) and at TotalTime
// (main.cpp:104,111; maincycle.cpp:94-98).  --convert writes one field.h5 on
// the host (no device work): from a binary initial state (conserved fields by
// primitiveToConserved) or by reading and rewriting a field.h5.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../include/ns3d_dropin.h"

namespace {

int time_advance(Field3D &F, CartDecomp &C, GridDesc &G, SolverParams &P, bool write_fields,
                 std::vector<double> &dts, std::vector<double> &energy) {
    const double eps = 1e-12;
    double t = 0.0, out_t = P.output_Timefreq;
    HaloRequests reqs;
    for (int step = 1; step <= P.max_steps; ++step) {
        F.recordConservedTo0();
        if (std::abs(P.TotalTime - t) <= eps) break;
        double dt = P.dt_fixed;
        if (P.dt_fixed < 0.0) dt = compute_timestep(F, G, P);
        bool if_output = false;
        if (t >= out_t - eps) {
            if_output = true;
            out_t += P.output_Timefreq;
        }
        if (t + dt > P.TotalTime) dt = P.TotalTime - t;
        if (t + dt > out_t) {
            dt = out_t - t;
            out_t += P.output_Timefreq;
            if_output = true;
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
        if (if_output && write_fields && P.post_basicfield) write_tecplot_field(F, G, C, P, t);
        if (t >= P.TotalTime) break;
    }
    return (int)dts.size();
}

// interior primitives rho, u, v, w, p from an initial-state file
bool read_initial(const char *path, const GridDesc &G, Field3D &F) {
    FILE *fi = std::fopen(path, "rb");
    if (!fi) {
        std::cerr << "cannot open " << path << "\n";
        return false;
    }
    int32_t dims[3];
    if (std::fread(dims, sizeof(int32_t), 3, fi) != 3 || dims[0] != G.global_nx || dims[1] != G.global_ny ||
        dims[2] != G.global_nz) {
        std::cerr << "initial state does not match the grid\n";
        std::fclose(fi);
        return false;
    }
    const LocalDesc &L = F.L;
    const size_t nint = (size_t)L.nx * L.ny * L.nz;
    std::vector<double> buf(nint);
    std::vector<double> *dst[5] = {&F.rho, &F.u, &F.v, &F.w, &F.p};
    for (auto *d : dst) {
        if (std::fread(buf.data(), sizeof(double), nint, fi) != nint) {
            std::cerr << "short initial state\n";
            std::fclose(fi);
            return false;
        }
        size_t q = 0;
        for (int k = 0; k < L.nz; ++k)
            for (int j = 0; j < L.ny; ++j)
                for (int i = 0; i < L.nx; ++i) (*d)[F.I(i + L.ngx, j + L.ngy, k + L.ngz)] = buf[q++];
    }
    std::fclose(fi);
    return true;
}

bool ends_with(const std::string &s, const std::string &x) {
    return s.size() >= x.size() && s.compare(s.size() - x.size(), x.size(), x) == 0;
}

}  // namespace

int main(int argc, char **argv) {
    const bool print_only = argc == 3 && std::strcmp(argv[2], "--print-params") == 0;
    if (argc < 4 && !print_only) {
        std::cerr << "usage: " << argv[0] << " solver.in initial.bin|riemann2d|sod|- final.bin [resident] [--write-fields]\n"
                  << "       " << argv[0] << " solver.in --print-params\n"
                  << "       " << argv[0] << " solver.in --convert initial.bin|in.h5 out.h5 [time]\n";
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
        std::printf("restart=%d\nrestart_file=%s\npost_basicfield=%d\nmonitor_energy=%d\n", (int)P.restart,
                    P.restart_file.c_str(), (int)P.post_basicfield, (int)P.monitor_energy);
        return 0;
    }
    Field3D F;
    F.allocate(make_local_desc(G, P));
    if (std::strcmp(argv[2], "--convert") == 0) {
        // host only: no hot-path function is called
        if (argc < 5) {
            std::cerr << "--convert needs an input and an output file\n";
            return 2;
        }
        const std::string in = argv[3];
        if (ends_with(in, ".h5")) {
            if (!initialize_from_hdf5(F, G, C, P, in)) return 1;
        } else {
            if (!read_initial(argv[3], G, F)) return 1;
            F.primitiveToConserved(P);
        }
        return write_hdf5_field(F, G, C, argv[4], argc > 5 ? std::atof(argv[5]) : 0.0) ? 0 : 1;
    }
    bool res = false, write_fields = false;
    for (int a = 4; a < argc; ++a) {
        if (std::strcmp(argv[a], "resident") == 0) res = true;
        else if (std::strcmp(argv[a], "--write-fields") == 0) write_fields = true;
        else {
            std::cerr << "unknown option " << argv[a] << "\n";
            return 2;
        }
    }
    if (P.restart) {
        // main.cpp:69-84: the state comes from the restart file
        if (!initialize_from_hdf5(F, G, C, P, P.restart_file)) {
            std::cerr << "Restart initialization from HDF5 failed, file='" << P.restart_file << "'\n";
            return 1;
        }
    } else if (std::strcmp(argv[2], "riemann2d") == 0) {
        initialize_riemann_2d(F, G, P);
    } else if (std::strcmp(argv[2], "sod") == 0) {
        initialize_sod_shock_tube(F, G, P);
    } else if (!read_initial(argv[2], G, F)) {
        return 1;
    }
    // main.cpp:98-99: ghost layers, then the conserved fields
    apply_boundary(F, G, C, P);
    F.primitiveToConserved(P);
    if (res) ns3d_set_resident(F, true);
    if (write_fields) write_tecplot_field(F, G, C, P, 0.0);
    std::vector<double> dts, energy;
    const int steps = time_advance(F, C, G, P, write_fields, dts, energy);
    if (write_fields) write_tecplot_field(F, G, C, P, P.TotalTime);
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
