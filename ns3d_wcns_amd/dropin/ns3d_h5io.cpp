// ns3d_h5io.cpp -- field.h5 output and restart of the drop-in
// (write_hdf5_field / write_tecplot_field, output.cpp:19-203;
// initialize_from_hdf5, initialization.cpp:270-420), serial HDF5 on the one
// rank.
//
// File layout (the reference's): scalar attributes global_nx, global_ny,
// global_nz (native int) and time, dx, dy, dz (native double) on the root
// group; datasets X, Y, Z, rho, u, v, w, E, p, T, each IEEE F64LE of shape
// (global_nz, global_ny, global_nx) over the interior, x fastest, with
// X = x0 + i dx (likewise Y, Z).
//
// The HDF5 C library is opened with dlopen on first use, so the rest of the
// drop-in neither links nor needs it; its ABI (HDF5 >= 1.10: 64-bit hid_t)
// comes from hdf5.h.  Only functions and the three datatype globals listed in
// H5Api are used.
#include <dlfcn.h>
#include <hdf5.h>

#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ns3d_dropin.h"

namespace {

// H5F_ACC_TRUNC / H5F_ACC_RDONLY without the library calls their macros
// carry (H5Fpublic.h: 0x0002u, 0x0000u)
constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;

struct H5Api {
    bool ok = false;
    std::string why;
    decltype(&H5open) open = nullptr;
    decltype(&H5Fcreate) fcreate = nullptr;
    decltype(&H5Fopen) fopen = nullptr;
    decltype(&H5Fclose) fclose = nullptr;
    decltype(&H5Screate) screate = nullptr;
    decltype(&H5Screate_simple) screate_simple = nullptr;
    decltype(&H5Sselect_hyperslab) sselect_hyperslab = nullptr;
    decltype(&H5Sclose) sclose = nullptr;
    decltype(&H5Acreate2) acreate = nullptr;
    decltype(&H5Awrite) awrite = nullptr;
    decltype(&H5Aclose) aclose = nullptr;
    decltype(&H5Dcreate2) dcreate = nullptr;
    decltype(&H5Dopen2) dopen = nullptr;
    decltype(&H5Dget_space) dget_space = nullptr;
    decltype(&H5Dwrite) dwrite = nullptr;
    decltype(&H5Dread) dread = nullptr;
    decltype(&H5Dclose) dclose = nullptr;
    hid_t native_int = -1, native_double = -1, ieee_f64le = -1;
};

template <class F>
bool sym(void *h, const char *name, F &f, std::string &why) {
    f = reinterpret_cast<F>(dlsym(h, name));
    if (!f) why = std::string("symbol ") + name + " missing";
    return f != nullptr;
}

const H5Api &h5() {
    static H5Api api;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> names;
        if (const char *e = std::getenv("NS3D_HDF5_LIBRARY")) names.push_back(e);
        names.push_back("libhdf5.so.103");
        names.push_back("/opt/conda/lib/libhdf5.so.103");
        names.push_back("libhdf5.so");
        void *h = nullptr;
        for (const auto &n : names)
            if ((h = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
        if (!h) {
            api.why = "no HDF5 library found (set NS3D_HDF5_LIBRARY)";
            return;
        }
        std::string &w = api.why;
        const bool all =
            sym(h, "H5open", api.open, w) && sym(h, "H5Fcreate", api.fcreate, w) && sym(h, "H5Fopen", api.fopen, w) &&
            sym(h, "H5Fclose", api.fclose, w) && sym(h, "H5Screate", api.screate, w) &&
            sym(h, "H5Screate_simple", api.screate_simple, w) &&
            sym(h, "H5Sselect_hyperslab", api.sselect_hyperslab, w) && sym(h, "H5Sclose", api.sclose, w) &&
            sym(h, "H5Acreate2", api.acreate, w) && sym(h, "H5Awrite", api.awrite, w) &&
            sym(h, "H5Aclose", api.aclose, w) && sym(h, "H5Dcreate2", api.dcreate, w) &&
            sym(h, "H5Dopen2", api.dopen, w) && sym(h, "H5Dget_space", api.dget_space, w) &&
            sym(h, "H5Dwrite", api.dwrite, w) && sym(h, "H5Dread", api.dread, w) && sym(h, "H5Dclose", api.dclose, w);
        if (!all) return;
        if (api.open() < 0) {
            w = "H5open failed";
            return;
        }
        // the predefined datatype ids are globals, valid after H5open
        const hid_t *ni = static_cast<const hid_t *>(dlsym(h, "H5T_NATIVE_INT_g"));
        const hid_t *nd = static_cast<const hid_t *>(dlsym(h, "H5T_NATIVE_DOUBLE_g"));
        const hid_t *fl = static_cast<const hid_t *>(dlsym(h, "H5T_IEEE_F64LE_g"));
        if (!ni || !nd || !fl) {
            w = "HDF5 datatype globals missing";
            return;
        }
        api.native_int = *ni;
        api.native_double = *nd;
        api.ieee_f64le = *fl;
        api.ok = true;
    });
    return api;
}

size_t interior_cells(const LocalDesc &L) { return (size_t)L.nx * L.ny * L.nz; }

// interior linear index (k, j, i) -> (k * ny + j) * nx + i
template <class Fn>
void for_interior(const Field3D &F, Fn fn) {
    const LocalDesc &L = F.L;
    size_t lid = 0;
    for (int k = L.ngz; k < L.ngz + L.nz; ++k)
        for (int j = L.ngy; j < L.ngy + L.ny; ++j)
            for (int i = L.ngx; i < L.ngx + L.nx; ++i, ++lid) fn(lid, F.I(i, j, k), i - L.ngx, j - L.ngy, k - L.ngz);
}

}  // namespace

bool write_hdf5_field(const Field3D &Fc, const GridDesc &G, const CartDecomp &C, const std::string &filepath,
                      double time) {
    const H5Api &H = h5();
    if (!H.ok) {
        if (C.rank == 0) std::cerr << "HDF5 output unavailable: " << H.why << "\n";
        return false;
    }
    // a device-resident state is brought to the host vectors first (they are
    // its cache; the state itself does not change)
    if (ns3d_is_resident(Fc)) ns3d_sync_host(const_cast<Field3D &>(Fc));
    const Field3D &F = Fc;
    const LocalDesc &L = F.L;
    const size_t n = interior_cells(L);
    std::vector<double> X(n), Y(n), Z(n), rho(n), u(n), v(n), w(n), E(n), p(n), T(n);
    for_interior(F, [&](size_t lid, int gid, int i, int j, int k) {
        X[lid] = G.x0 + (L.ox + i) * G.dx;
        Y[lid] = G.y0 + (L.oy + j) * G.dy;
        Z[lid] = G.z0 + (L.oz + k) * G.dz;
        rho[lid] = F.rho[gid];
        u[lid] = F.u[gid];
        v[lid] = F.v[gid];
        w[lid] = F.w[gid];
        E[lid] = F.E[gid];
        p[lid] = F.p[gid];
        T[lid] = F.T[gid];
    });
    const hsize_t gdims[3] = {(hsize_t)G.global_nz, (hsize_t)G.global_ny, (hsize_t)G.global_nx};
    const hsize_t ldims[3] = {(hsize_t)L.nz, (hsize_t)L.ny, (hsize_t)L.nx};
    const hsize_t start[3] = {(hsize_t)L.oz, (hsize_t)L.oy, (hsize_t)L.ox};
    const hid_t file = H.fcreate(filepath.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) {
        std::cerr << "Failed to create HDF5 file: " << filepath << "\n";
        return false;
    }
    bool ok = true;
    auto attr = [&](const char *name, hid_t type, const void *value) {
        const hid_t sp = H.screate(H5S_SCALAR);
        const hid_t a = sp < 0 ? -1 : H.acreate(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
        ok = ok && a >= 0 && H.awrite(a, type, value) >= 0;
        if (a >= 0) H.aclose(a);
        if (sp >= 0) H.sclose(sp);
    };
    attr("global_nx", H.native_int, &G.global_nx);
    attr("global_ny", H.native_int, &G.global_ny);
    attr("global_nz", H.native_int, &G.global_nz);
    attr("time", H.native_double, &time);
    attr("dx", H.native_double, &G.dx);
    attr("dy", H.native_double, &G.dy);
    attr("dz", H.native_double, &G.dz);
    const hid_t filespace = H.screate_simple(3, gdims, nullptr);
    const hid_t memspace = H.screate_simple(3, ldims, nullptr);
    if (filespace < 0 || memspace < 0 ||
        H.sselect_hyperslab(filespace, H5S_SELECT_SET, start, nullptr, ldims, nullptr) < 0) {
        std::cerr << "Failed to select hyperslab for HDF5 write\n";
        ok = false;
    }
    auto dataset = [&](const char *name, const std::vector<double> &data) {
        if (!ok) return;
        const hid_t d = H.dcreate(file, name, H.ieee_f64le, filespace, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        ok = d >= 0 && H.dwrite(d, H.native_double, memspace, filespace, H5P_DEFAULT, data.data()) >= 0;
        if (d >= 0) H.dclose(d);
    };
    dataset("X", X);
    dataset("Y", Y);
    dataset("Z", Z);
    dataset("rho", rho);
    dataset("u", u);
    dataset("v", v);
    dataset("w", w);
    dataset("E", E);
    dataset("p", p);
    dataset("T", T);
    if (memspace >= 0) H.sclose(memspace);
    if (filespace >= 0) H.sclose(filespace);
    H.fclose(file);
    if (!ok) std::cerr << "Failed to write one or more datasets to HDF5 file: " << filepath << "\n";
    return ok;
}

void write_tecplot_field(const Field3D &F, const GridDesc &G, const CartDecomp &C, const SolverParams &,
                         double time) {
    // directory tag int(time * 1e5), as the reference names it; a time whose
    // tag does not fit an int has no such name and is refused
    const double tag = time * 100000;
    if (!(std::fabs(tag) < 2147483647.0)) {
        std::cerr << "ns3d: write_tecplot_field: time " << time << " has no output directory tag\n";
        return;
    }
    const std::filesystem::path dir = std::filesystem::path("output") / ("time_" + std::to_string((int)tag));
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) std::cerr << "Warning: could not create output directory 'output': " << ec.message() << "\n";
    const std::filesystem::path h5path = dir / "field.h5";
    if (write_hdf5_field(F, G, C, h5path.string(), time) && C.rank == 0)
        std::cerr << "Wrote HDF5 file: " << h5path.string() << " (time=" << std::scientific << std::setprecision(8)
                  << time << ")\n";
}

bool initialize_from_hdf5(Field3D &F, const GridDesc &, const CartDecomp &C, const SolverParams &,
                          const std::string &filename) {
    const H5Api &H = h5();
    if (!H.ok) {
        if (C.rank == 0) std::cerr << "HDF5 restart unavailable: " << H.why << "\n";
        return false;
    }
    const LocalDesc &L = F.L;
    const size_t n = interior_cells(L);
    const hsize_t ldims[3] = {(hsize_t)L.nz, (hsize_t)L.ny, (hsize_t)L.nx};
    const hsize_t start[3] = {(hsize_t)L.oz, (hsize_t)L.oy, (hsize_t)L.ox};
    std::vector<double> rho(n), u(n), v(n), w(n), E(n), p(n), T(n);
    const hid_t file = H.fopen(filename.c_str(), kAccRdonly, H5P_DEFAULT);
    if (file < 0) {
        std::cerr << "Failed to open HDF5 file: " << filename << "\n";
        return false;
    }
    const hid_t memspace = H.screate_simple(3, ldims, nullptr);
    if (memspace < 0) {
        std::cerr << "Failed to create HDF5 memory dataspace\n";
        H.fclose(file);
        return false;
    }
    auto read = [&](const char *name, std::vector<double> &data) -> bool {
        const hid_t d = H.dopen(file, name, H5P_DEFAULT);
        if (d < 0) {
            std::cerr << "Failed to open dataset '" << name << "' in " << filename << "\n";
            return false;
        }
        const hid_t fs = H.dget_space(d);
        if (fs < 0) {
            H.dclose(d);
            std::cerr << "Failed to get filespace for dataset '" << name << "'\n";
            return false;
        }
        if (H.sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, ldims, nullptr) < 0) {
            H.sclose(fs);
            H.dclose(d);
            std::cerr << "Failed to select hyperslab for dataset '" << name << "'\n";
            return false;
        }
        const herr_t rc = H.dread(d, H.native_double, memspace, fs, H5P_DEFAULT, data.data());
        H.sclose(fs);
        H.dclose(d);
        if (rc < 0) std::cerr << "Failed to read dataset '" << name << "'\n";
        return rc >= 0;
    };
    const bool ok = read("rho", rho) && read("u", u) && read("v", v) && read("w", w) && read("E", E) &&
                    read("p", p) && read("T", T);
    H.sclose(memspace);
    H.fclose(file);
    if (!ok) return false;
    for_interior(F, [&](size_t lid, int id, int, int, int) {
        F.rho[id] = rho[lid];
        F.rhou[id] = rho[lid] * u[lid];
        F.rhov[id] = rho[lid] * v[lid];
        F.rhow[id] = rho[lid] * w[lid];
        F.E[id] = E[lid];
        F.u[id] = u[lid];
        F.v[id] = v[lid];
        F.w[id] = w[lid];
        F.p[id] = p[lid];
        F.T[id] = T[lid];
    });
    if (ns3d_is_resident(F)) ns3d_sync_device(F);
    return true;
}
