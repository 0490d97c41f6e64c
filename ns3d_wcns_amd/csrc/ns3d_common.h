// ns3d_common.h -- kernel argument blocks and launch helpers shared by the
// HIP translation units of libns3d_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ns3d_device.h"

namespace ns3d {

constexpr int NG = 3;  // ghost layers (initialization.cpp:175-191)

// Geometry of one periodic box: interior n*, ghosted s* = n* + 2*NG.
struct Box {
    int nx, ny, nz;
    int sx, sy, sz;
    __host__ __device__ long gidx(int i, int j, int k) const {  // interior coords -> ghosted index
        return ((long)(k + NG) * sy + (j + NG)) * sx + (i + NG);
    }
    __host__ __device__ long oidx(int i, int j, int k) const {  // interior coords -> compact index
        return ((long)k * ny + j) * nx + i;
    }
};

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
// threadIdx.x of the calling wave's first lane (1-D blocks of whole waves,
// all lanes active): a scalar value, so that indices derived from it are
// formed on the scalar unit instead of once per lane
__device__ __forceinline__ int wave_first_thread() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63; }

// |u| + a with the reference's x-velocity formula (compute_timestep,
// timestep.cpp:12-20).  Evaluated without contraction so that
// dt = cfl * min(1e9, dx / smax) reproduces the reference's bits (correctly
// rounded division is monotone, so min_i dx/s_i = dx / max_i s_i).
__device__ __forceinline__ double cfl_speed(double rho, double rhou, double E, double gamma) {
#pragma clang fp contract(off)
    const double u = rhou / rho;
    const double p = (gamma - 1.) * (E - 0.5 * rho * u * u);
    const double a = sqrt(gamma * p / rho);
    return fabs(u) + a;
}

// Reference state subtracted before the FP32 arithmetic of the mixed method
// (zero for the plain method): E' = E - E_ref, and the pressure in the
// momentum flux is carried as p - p_ref.
struct MixedRef {
    double E_ref, p_ref;
};

// Face-flux kernel arguments.  U: ghosted conserved fields; Q: ghosted u, v,
// w, p (read for characteristic-wise interpolation); F: compact face fluxes
// of this direction, face c+1/2 stored at the interior cell c.
template <class R>
struct FaceArgs {
    const double *U[5];
    const double *Q[4];
    double *F[5];
    Box b;
    SchemeCoef<R> c;
    MixedRef ref;
};

// Compact slots of the nine distinct viscous-flux components.  The reference
// stores twelve (Fvflux_* = x, Hvflux_* = y, Gvflux_* = z for momx, momy, momz,
// E; ViscousFlux.cpp:187-206); the off-diagonal stresses appear twice there.
enum { VXX = 0, VYY, VZZ, VXY, VXZ, VYZ, VEX, VEY, VEZ, NVISC };

// Shortest marching segment of the forward pencils (the launch picks the
// segment count from the kernel's occupancy, march_seg); the adjoint's
// marches, with fewer cells of lead-in, go down to ADJ_SEG_MIN.
constexpr int PENCIL_SEG_MIN = 32;
constexpr int ADJ_SEG_MIN = 16;
// Shortest segment of a box too small to give every CU a block with
// segments of min_len cells.
constexpr int SEG_FINE = 8;

// Segments per column of a marching launch: the blocks the device holds at
// once (blocks per CU at the kernel's occupancy x CUs) over the blocks of one
// segment layer, rounded, so that the launch fills the resident slots a whole
// number of times with the fewest segments (each segment repeats its
// stencil's lead-in); segments of at least min_len cells, or of at least
// SEG_FINE cells as far as a box needs them to give every CU one block.
// E.g. 256^3: a layer is 4 x 64 blocks of 256 threads; two blocks per CU
// hold 512 -> 2 segments of 128 (three per CU: 3 of 86).  128^3: 4 segments
// of 32 (one block per CU).  64^3: a layer is 16 blocks -> 8 segments of 8.
inline int march_segments(int blocks_per_cu, long layer_blocks, int N, int min_len) {
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            n = 0;
        return n;
    }();
    const int coarse = N / min_len > 1 ? N / min_len : 1;
    if (cus <= 0 || blocks_per_cu <= 0 || layer_blocks <= 0) return (N + 63) / 64 < coarse ? (N + 63) / 64 : coarse;
    const long fine = N / SEG_FINE > 1 ? N / SEG_FINE : 1;
    const long fill = (cus + layer_blocks - 1) / layer_blocks;  // segments for one block per CU
    const long most = std::min(fine, std::max<long>(coarse, fill));
    long S = ((long)blocks_per_cu * cus + layer_blocks / 2) / layer_blocks;
    if (S < 1) S = 1;
    return (int)std::min(S, most);
}

// Cells per segment for kernel K (256-thread blocks) marching columns of
// length N, layer_blocks blocks per segment layer; K's occupancy is queried
// once.
template <auto K>
inline int march_seg(long layer_blocks, int N, int min_len = PENCIL_SEG_MIN) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, K, 256, 0) != hipSuccess) n = 0;
        return n;
    }();
    const int S = march_segments(per_cu, layer_blocks, N, min_len);
    return (N + S - 1) / S;
}

// The recorded state U0 of stage 1 (U1 = U0 + dt RHS): U0_READ loads the
// U0 arrays; U0_RECORD takes the stage's own state and also stores it as U0
// for the interior (recordConservedTo0 fused into the stage); U0_INPUT takes
// the stage's own state and stores nothing.  Stages 2 and 3 always load U0.
enum { U0_READ = 0, U0_RECORD = 1, U0_INPUT = 2 };

// Arguments of the fused pencil kernels of one stage.
struct PencilArgs {
    const double *U[5];       // ghosted current state
    double *RHS[5];           // ghosted RHS (stage 0 output)
    double *Rz[5], *Ry[5];    // compact partial RHS of the z and y pencils
    const double *V[NVISC];   // compact viscous fluxes
    double *U0[5];            // ghosted recorded state (written when record is set)
    double *Unew[5];          // ghosted state written by the stage update (stage 1..3)
    double *Q[5];             // ghosted u, v, w, p, T (written at stage 1..3)
    Box b;
    MixedRef ref;
    double ihd[3];            // 1/dx, 1/dy, 1/dz
    double ih, i60;           // this direction: 1/h, 1/(60 h)
    double gamma, Rgas;
    const double *dt;         // device scalar, or nullptr to use dtv
    double dtv;
    int stage;                // 0: store the RHS; 1..3: RK3 stage
    int record;               // stage 1: where U0 comes from (U0_READ / U0_RECORD / U0_INPUT)
    int seg;                  // y / z segment length
    int line_threads;         // x pencil: threads per line
    float *Q32[4];            // mixed, stages 1-2: compact FP32 u, v, w, T - 1 written instead of Q (else null)
    unsigned long long *cfl_slots;  // stage 3: CFL speed maxima of the new state (atomic max of the bit
                                    // patterns into CFL_SLOTS slots), for the next step's dt; else null
};

// Slots of the CFL maxima that the stage-3 x pencils fold into (each wave
// takes slot (its index in the launch) mod CFL_SLOTS; non-negative doubles
// order as their bit patterns do).
constexpr int CFL_SLOTS = 2048;

template <int RS>
void launch_pencils(int dir, int interp, bool chr, int precision, const PencilArgs &a, const SchemeCoef<double> &c64,
                    const SchemeCoef<float> &c32, hipStream_t s);

template <int RS>
void launch_face(int dir, int interp, bool chr, int precision, const FaceArgs<double> &a64,
                 const FaceArgs<float> &a32, hipStream_t s);

// 256-thread blocks of 64 x 4 interior cells of one z plane.
inline dim3 cell_block() { return dim3(64, 4, 1); }
inline dim3 cell_grid(const Box &b) { return dim3((b.nx + 63) / 64, (b.ny + 3) / 4, b.nz); }

// XCD-aware tile order.  Workgroup ids are dispatched round-robin over the
// eight XCDs, each with its own L2, so neighbouring ids land on different L2s.
// The remap gives XCD x the contiguous tile range [x*per, (x+1)*per) and walks
// the tiles with axis FAST (1 = y, 2 = z) varying fastest: one XCD then sweeps
// neighbouring tiles along the stencil axis while their data are in its L2.
enum { TILE_Y_FAST = 1, TILE_Z_FAST = 2 };

inline long tile_count(const Box &b) { return (long)((b.nx + 63) / 64) * ((b.ny + 3) / 4) * b.nz; }
inline dim3 tile_grid(const Box &b) { return dim3((unsigned)(((tile_count(b) + 7) / 8) * 8)); }

template <int FAST>
__device__ __forceinline__ bool tile_cell(const Box &b, int &i, int &j, int &k) {
    const int ntx = (b.nx + 63) >> 6, nty = (b.ny + 3) >> 2;
    const long nblk = (long)ntx * nty * b.nz;
    const long per = (nblk + 7) >> 3;
    const long t = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= nblk) return false;
    int tx, ty;
    if constexpr (FAST == TILE_Z_FAST) {
        k = (int)(t % b.nz);
        const long r = t / b.nz;
        tx = (int)(r % ntx);
        ty = (int)(r / ntx);
    } else {
        ty = (int)(t % nty);
        const long r = t / nty;
        k = (int)(r % b.nz);
        tx = (int)(r / b.nz);
    }
    i = tx * 64 + threadIdx.x;
    j = ty * 4 + threadIdx.y;
    return true;
}

}  // namespace ns3d
