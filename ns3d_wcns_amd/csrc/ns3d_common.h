// ns3d_common.h -- kernel argument blocks and launch helpers shared by the
// HIP translation units of libns3d_hip.
#pragma once
#include <hip/hip_runtime.h>

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

template <int RS>
void launch_face(int dir, int interp, bool chr, int precision, const FaceArgs<double> &a64,
                 const FaceArgs<float> &a32, hipStream_t s);

// 256-thread blocks of 64 x 4 interior cells, one z plane per block row.
inline dim3 cell_block() { return dim3(64, 4, 1); }
inline dim3 cell_grid(const Box &b) { return dim3((b.nx + 63) / 64, (b.ny + 3) / 4, b.nz); }

}  // namespace ns3d
