// ns3d_observe.hip -- isotropic-turbulence observables on the device: the
// shell-summed energy spectrum E(k) (compute_energy_spectrum_rank0,
// EnergySpectrum.cpp:63-156) and the one-point statistics of
// compute_turbulence_statistics (ana_isotropic.cpp:238-305).
//
// E(k): each velocity component's interior is packed into a complex array,
// transformed by a 3D forward hipFFT (Z2Z) and normalised by 1/N; the modal
// energies 0.5 (|U|^2 + |V|^2 + |W|^2) accumulate per cell, and a shell
// kernel sums them into shells lround(|k|) (wavenumber i <= N/2 counts as +i,
// others as i - N).  Shell sums: per block in LDS, then over blocks in a
// fixed order.
//
// Statistics: fixed-order two-level sums of rho, mu(T), u, v, w, u^2, v^2,
// w^2 over the interior; the dissipation 2 nu sum k^2 E(k) uses the spectrum.
#include <hipfft/hipfft.h>

#include "ns3d_kernels.h"
#include "ns3d_observe.h"

namespace ns3d {

namespace {

constexpr int OBS_BLOCKS = 2048;

__global__ void __launch_bounds__(256) k_pack(const double *f, Box b, double2 *z) {
    const long n = (long)b.nx * b.ny * b.nz;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id % b.nx);
        const long r = id / b.nx;
        const int j = (int)(r % b.ny), k = (int)(r / b.ny);
        z[id] = make_double2(f[b.gidx(i, j, k)], 0.0);
    }
}

// e[id] (+)= 0.5 |z[id] / N|^2
__global__ void __launch_bounds__(256) k_modal(const double2 *z, long n, double inv_n, int first, double *e) {
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const double re = z[id].x * inv_n, im = z[id].y * inv_n;
        const double v = (re * re + im * im) * 0.5;
        e[id] = first ? v : e[id] + v;
    }
}

__device__ __forceinline__ int wavenumber(int i, int n) { return i <= n / 2 ? i : i - n; }

// Per block: shell sums of a contiguous cell range in LDS (shared atomics),
// written to partial[block][shell].
__global__ void __launch_bounds__(256) k_shells(const double *e, Box b, int nshell, double *partial) {
    extern __shared__ double sh[];
    for (int s = threadIdx.x; s < nshell; s += blockDim.x) sh[s] = 0.0;
    __syncthreads();
    const long n = (long)b.nx * b.ny * b.nz;
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per, hi = min(lo + per, n);
    for (long id = lo + threadIdx.x; id < hi; id += blockDim.x) {
        const int i = (int)(id % b.nx);
        const long r = id / b.nx;
        const int j = (int)(r % b.ny), k = (int)(r / b.ny);
        const int k1 = wavenumber(i, b.nx), k2 = wavenumber(j, b.ny), k3 = wavenumber(k, b.nz);
        const int s = (int)lround(sqrt(1.0 * (k1 * k1 + k2 * k2 + k3 * k3)));
        if (s < nshell) atomicAdd(&sh[s], e[id]);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nshell; s += blockDim.x) partial[(long)blockIdx.x * nshell + s] = sh[s];
}

// out[s] = sum over blocks of partial[block][s], in block order
__global__ void __launch_bounds__(256) k_shell_final(const double *partial, int nblocks, int nshell, double *out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshell) return;
    double acc = 0.0;
    for (int q = 0; q < nblocks; ++q) acc += partial[(long)q * nshell + s];
    out[s] = acc;
}

// Per block: sums of rho, mu(T), u, v, w, u^2, v^2, w^2 over a grid-stride
// subset of the interior, reduced in the block in a fixed tree order.
__global__ void __launch_bounds__(256) k_moments(MomentArgs a) {
    __shared__ double sh[8][256];
    const Box &b = a.b;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long n = (long)b.nx * b.ny * b.nz;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id % b.nx);
        const long r = id / b.nx;
        const int j = (int)(r % b.ny), k = (int)(r / b.ny);
        const long g = b.gidx(i, j, k);
        const double u = a.u[g], v = a.v[g], w = a.w[g], T = a.T[g];
        s[0] += a.rho[g];
        s[1] += sutherland(a.mu, a.S_ref, T);
        s[2] += u;
        s[3] += v;
        s[4] += w;
        s[5] += u * u;
        s[6] += v * v;
        s[7] += w * w;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half)
#pragma unroll
            for (int q = 0; q < 8; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x < 8) a.partial[(long)blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(64) k_moments_final(const double *partial, int nblocks, double *out) {
    const int q = threadIdx.x;
    if (q >= 8) return;
    double acc = 0.0;
    for (int blk = 0; blk < nblocks; ++blk) acc += partial[(long)blk * 8 + q];
    out[q] = acc;
}

}  // namespace

long spectrum_work_doubles(const Box &b, int nshell) {
    const long n = (long)b.nx * b.ny * b.nz;
    return 2 * n + n + (long)OBS_BLOCKS * nshell + nshell;  // complex field, modal energy, partials, result
}

long moment_work_doubles() { return (long)OBS_BLOCKS * 8 + 8; }

int launch_spectrum(const Box &b, const double *const *vel, int nshell, double *work, void *&plan,
                    hipStream_t s) {
    const long n = (long)b.nx * b.ny * b.nz;
    double2 *z = reinterpret_cast<double2 *>(work);
    double *e = work + 2 * n;
    double *partial = e + n;
    double *out = partial + (long)OBS_BLOCKS * nshell;
    if (!plan) {
        hipfftHandle h;
        if (hipfftPlan3d(&h, b.nz, b.ny, b.nx, HIPFFT_Z2Z) != HIPFFT_SUCCESS) return -1;
        plan = (void *)h;
    }
    const hipfftHandle h = (hipfftHandle)plan;
    if (hipfftSetStream(h, s) != HIPFFT_SUCCESS) return -1;
    const unsigned g = (unsigned)std::min<long>((n + 255) / 256, 65536);
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_pack, dim3(g), dim3(256), 0, s, vel[c], b, z);
        if (hipfftExecZ2Z(h, reinterpret_cast<hipfftDoubleComplex *>(z), reinterpret_cast<hipfftDoubleComplex *>(z),
                          HIPFFT_FORWARD) != HIPFFT_SUCCESS)
            return -1;
        hipLaunchKernelGGL(k_modal, dim3(g), dim3(256), 0, s, (const double2 *)z, n, 1.0 / (double)n, c == 0 ? 1 : 0,
                           e);
    }
    hipLaunchKernelGGL(k_shells, dim3(OBS_BLOCKS), dim3(256), (size_t)nshell * sizeof(double), s, (const double *)e,
                       b, nshell, partial);
    hipLaunchKernelGGL(k_shell_final, dim3((nshell + 255) / 256), dim3(256), 0, s, (const double *)partial,
                       OBS_BLOCKS, nshell, out);
    return 0;
}

double *spectrum_result(const Box &b, int nshell, double *work) {
    const long n = (long)b.nx * b.ny * b.nz;
    return work + 3 * n + (long)OBS_BLOCKS * nshell;
}

void destroy_fft_plan(void *plan) {
    if (plan) hipfftDestroy((hipfftHandle)plan);
}

void launch_moments(const MomentArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_moments, dim3(OBS_BLOCKS), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, s, (const double *)a.partial, OBS_BLOCKS, a.out);
}

}  // namespace ns3d
