// ns3d_observe.h -- launchers of the device observables (ns3d_observe.hip).
#pragma once
#include "ns3d_common.h"

namespace ns3d {

struct MomentArgs {
    const double *rho, *u, *v, *w, *T;  // ghosted
    Box b;
    double mu, S_ref;
    double *partial;  // moment_work_doubles() - 8
    double *out;      // 8 sums: rho, mu, u, v, w, u^2, v^2, w^2
};

// Doubles of device work memory for one spectrum with nshell shells.
long spectrum_work_doubles(const Box &b, int nshell);
long moment_work_doubles();
// E(k) of the three velocity fields (ghosted) into spectrum_result(); plan
// is created on first use (returns nonzero on an FFT library error).
int launch_spectrum(const Box &b, const double *const *vel, int nshell, double *work, void *&plan, hipStream_t s);
double *spectrum_result(const Box &b, int nshell, double *work);
void destroy_fft_plan(void *plan);
void launch_moments(const MomentArgs &a, hipStream_t s);

}  // namespace ns3d
