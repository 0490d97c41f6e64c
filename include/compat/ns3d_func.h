/* ns3d_func.h -- the drop-in's free functions under the reference's header
 * name, with the standard headers the reference's one brings in
 * (ns3d_func.h:1-13 there), so that translation units written against it
 * (maincycle.cpp) compile unchanged against libns3d_dropin.
 *
 * One rank, no MPI: MPI_Wtime() is the steady clock in seconds unless an
 * MPI header was included first. */
#ifndef NS3D_COMPAT_NS3D_FUNC_H
#define NS3D_COMPAT_NS3D_FUNC_H

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "field_structures.h"

#ifndef MPI_VERSION
inline double MPI_Wtime() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
#endif

// main time-advance loop with monitor and output (maincycle.cpp:7-113); the
// library does not define it: the caller's maincycle.cpp does
void time_advance(Field3D &F, CartDecomp &C, GridDesc &G, SolverParams &P);

#endif
