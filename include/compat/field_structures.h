/* field_structures.h -- the drop-in's types (GridDesc, LocalDesc,
 * SolverParams, HaloRequests, Field3D, idx3 / idx_f*) under the reference's
 * header name, with the standard headers the reference's one brings in
 * (field_structures.h:19-28 there), so that translation units that include
 * it compile unchanged against libns3d_dropin.  One rank, no MPI. */
#ifndef NS3D_COMPAT_FIELD_STRUCTURES_H
#define NS3D_COMPAT_FIELD_STRUCTURES_H

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../ns3d_dropin.h"

#endif
