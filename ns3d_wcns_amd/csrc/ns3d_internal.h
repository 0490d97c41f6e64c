// ns3d_internal.h -- library state shared between translation units.
#pragma once
#include "../../include/ns3d_hip.h"
#include "ns3d_device.h"

namespace ns3d {

// Records msg as the library's last error (ns3d_last_error) and returns code.
int set_error(int code, const char *msg);
// The FP64 scheme coefficients of a parameter set (linear MDCD weights,
// MDCD-WENO linear weights; no reference-state offsets).
SchemeCoef<double> scheme_coef64(const ns3d_params &P);

}  // namespace ns3d
