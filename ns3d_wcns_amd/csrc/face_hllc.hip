// Face-flux and pencil kernels for the hllc Riemann solver (all directions, interpolations
// and interpolation variables).
#include "ns3d_face_impl.h"
#include "ns3d_pencil_impl.h"

NS3D_INSTANTIATE_FACE(ns3d::R_HLLC)
NS3D_INSTANTIATE_PENCILS(ns3d::R_HLLC)
