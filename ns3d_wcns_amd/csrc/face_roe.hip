// Face-flux kernels for the roe Riemann solver (all directions, interpolations
// and interpolation variables).
#include "ns3d_face_impl.h"

NS3D_INSTANTIATE_FACE(ns3d::R_ROE)
