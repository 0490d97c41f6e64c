// ns3d_dual.h -- forward-mode dual number (value, directional derivative)
// for the tangent-linear model.  The device numerics are templated on their
// arithmetic type, so the tangent of any face flux or viscous formula is the
// same code instantiated on Dual.
//
// Non-smooth operations take a fixed subgradient: d|x| = sign(x) with
// sign(0) = +1; fmin / fmax follow the operand they return, the first one
// on ties.  The hand-written reverse mode (ns3d_adjoint.h) makes the same
// choices, so tangent and adjoint are transposes of one linear map.
#pragma once
#include <cmath>

#ifndef NS3D_HD
#define NS3D_HD __host__ __device__ __forceinline__
#endif

namespace ns3d {
// Own namespace: its sqrt / fabs / fmin / fmax are found by argument-dependent
// lookup for Dual only and never hide the double overloads inside ns3d.
namespace ad {

struct Dual {
    double v, d;
    NS3D_HD Dual() : v(0.0), d(0.0) {}
    NS3D_HD Dual(double x) : v(x), d(0.0) {}
    NS3D_HD Dual(double x, double dx) : v(x), d(dx) {}
};

NS3D_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
NS3D_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
NS3D_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
NS3D_HD Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
NS3D_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
NS3D_HD Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
NS3D_HD Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
NS3D_HD Dual &operator*=(Dual &a, Dual b) { return a = a * b; }
NS3D_HD bool operator<(Dual a, Dual b) { return a.v < b.v; }
NS3D_HD bool operator>(Dual a, Dual b) { return a.v > b.v; }
NS3D_HD bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
NS3D_HD bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
NS3D_HD Dual sqrt(Dual a) {
    const double s = ::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
NS3D_HD Dual fabs(Dual a) { return a.v >= 0.0 ? a : -a; }
NS3D_HD Dual fmin(Dual a, Dual b) { return b.v < a.v ? b : a; }
NS3D_HD Dual fmax(Dual a, Dual b) { return b.v > a.v ? b : a; }

}  // namespace ad
using ad::Dual;
}  // namespace ns3d
