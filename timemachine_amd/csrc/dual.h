// Forward-mode dual numbers: a value and its derivative along one tangent
// direction. Evaluating a gradient in dual arithmetic along (vx, vp), with
// tangent vx on the coordinates and vp on the parameters, gives the
// gradient's directional derivative: H_xx vx + H_xp vp for du/dx and
// H_px vx + H_pp vp for du/dp (H the Hessian of U in (x, p)). By the symmetry
// of H this is also the vector-Jacobian product of the gradient with (vx, vp),
// what reverse-mode differentiation of a loss on forces needs (the adjoint of
// du_dx).
//
// Derivatives are the exact ones of each function; values use the FP64
// library functions. The type and its functions live in a namespace of their
// own, found by argument-dependent lookup: unqualified math on doubles
// elsewhere in tmamd never resolves to them.
#pragma once

#include <cmath>

#include "hd.h"

namespace tmamd {
namespace dual {

struct Dual {
    double v; // value
    double d; // derivative along the tangent
    TM_HD Dual() : v(0.0), d(0.0) {}
    TM_HD Dual(double x) : v(x), d(0.0) {} // constants carry no tangent
    TM_HD Dual(double x, double dx) : v(x), d(dx) {}
};

TM_HD inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
TM_HD inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
TM_HD inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
TM_HD inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
TM_HD inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
TM_HD inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
TM_HD inline bool operator<(Dual a, Dual b) { return a.v < b.v; }

TM_HD inline Dual sqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
TM_HD inline Dual exp(Dual a) {
    const double e = std::exp(a.v);
    return Dual(e, e * a.d);
}
TM_HD inline Dual sin(Dual a) { return Dual(std::sin(a.v), std::cos(a.v) * a.d); }
TM_HD inline Dual cos(Dual a) { return Dual(std::cos(a.v), -std::sin(a.v) * a.d); }
TM_HD inline Dual atan2(Dual y, Dual x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return Dual(std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2);
}
// erfc'(x) = -2 / sqrt(pi) exp(-x^2)
TM_HD inline Dual erfc(Dual a) {
    return Dual(std::erfc(a.v), -1.1283791670955126 * std::exp(-a.v * a.v) * a.d);
}

} // namespace dual

using dual::Dual;

} // namespace tmamd
