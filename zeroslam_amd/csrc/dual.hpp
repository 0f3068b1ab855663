// First-order forward-mode dual numbers v + d eps (eps^2 = 0) for the templated
// bundle-adjustment arithmetic of ba_math.hpp: evaluated on duals, the
// reference's retraction, residual, Jacobians, robust weight and gradient give
// their exact directional derivatives.  Comparisons look at the value only, so a
// dual evaluation takes the same branches as the binary64 one.
#pragma once

#include "ba_math.hpp"

namespace zb {

struct dual {
    double v, d;
    ZB_HD dual() : v(0.0), d(0.0) {}
    ZB_HD dual(double x) : v(x), d(0.0) {}  // NOLINT: constants in the templates
    ZB_HD dual(double x, double t) : v(x), d(t) {}
    ZB_HD dual& operator+=(dual b) { v += b.v; d += b.d; return *this; }
    ZB_HD dual& operator-=(dual b) { v -= b.v; d -= b.d; return *this; }
    ZB_HD dual& operator*=(dual b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
};
ZB_HD dual operator+(dual a, dual b) { return dual(a.v + b.v, a.d + b.d); }
ZB_HD dual operator-(dual a, dual b) { return dual(a.v - b.v, a.d - b.d); }
ZB_HD dual operator-(dual a) { return dual(-a.v, -a.d); }
ZB_HD dual operator+(dual a) { return a; }
ZB_HD dual operator*(dual a, dual b) { return dual(a.v * b.v, a.d * b.v + a.v * b.d); }
ZB_HD dual operator/(dual a, dual b) { return dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
ZB_HD bool operator<(dual a, dual b) { return a.v < b.v; }
ZB_HD bool operator<=(dual a, dual b) { return a.v <= b.v; }
ZB_HD bool operator>(dual a, dual b) { return a.v > b.v; }
ZB_HD bool operator>=(dual a, dual b) { return a.v >= b.v; }
ZB_HD dual sqrt(dual a) {
    const double s = ::sqrt(a.v);
    return dual(s, a.d / (2.0 * s));
}
ZB_HD dual log(dual a) { return dual(::log(a.v), a.d / a.v); }
ZB_HD dual sin(dual a) { return dual(::sin(a.v), a.d * ::cos(a.v)); }
ZB_HD dual cos(dual a) { return dual(::cos(a.v), -a.d * ::sin(a.v)); }
ZB_HD dual fabs(dual a) { return a.v < 0.0 ? -a : a; }

}  // namespace zb
