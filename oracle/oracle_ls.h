// Line-search coordinate update -- test infrastructure (the oracle), never
// the product path.
//
// Restates CoordinateUpdateLS::operator() (src/solver.h:254-314): for the
// picked coordinate with value x, z = (Hx)_i, diagonal h = H_ii and
// xx = |x|^2, the new value t minimises f = |H + b b^T|_F^2 along the
// coordinate, a root of g(t) = t^3 + p1 t + q with p1 = xx - x^2 + h and
// q = z - h x (solver.h:268-272).  The reference picks the root by Cardano
// or the trigonometric form (solver.h:274-290): the only real root, or with
// three real roots the largest when q < 0 and the smallest when q >= 0 --
// the outer root on the side opposite to q.  Here that root comes from
// Newton's iteration started outside every root (|t| <= max(1, |p1|+|q|),
// Cauchy), which approaches the outer root monotonically, with IEEE basic
// operations only so that host and device agree bit for bit; then the
// reference's Newton polish on dx = t - x runs unchanged (solver.h:292-307,
// returning the last dx it accepted).
//
// The reference never updates xx or scale after a line-search step
// (solver.h:743-744 updates them only inside CoordinateUpdateEig), so its
// "ls" runs use a stale |x|^2.  This restatement keeps them: scale = 1
// (no rescaling, z = Hx exactly) and xx += dx (2x + dx) in double-double.
#pragma once

#include <cmath>

#include "oracle_num.h"

namespace oc {

// Outer root of t^3 + p1 t + q on the side opposite to q (q = 0: the
// smallest root).
static inline double ls_root(double p1, double q) {
  const double s = q < 0.0 ? 1.0 : -1.0;
  double t = s * std::fmax(1.0, std::fabs(p1) + std::fabs(q));
  for (int it = 0; it < 400; ++it) {
    const double g = (t * t + p1) * t + q;
    const double gp = 3.0 * (t * t) + p1;
    const double tn = t - g / gp;
    if (!(s * tn < s * t)) break;  // no longer moving inward: at the root to rounding
    t = tn;
  }
  return t;
}

// dx of the picked coordinate; updates xx and sets scale = 1.
static inline double ls_update(double h, double x, double z, dd& xx, dd& sc) {
  const dd x2 = two_prod(x, x);
  const double P1 = dd_add_d(dd_add(xx, dd{-x2.hi, -x2.lo}), h).hi;  // xx - x^2 - dA, dA = -h
  const double Q = z + (-h) * x;                                       // z + dA x
  double dx = ls_root(P1, Q) - x;
  auto newton = [&](double d) {
    return d - (d * (d * (d + 3 * x) + (3 * x * x + P1)) + P1 * x + Q + x * x * x) /
                   (d * (3 * d + 6 * x) + 3 * x * x + P1);
  };
  double dxn = newton(dx);
  int iter = 0;
  while (std::fabs((dxn - dx) / dx) > 1e-12 && iter < 10) {
    dx = dxn;
    dxn = newton(dx);
    ++iter;
  }
  const dd t = two_sum(2.0 * x, dx);
  xx = dd_add(xx, dd_mul_d(t, dx));
  sc = dd_make(1.0);
  return dx;
}

}  // namespace oc
