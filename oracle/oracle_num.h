// Oracle numerics: double-double arithmetic, exact fixed-point sums and the
// canonical 64-lane reduction.  TEST INFRASTRUCTURE ONLY: the oracle is the
// checker for the HIP path and is never linked into the product.
//
// Every routine here is written so that IEEE-754 binary64 with
// round-to-nearest, no FMA contraction (-ffp-contract=off) and correctly
// rounded fma/sqrt/div gives one result on any conforming machine.
//
// Double-double algorithms: Dekker (1971) and Knuth TAOCP vol. 2 (TwoSum,
// FastTwoSum, TwoProd via fma); the composed add/mul/div/sqrt follow the
// published QD library algorithms of Hida, Li and Bailey (2000).
#pragma once

#include <cmath>
#include <cstdint>

namespace oc {

typedef __int128 i128;

struct dd {
  double hi, lo;
};

static inline dd dd_make(double h) { return dd{h, 0.0}; }

static inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return dd{s, e};
}

static inline dd quick_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return dd{s, e};
}

static inline dd two_prod(double a, double b) {
  double p = a * b;
  double e = std::fma(a, b, -p);
  return dd{p, e};
}

static inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  double s2 = s.lo + t.hi;
  dd u = quick_two_sum(s.hi, s2);
  double u2 = u.lo + t.lo;
  return quick_two_sum(u.hi, u2);
}

static inline dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }
static inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }

static inline dd dd_add_d(dd a, double b) {
  dd s = two_sum(a.hi, b);
  double s2 = s.lo + a.lo;
  return quick_two_sum(s.hi, s2);
}

static inline dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  double t = a.hi * b.lo;
  double u = a.lo * b.hi;
  double p2 = p.lo + (t + u);
  return quick_two_sum(p.hi, p2);
}

static inline dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  double t = a.lo * b;
  double p2 = p.lo + t;
  return quick_two_sum(p.hi, p2);
}

static inline dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.hi / b.hi;
  dd q = quick_two_sum(q1, q2);
  return dd_add_d(q, q3);
}

static inline dd dd_div_d(dd a, double b) { return dd_div(a, dd_make(b)); }

// double-double divided by a double (QD's div(dd, double)).
static inline dd dd_div_dbl(dd a, double b) {
  double q1 = a.hi / b;
  dd p = two_prod(q1, b);
  dd s = two_sum(a.hi, -p.hi);
  double sl = s.lo - p.lo;
  sl = sl + a.lo;
  double q2 = (s.hi + sl) / b;
  return quick_two_sum(q1, q2);
}

// double-double divided by a double b, given rb = 1/b rounded: the same
// two quotient digits as dd_div_dbl, each formed by a multiplication with
// rb and corrected through the exact remainder.
static inline dd dd_div_rcp(dd a, double b, double rb) {
  double q1 = a.hi * rb;
  dd p = two_prod(q1, b);
  dd s = two_sum(a.hi, -p.hi);
  double sl = s.lo - p.lo;
  sl = sl + a.lo;
  double q2 = (s.hi + sl) * rb;
  return quick_two_sum(q1, q2);
}

static inline dd dd_sqrt(dd a) {
  if (!(a.hi > 0.0)) return dd{std::sqrt(a.hi), 0.0};  // 0 stays 0, negative gives NaN
  double x = 1.0 / std::sqrt(a.hi);
  double ax = a.hi * x;
  dd sq = two_prod(ax, ax);
  dd diff = dd_sub(a, sq);
  double corr = diff.hi * (x * 0.5);
  return two_sum(ax, corr);
}

static inline bool dd_lt0(dd a) { return a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0); }

// |a| > t for a double threshold t.
static inline bool dd_abs_gt(dd a, double t) {
  double h = a.hi, l = a.lo;
  if (h < 0.0) {
    h = -h;
    l = -l;
  }
  return h > t || (h == t && l > 0.0);
}

// ---- exact fixed point: value = v * 2^-FIX_BITS ----------------------------
static const int FIX_BITS = 64;

// Round-half-even of d * 2^64, exact for |d| < 2^62.
static inline i128 to_fixed(double d) {
  double y = std::ldexp(d, FIX_BITS);
  if (std::fabs(y) < 4503599627370496.0) {
    double r = std::rint(y);
    return (i128)(int64_t)r;
  }
  // |y| >= 2^52: y is an integer; split |y| exactly into 64 + 32 + 32 bits.
  const bool neg = y < 0.0;
  const double ay = neg ? -y : y;
  const double yh = std::floor(std::ldexp(ay, -64));
  const double ym = ay - std::ldexp(yh, 64);  // in [0, 2^64), exact
  const double ymh = std::floor(std::ldexp(ym, -32));
  const double yml = ym - std::ldexp(ymh, 32);  // in [0, 2^32), exact
  const i128 v = (((i128)(int64_t)yh) << 64) + (((i128)(int64_t)ymh) << 32) + (i128)(int64_t)yml;
  return neg ? -v : v;
}

static inline double fixed_to_double(i128 v) {
  if (v < 0) return -fixed_to_double(-v);
  int64_t hi = (int64_t)(v >> 64);
  uint64_t lo = (uint64_t)v;
  double dlo = (double)(uint32_t)(lo >> 32) * 4294967296.0 + (double)(uint32_t)lo;
  double d = (double)hi * 18446744073709551616.0 + dlo;
  return std::ldexp(d, -FIX_BITS);
}

static inline dd fixed_to_dd(i128 v) {
  double h = fixed_to_double(v);
  i128 rem = v - to_fixed(h);
  return quick_two_sum(h, fixed_to_double(rem));
}

// ---- canonical reduction: 64 strided partial sums, then a binary tree -----
// partial[l] = sum over i = l, l+64, l+128, ... (ascending) starting from 0;
// then for off = 32, 16, ..., 1: partial[l] += partial[l + off] (l < off).
template <class F>
static inline double red64(int n, F term) {
  double part[64];
  for (int l = 0; l < 64; ++l) {
    double acc = 0.0;
    for (int i = l; i < n; i += 64) acc = acc + term(i);
    part[l] = acc;
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) part[l] = part[l] + part[l + off];
  return part[0];
}

template <class F>
static inline dd red64_dd(int n, F term) {
  dd part[64];
  for (int l = 0; l < 64; ++l) {
    dd acc{0.0, 0.0};
    for (int i = l; i < n; i += 64) acc = dd_add(acc, term(i));
    part[l] = acc;
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) part[l] = dd_add(part[l], part[l + off]);
  return part[0];
}

}  // namespace oc
