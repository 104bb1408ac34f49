// Numeric building blocks shared by host and device code.
//
// Everything here is deterministic under IEEE-754 binary64 with
// round-to-nearest and no FMA contraction (the library is built with
// -ffp-contract=off); fma is only used where it is written out, so host and
// device produce bit-identical results.
//
//  * double-double (Dekker 1971; Knuth TAOCP 4.2.2 TwoSum; QD library
//    add/mul/div/sqrt of Hida, Li & Bailey 2000) for the scalars xx, scale
//    and the subspace refinement -- the reference keeps these in binary128
//    (src/config.h:44-51, src/wavefunction.h:163-198, src/solver.h:455-538);
//  * an exact fixed-point format with 2^-64 units for z, new_z and x.z, so
//    that scatter-adds from many columns commute exactly (z within +-2^31,
//    at most 2^32 adds into one slot between two reinserts);
//  * the 64-lane reduction order used by every non-exact reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define MC_HD __host__ __device__ __forceinline__

namespace mc {

typedef __int128 i128;

struct dd {
  double hi, lo;
};

MC_HD dd mkdd(double h) { return dd{h, 0.0}; }

MC_HD dd two_sum(double a, double b) {
  double s = a + b;
  double v = s - a;
  double e = (a - (s - v)) + (b - v);
  return dd{s, e};
}

MC_HD dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return dd{s, e};
}

MC_HD dd two_prod(double a, double b) {
  double p = a * b;
  return dd{p, __builtin_fma(a, b, -p)};
}

MC_HD dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  dd u = fast_two_sum(s.hi, s.lo + t.hi);
  return fast_two_sum(u.hi, u.lo + t.lo);
}

MC_HD dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }
MC_HD dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }

MC_HD dd dd_add_d(dd a, double b) {
  dd s = two_sum(a.hi, b);
  return fast_two_sum(s.hi, s.lo + a.lo);
}

MC_HD dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  double cross = a.hi * b.lo;
  double cross2 = a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo + (cross + cross2));
}

MC_HD dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  double c = a.lo * b;
  return fast_two_sum(p.hi, p.lo + c);
}

MC_HD dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.hi / b.hi;
  return dd_add_d(fast_two_sum(q1, q2), q3);
}

MC_HD dd dd_div_d(dd a, double b) { return dd_div(a, mkdd(b)); }

// double-double divided by a double (QD's div(dd, double)).
MC_HD dd dd_div_dbl(dd a, double b) {
  double q1 = a.hi / b;
  dd p = two_prod(q1, b);
  dd s = two_sum(a.hi, -p.hi);
  double sl = (s.lo - p.lo) + a.lo;
  double q2 = (s.hi + sl) / b;
  return fast_two_sum(q1, q2);
}

// double-double divided by a double b, given rb = 1/b rounded: the two
// quotient digits of dd_div_dbl, each formed by a multiplication with rb and
// corrected through the exact remainder.
MC_HD dd dd_div_rcp(dd a, double b, double rb) {
  double q1 = a.hi * rb;
  dd p = two_prod(q1, b);
  dd s = two_sum(a.hi, -p.hi);
  double sl = (s.lo - p.lo) + a.lo;
  double q2 = (s.hi + sl) * rb;
  return fast_two_sum(q1, q2);
}

MC_HD dd dd_sqrt(dd a) {
  if (!(a.hi > 0.0)) return dd{__builtin_sqrt(a.hi), 0.0};
  double r = 1.0 / __builtin_sqrt(a.hi);
  double s = a.hi * r;
  dd d = dd_sub(a, two_prod(s, s));
  return two_sum(s, d.hi * (r * 0.5));
}

MC_HD bool dd_neg_p(dd a) { return a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0); }
MC_HD bool dd_zero_p(dd a) { return a.hi == 0.0 && a.lo == 0.0; }
MC_HD dd dd_abs(dd a) { return dd_neg_p(a) ? dd_neg(a) : a; }
MC_HD bool dd_ge(dd a, dd b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }

// |a| > t
MC_HD bool dd_abs_gt(dd a, double t) {
  double h = a.hi, l = a.lo;
  if (h < 0.0) {
    h = -h;
    l = -l;
  }
  return h > t || (h == t && l > 0.0);
}

// ---------------------------------------------------------------- fixed point
constexpr int kFixBits = 64;

// round-half-even(d * 2^64), exact for |d| < 2^62
MC_HD i128 to_fix(double d) {
  double y = __builtin_ldexp(d, kFixBits);
  if (__builtin_fabs(y) < 4503599627370496.0) return (i128)(int64_t)__builtin_rint(y);
  // |y| >= 2^52: y is an integer; split |y| exactly into 64 + 32 + 32 bits.
  const bool neg = y < 0.0;
  const double ay = neg ? -y : y;
  const double yh = __builtin_floor(__builtin_ldexp(ay, -64));
  const double ym = ay - __builtin_ldexp(yh, 64);  // in [0, 2^64), exact
  const double ymh = __builtin_floor(__builtin_ldexp(ym, -32));
  const double yml = ym - __builtin_ldexp(ymh, 32);  // in [0, 2^32), exact
  const i128 v = (((i128)(int64_t)yh) << 64) + (((i128)(int64_t)ymh) << 32) + (i128)(int64_t)yml;
  return neg ? -v : v;
}

MC_HD double fix_to_d_pos(i128 v) {
  int64_t hi = (int64_t)(v >> 64);
  uint64_t lo = (uint64_t)v;
  double dlo = (double)(uint32_t)(lo >> 32) * 4294967296.0 + (double)(uint32_t)lo;
  double d = (double)hi * 18446744073709551616.0 + dlo;
  return __builtin_ldexp(d, -kFixBits);
}

// Sign-magnitude conversion: the magnitude is rounded, then negated, so
// small negative values do not cancel against a two's-complement high word.
MC_HD double fix_to_d(i128 v) { return v < 0 ? -fix_to_d_pos(-v) : fix_to_d_pos(v); }

MC_HD dd fix_to_dd(i128 v) {
  double h = fix_to_d(v);
  return fast_two_sum(h, fix_to_d(v - to_fix(h)));
}

// A table z is stored as two int64 accumulators: z = (A * 2^32 + B) * 2^-64.
// Adding v splits it so that both parts can be added with 64-bit atomics.
MC_HD void fix_split(i128 v, int64_t& a, int64_t& b) {
  int64_t lo = (int64_t)(int32_t)(uint32_t)(uint64_t)v;
  b = lo;
  a = (int64_t)((v - (i128)lo) >> 32);
}

MC_HD i128 fix_join(int64_t a, int64_t b) { return (((i128)a) << 32) + (i128)b; }

// A table z is two int64 words, z = (A * 2^53 + B) * 2^-64 exactly.  The
// scatter adds a contribution |v| < 2^53 to B alone (one atomic) and splits
// a larger one into A and B in [0, 2^53) (two).  One determinant receives at
// most one add per column, k <= 256 per iteration, so B moves by less than
// 2^61 per iteration; every iteration folds each touched B with |B| > 2^61
// into A (pass 2 lists them, the merge folds them), so |B| stays below 2^62.
constexpr int kZShift = 53;
MC_HD i128 tz_join(int64_t a, int64_t b) { return (((i128)a) << kZShift) + (i128)b; }
MC_HD void tz_split(i128 v, int64_t& a, int64_t& b) {
  a = (int64_t)(v >> kZShift);                  // floor
  b = (int64_t)(v - (((i128)a) << kZShift));    // in [0, 2^53)
}

// ---------------------------------------------------------------- hashing
MC_HD uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

}  // namespace mc
