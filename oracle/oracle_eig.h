// Oracle: canonical lowest-eigenpair solver and the CDFCI subspace update.
// TEST INFRASTRUCTURE ONLY (checker for the HIP path, never shipped).
//
// The reference calls LAPACK dsyevr (RANGE='I', IL=IU=1) and then runs one
// step of inverse iteration in __float128 on an LU factorisation
// (src/solver.h:333-609).  LAPACK is an external, unversioned dependency
// (Makefile:25), so this restatement fixes one deterministic algorithm that
// the device solver reproduces operation for operation:
//   1. Householder tridiagonalisation (LAPACK dsytd2 conventions),
// with every matrix-vector row sum (here and in the subspace update) taken
// as four interleaved partial sums, c = j mod 4, combined as
// (S0 + S1) + (S2 + S3) (dot4),
//   2. Laguerre's iteration on det(T - lambda I) from the Gershgorin lower
//      bound for the lowest eigenvalue of the tridiagonal matrix T (for a
//      polynomial with real roots it rises monotonically to the smallest
//      root, cubically near it); each pass is one leading-minor recurrence
//      with its first and second derivatives (no division on the chain),
//      and lambda stays a point where every leading minor is positive,
//   3. three steps of inverse iteration on T in double, sharing one LU
//      factorisation with partial pivoting (dgttrf style); back substitution
//      multiplies by the rounded reciprocals of the pivots,
//   4. the reference's extended-precision refinement step: one more solve
//      with shift lambda - 1e-14 (solver.h:439) using double LU factors and
//      a double-double vector (as the reference uses double LU factors with
//      binary128 vectors, solver.h:437-503), back-transformed with the
//      Householder vectors in double-double and normalised.
// Reductions use the canonical 64-lane order of oracle_num.h.
#pragma once

#include <algorithm>
#include <vector>

#include "oracle_num.h"

namespace oc {

struct EigPair {
  double lambda = 0.0;  // bisection eigenvalue of the tridiagonal form
  std::vector<dd> q;    // refined, normalised eigenvector (q[0] >= 0)
};

// Row sum in the canonical four-way order: S_j = sum over c = j mod 4 in
// ascending c, combined as (S0 + S1) + (S2 + S3).
template <class F>
static inline double dot4(int m, F term) {
  double S[4];
  for (int j = 0; j < 4; ++j) {
    double acc = 0.0;
    for (int c = j; c < m; c += 4) acc = acc + term(c);
    S[j] = acc;
  }
  return (S[0] + S[1]) + (S[2] + S[3]);
}

static inline bool dd_is_zero(dd a) { return a.hi == 0.0 && a.lo == 0.0; }
static inline dd dd_abs(dd a) { return dd_lt0(a) ? dd_neg(a) : a; }
static inline bool dd_ge(dd a, dd b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }

// LU factorisation with partial pivoting of T - shift*I, T = tridiag(e, d, e)
// (LAPACK dgttrf conventions: row i and i+1 swap when |sub| > |diag|).
struct TriLU {
  std::vector<double> dg, du, du2, mul;
  std::vector<char> swp;
};

static TriLU tri_factor(int n, const double* d, const double* e, double shift) {
  const double tiny = 1e-290;
  TriLU F;
  F.dg.assign(n, 0.0);
  F.du.assign(n > 1 ? n - 1 : 1, 0.0);
  F.du2.assign(n > 2 ? n - 2 : 1, 0.0);
  F.mul.assign(n > 1 ? n - 1 : 1, 0.0);
  F.swp.assign(n > 1 ? n - 1 : 1, 0);
  std::vector<double> dl(n > 1 ? n - 1 : 1);
  for (int i = 0; i < n; ++i) F.dg[i] = d[i] - shift;
  for (int i = 0; i + 1 < n; ++i) {
    F.du[i] = e[i];
    dl[i] = e[i];
  }
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(F.dg[i]) >= std::fabs(dl[i])) {
      if (F.dg[i] == 0.0) F.dg[i] = tiny;
      double m = dl[i] / F.dg[i];
      F.mul[i] = m;
      F.swp[i] = 0;
      F.dg[i + 1] = F.dg[i + 1] - m * F.du[i];
    } else {
      double m = F.dg[i] / dl[i];
      F.mul[i] = m;
      F.swp[i] = 1;
      double ndgi = dl[i];
      double ndui = F.dg[i + 1];
      double ndgi1 = F.du[i] - m * F.dg[i + 1];
      if (i < n - 2) {
        F.du2[i] = F.du[i + 1];
        F.du[i + 1] = -m * F.du[i + 1];
      }
      F.dg[i] = ndgi;
      F.du[i] = ndui;
      F.dg[i + 1] = ndgi1;
    }
  }
  if (F.dg[n - 1] == 0.0) F.dg[n - 1] = tiny;
  return F;
}

static std::vector<double> tri_rcp(int n, const TriLU& F) {
  std::vector<double> r(n);
  for (int i = 0; i < n; ++i) r[i] = 1.0 / F.dg[i];
  return r;
}

static void tri_solve_d(int n, const TriLU& F, const double* rdg, double* b) {
  for (int i = 0; i + 1 < n; ++i) {
    if (!F.swp[i]) {
      b[i + 1] = b[i + 1] - F.mul[i] * b[i];
    } else {
      double bi = b[i], bi1 = b[i + 1];
      b[i] = bi1;
      b[i + 1] = bi - F.mul[i] * bi1;
    }
  }
  b[n - 1] = b[n - 1] * rdg[n - 1];
  if (n >= 2) b[n - 2] = (b[n - 2] - F.du[n - 2] * b[n - 1]) * rdg[n - 2];
  for (int i = n - 3; i >= 0; --i) {
    double t = b[i] - F.du[i] * b[i + 1];
    t = t - F.du2[i] * b[i + 2];
    b[i] = t * rdg[i];
  }
}

// Substitution with double factors and a double-double right-hand side
// (the reference's double LU with binary128 vectors, src/solver.h:459-478).
static void tri_solve_dd(int n, const TriLU& F, const double* rdg, dd* b) {
  for (int i = 0; i + 1 < n; ++i) {
    if (!F.swp[i]) {
      b[i + 1] = dd_sub(b[i + 1], dd_mul_d(b[i], F.mul[i]));
    } else {
      dd bi = b[i], bi1 = b[i + 1];
      b[i] = bi1;
      b[i + 1] = dd_sub(bi, dd_mul_d(bi1, F.mul[i]));
    }
  }
  b[n - 1] = dd_div_rcp(b[n - 1], F.dg[n - 1], rdg[n - 1]);
  if (n >= 2)
    b[n - 2] = dd_div_rcp(dd_sub(b[n - 2], dd_mul_d(b[n - 1], F.du[n - 2])), F.dg[n - 2], rdg[n - 2]);
  for (int i = n - 3; i >= 0; --i) {
    dd t = dd_sub(b[i], dd_mul_d(b[i + 1], F.du[i]));
    t = dd_sub(t, dd_mul_d(b[i + 2], F.du2[i]));
    b[i] = dd_div_rcp(t, F.dg[i], rdg[i]);
  }
}

// One pass of the leading-minor recurrence of T - lam*I,
//   p_i = (d_i - lam) p_{i-1} - e_{i-1}^2 p_{i-2},
// with its first and second lam-derivatives, rescaled by exact powers of two
// (2^-300 / 2^300) whenever p_i leaves [2^-300, 2^300].  Returns false as
// soon as some p_i is not positive (T - lam*I is not positive definite in
// floating point); otherwise G = f'/f and Hh = (f'/f)^2 - f''/f for
// f = det(T - lam*I).
static bool minor_pass(int n, const double* d, const double* e2, double lam, double& G, double& Hh) {
  const double big = 0x1p300, small = 0x1p-300;
  double pm = 1.0, dpm = 0.0, ddpm = 0.0;
  double p = d[0] - lam, dp = -1.0, ddp = 0.0;
  if (!(p > 0.0)) return false;
  for (int i = 1; i < n; ++i) {
    const double a = d[i] - lam, b = e2[i - 1];
    double pn = a * p - b * pm;
    double dpn = (a * dp - b * dpm) - p;
    double ddpn = (a * ddp - b * ddpm) - 2.0 * dp;
    if (!(pn > 0.0)) return false;
    if (pn > big) {
      pn *= small, dpn *= small, ddpn *= small, p *= small, dp *= small, ddp *= small;
    } else if (pn < small) {
      pn *= big, dpn *= big, ddpn *= big, p *= big, dp *= big, ddp *= big;
    }
    pm = p, dpm = dp, ddpm = ddp;
    p = pn, dp = dpn, ddp = ddpn;
  }
  G = dp / p;
  Hh = G * G - ddp / p;
  return true;
}

// Lowest eigenvalue of tridiag(e, d, e) by Laguerre's iteration from the
// Gershgorin lower bound.  lam always passes minor_pass (a lower bound).  In
// exact arithmetic the iteration never passes the smallest root, so a
// candidate that fails is within rounding of it: the point
// 1e-13 * (|cand| + 1) below it is tried once, and if it passes it is the
// answer; otherwise the failed points bound the bisection that follows.  At
// most 100 iterations.
static double tri_lowest(int n, const double* d, const double* e) {
  if (n == 1) return d[0];
  std::vector<double> e2(n - 1);
  for (int i = 0; i + 1 < n; ++i) e2[i] = e[i] * e[i];
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    if (i > 0) r = r + std::fabs(e[i - 1]);
    if (i < n - 1) r = r + std::fabs(e[i]);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double pad = (hi - lo) * 1e-6 + 1e-200;
  double lam = lo - pad;
  hi = hi + pad;
  const double nn = (double)n;
  double G, Hh;
  if (!minor_pass(n, d, e2.data(), lam, G, Hh)) return lam;
  for (int it = 0; it < 100; ++it) {
    double disc = (nn - 1.0) * (nn * Hh - G * G);
    if (!(disc > 0.0)) disc = 0.0;
    double cand = lam - nn / (G - std::sqrt(disc));
    if (!(cand < hi)) cand = lam + 0.5 * (hi - lam);
    if (!(cand > lam) || !(cand < hi)) break;
    double G2, H2;
    if (minor_pass(n, d, e2.data(), cand, G2, H2)) {
      lam = cand, G = G2, Hh = H2;
      continue;
    }
    hi = cand;
    const double t = cand - 1e-13 * (std::fabs(cand) + 1.0);
    if (!(t > lam)) break;
    if (minor_pass(n, d, e2.data(), t, G2, H2)) {
      lam = t;
      break;
    }
    hi = t;
  }
  return lam;
}

// A: n x n symmetric, row-major, both triangles filled with identical bits.
static EigPair lowest_eigpair(int n, std::vector<double> A) {
  EigPair out;
  out.q.assign(n, dd{0.0, 0.0});
  if (n == 1) {
    out.lambda = A[0];
    out.q[0] = dd{1.0, 0.0};
    return out;
  }
  std::vector<double> d(n), e(n - 1), tau(n, 0.0);
  std::vector<std::vector<double>> vs(n);
  std::vector<double> v(n), p(n), w(n);
  for (int s = 0; s <= n - 3; ++s) {
    int m = n - s - 1;
    double alpha = A[(s + 1) * n + s];
    double xnorm2 = red64(m - 1, [&](int i) {
      double t = A[(s + 2 + i) * n + s];
      return t * t;
    });
    d[s] = A[s * n + s];
    if (xnorm2 == 0.0) {
      tau[s] = 0.0;
      e[s] = alpha;
      continue;
    }
    double beta = -std::copysign(std::sqrt(alpha * alpha + xnorm2), alpha);
    double t = (beta - alpha) / beta;
    double scal = 1.0 / (alpha - beta);
    v[0] = 1.0;
    for (int i = 1; i < m; ++i) v[i] = A[(s + 1 + i) * n + s] * scal;
    tau[s] = t;
    e[s] = beta;
    vs[s].assign(v.begin(), v.begin() + m);
    for (int r = 0; r < m; ++r) {
      const double* row = &A[(s + 1 + r) * n + (s + 1)];
      p[r] = t * dot4(m, [&](int c) { return row[c] * v[c]; });
    }
    double K = (-0.5 * t) * red64(m, [&](int i) { return p[i] * v[i]; });
    for (int r = 0; r < m; ++r) w[r] = p[r] + K * v[r];
    for (int r = 0; r < m; ++r) {
      double* row = &A[(s + 1 + r) * n + (s + 1)];
      for (int c = 0; c < m; ++c) row[c] = row[c] - (v[r] * w[c] + w[r] * v[c]);
    }
  }
  d[n - 2] = A[(n - 2) * n + (n - 2)];
  d[n - 1] = A[(n - 1) * n + (n - 1)];
  e[n - 2] = A[(n - 1) * n + (n - 2)];

  double lam = tri_lowest(n, d.data(), e.data());
  out.lambda = lam;

  // Inverse iteration on T in double: one factorisation, three solves.
  std::vector<double> y(n, 1.0);
  {
    TriLU F1 = tri_factor(n, d.data(), e.data(), lam);
    std::vector<double> r1 = tri_rcp(n, F1);
    for (int it = 0; it < 3; ++it) {
      tri_solve_d(n, F1, r1.data(), y.data());
      double mx = 0.0;
      for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(y[i]));
      for (int i = 0; i < n; ++i) y[i] = y[i] / mx;
    }
  }

  // Extended-precision refinement step (shift lambda - 1e-14, solver.h:439).
  std::vector<dd> q(n);
  for (int i = 0; i < n; ++i) q[i] = dd_make(y[i]);
  {
    TriLU F2 = tri_factor(n, d.data(), e.data(), lam - 1e-14);
    std::vector<double> r2 = tri_rcp(n, F2);
    tri_solve_dd(n, F2, r2.data(), q.data());
  }
  for (int s = n - 3; s >= 0; --s) {
    if (tau[s] == 0.0) continue;
    int m = n - s - 1;
    const std::vector<double>& vv = vs[s];
    dd r = red64_dd(m, [&](int i) { return dd_mul_d(q[s + 1 + i], vv[i]); });
    dd t = dd_mul_d(r, tau[s]);
    for (int i = 0; i < m; ++i) q[s + 1 + i] = dd_sub(q[s + 1 + i], dd_mul_d(t, vv[i]));
  }
  dd nrm2 = red64_dd(n, [&](int i) { return dd_mul(q[i], q[i]); });
  dd nrm = dd_sqrt(nrm2);
  for (int i = 0; i < n; ++i) q[i] = dd_div(q[i], nrm);
  int lead = 0;
  while (lead < n - 1 && dd_is_zero(q[lead])) ++lead;
  if (dd_lt0(q[lead]))
    for (int i = 0; i < n; ++i) q[i] = dd_neg(q[i]);
  out.q = q;
  return out;
}

struct EigStats {
  double max_residual = 0.0;
  double curr_max_abs_evec = 0.0;
  long sum_iterations = 0;
  long sum_inv_iter = 0;
  long curr_inv_iter = 0;
  double sum_diff_inv_iter = 0.0;
  double max_diff_inv_iter = 0.0;
  double curr_diff_inv_iter = 0.0;
};

enum { EIG_OK = 0, EIG_NONNEG = 1 };

// The CDFCI subspace update (src/solver.h:333-609), canonical form.
// Hs: k x k row-major; only Hs[i*k+j] with j >= i is read (the reference's
// dsymv/dsyevr read the 'L' triangle of the column-major view, which is the
// upper triangle of the row-major array).  Returns dx; updates xx and sc.
static int eig_update(int k, const double* Hs, const double* x, const double* z, dd& xx, dd& sc,
                      dd xz, double* dx, EigStats& st) {
  auto H = [&](int r, int c) { return r <= c ? Hs[r * k + c] : Hs[c * k + r]; };
  double xTx = 0.0, xTz = 0.0, xTHx = 0.0;
  for (int i = 0; i < k; ++i) xTx = xTx + x[i] * x[i];
  for (int i = 0; i < k; ++i) xTz = xTz + x[i] * z[i];
  std::vector<double> Hx(k);
  for (int r = 0; r < k; ++r) Hx[r] = dot4(k, [&](int c) { return H(r, c) * x[c]; });
  for (int i = 0; i < k; ++i) xTHx = xTHx + x[i] * Hx[i];
  double yy = dd_add_d(xx, -xTx).hi;
  double y1 = std::sqrt(yy);

  if (std::fabs(yy) > 1e-14) {
    int n = k + 1;
    std::vector<double> A((size_t)n * n);
    dd a00 = dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx);
    A[0] = dd_div_d(a00, yy).hi;
    for (int i = 0; i < k; ++i) {
      double b = (z[i] - Hx[i]) / y1;
      A[(i + 1) * n] = b;
      A[i + 1] = b;
    }
    for (int i = 0; i < k; ++i)
      for (int j = i; j < k; ++j) {
        A[(i + 1) * n + (j + 1)] = Hs[i * k + j];
        A[(j + 1) * n + (i + 1)] = Hs[i * k + j];
      }
    EigPair ep = lowest_eigpair(n, A);
    double lam = ep.lambda;
    std::vector<double> ev(n), dl(n), Ae(n);
    for (int i = 0; i < n; ++i) {
      ev[i] = ep.q[i].hi;
      dl[i] = ep.q[i].lo;
    }
    for (int r = 0; r < n; ++r) Ae[r] = dot4(n, [&](int c) { return A[r * n + c] * ev[c]; });
    double res2 = red64(n, [&](int i) {
      double t = Ae[i] + (-lam) * ev[i];
      return t * t;
    });
    st.max_residual = std::max(st.max_residual, std::sqrt(res2));
    double mev = 0.0;
    for (int i = 1; i < n; ++i) mev = std::max(mev, std::fabs(ev[i]));
    st.curr_max_abs_evec = mev;
    dd t1 = red64_dd(n, [&](int i) { return two_prod(ev[i], Ae[i]); });
    dd t2 = red64_dd(n, [&](int i) { return two_prod(dl[i], Ae[i]); });
    dd rq = dd_add(t1, dd_mul_d(t2, 2.0));
    dd lam_q = dd_neg(rq);
    if (!(lam_q.hi > 0.0)) return EIG_NONNEG;
    dd gam = dd_div_d(ep.q[0], y1);
    dd sc_q = dd_mul(gam, dd_sqrt(lam_q));
    for (int i = 0; i < k; ++i) dx[i] = dd_add_d(dd_div(ep.q[i + 1], gam), -x[i]).hi;
    double sc_d = ev[0] / y1 * std::sqrt(-lam);
    dd diff = dd_add_d(dd_abs(sc_q), -std::fabs(sc_d));
    double ad = std::fabs(diff.hi);
    st.sum_iterations += 1;
    st.sum_inv_iter += 1;
    st.curr_inv_iter = 1;
    st.sum_diff_inv_iter += ad;
    st.max_diff_inv_iter = std::max(st.max_diff_inv_iter, ad);
    st.curr_diff_inv_iter = ad;
    sc = sc_q;
    xx = dd_div(dd_div(lam_q, sc_q), sc_q);
    return EIG_OK;
  }
  // First iteration: only the reference determinant carries weight.
  std::vector<double> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) A[i * k + j] = H(i, j);
  EigPair ep = lowest_eigpair(k, A);
  double lam = ep.lambda;
  if (!(lam < 0.0)) return EIG_NONNEG;
  double r = std::sqrt(-lam);
  for (int i = 0; i < k; ++i) dx[i] = ep.q[i].hi * r - x[i];
  xx = dd_make(-lam);
  sc = dd_make(1.0);
  return EIG_OK;
}

}  // namespace oc
