// Oracle: canonical lowest-eigenpair solver and the CDFCI subspace update.
// TEST INFRASTRUCTURE ONLY (checker for the HIP path, never shipped).
//
// The reference calls LAPACK dsyevr (RANGE='I', IL=IU=1) and then runs one
// step of inverse iteration in __float128 on an LU factorisation
// (src/solver.h:333-609).  LAPACK is an external, unversioned dependency
// (Makefile:25), so this restatement fixes one deterministic algorithm that
// the device solver reproduces operation for operation:
//   1. Lanczos on the (k+1)-dimensional matrix A from the current state
//      (sqrt(yy), x_I) (normalised, plus a fixed 2^-16 pseudo-random
//      perturbation so that no eigenvector is missed by symmetry):
//        w = A q_j - beta_{j-1} q_{j-1};  alpha_j = q_j.w;
//        w = w - alpha_j q_j;  w = w - sum_i (q_i.w) q_i (one classical
//        Gram-Schmidt pass over all previous vectors, each q_i.w a four-way
//        sum over the rows);
//        beta_j = |w|;  q_{j+1} = w / beta_j,
//      every matrix-vector row sum taken as four interleaved partial sums,
//      c = j mod 4, each a chain of fused multiply-adds, combined as
//      (S0 + S1) + (S2 + S3) (dot4) -- sixteen classes combined pairwise
//      (dot16) where the device spreads the rows over several workgroups
//      (k + 1 > kEigLdsN) -- every other dot product in the
//      canonical 64-lane order (red64); the vector updates w - beta q,
//      w - alpha q, w - sum h_i q_i and the Ritz vector sum_i y_i q_i are
//      fused multiply-adds too (std::fma here, the device's fma: both
//      correctly rounded, so they give one result);
//   2. at scheduled steps (from one less than the previous solve's
//      count, then every second step) the lowest eigenvalue theta of the
//      Lanczos matrix T_m by multisection on Sturm counts (128 points per
//      round, each count one division-free leading-minor recurrence; the
//      device runs two points on each lane of one wave), warm-started
//      from the previous check's theta - est and theta, and its
//      eigenvector y by two inverse-iteration solves on one LU
//      factorisation with partial pivoting (dgttrf style; back substitution
//      by reciprocal pivots);
//   3. stop when beta_j |y_m| <= 1e-15 |T_m|_inf |y|, at breakdown
//      (beta_j <= 1e-15 |T_m|_inf), at m = n or at the on-chip capacity
//      lz_capacity(n); the Ritz vector
//      v = sum_i y_i q_i is normalised in double-double;
//   4. the solve is accepted when |A v - rq v| <= 8e-15 |T_m|_inf (rq the
//      Rayleigh quotient); otherwise it restarts from v (at most five
//      times).
// The subspace update then follows src/solver.h:510-538 in double-double:
// the first-order Rayleigh correction rq = v.Av + 2 dv.Av, gamma, scale,
// dx and xx.  The reference's refinement step (one binary128 inverse
// iteration on double LU factors) improves the vector only to the accuracy
// of the double factorisation, which the converged Lanczos vector already
// has (residual ~ 1e-15 |A|).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "oracle_num.h"

namespace oc {

constexpr double kLzTol = 1e-15;
constexpr double kLzAccept = 8e-15;  // |A v - rq v| <= kLzAccept |T|_inf accepts a solve
constexpr double kLzPerturb = 0x1p-16;
constexpr int kLzAttempts = 6;
// The device keeps the Lanczos basis in on-chip memory: with A in 12288
// doubles up to n = 90, in 8192 doubles above (where the rows of A are
// spread over several workgroups that each hold the whole basis); a solve
// stops at that many vectors and restarts.
constexpr int kEigLdsN = 90;
constexpr int kQBudget = 12288;
constexpr int kQBudgetRep = 8192;

static inline int lz_capacity(int n) {
  const int avail = n <= kEigLdsN ? kQBudget - n * n : kQBudgetRep;
  return std::min(n, avail / n);
}

struct EigPair {
  double lambda = 0.0;  // lowest eigenvalue of the Lanczos matrix (multisection)
  std::vector<dd> q;    // normalised Ritz vector, first nonzero entry > 0
  int steps = 0;        // Lanczos vectors used
  double tn = 0.0;      // |T_m|_inf at the stop
  double alpha0 = 0.0;  // first Lanczos coefficient (diagnostics)
};

// Row sum of a(c) b(c) in the canonical four-way order: S_j = the fused
// multiply-add chain over c = j mod 4 in ascending c (from 0.0), combined as
// (S0 + S1) + (S2 + S3).
template <class FA, class FB>
static inline double dot4(int m, FA a, FB b) {
  double S[4];
  for (int j = 0; j < 4; ++j) {
    double acc = 0.0;
    for (int c = j; c < m; c += 4) acc = std::fma(a(c), b(c), acc);
    S[j] = acc;
  }
  return (S[0] + S[1]) + (S[2] + S[3]);
}

// The same with sixteen classes, c = j mod 16, combined pairwise over four
// levels: (((S0 + S1) + (S2 + S3)) + ((S4 + S5) + (S6 + S7))) + (... S15).
template <class FA, class FB>
static inline double dot16(int m, FA a, FB b) {
  double S[16];
  for (int j = 0; j < 16; ++j) {
    double acc = 0.0;
    for (int c = j; c < m; c += 16) acc = std::fma(a(c), b(c), acc);
    S[j] = acc;
  }
  for (int w = 16; w > 1; w >>= 1)
    for (int i = 0; i < w / 2; ++i) S[i] = S[2 * i] + S[2 * i + 1];
  return S[0];
}

// dot4 or dot16 (cls = 4 or 16).
template <class FA, class FB>
static inline double dotc(int cls, int m, FA a, FB b) {
  return cls == 16 ? dot16(m, a, b) : dot4(m, a, b);
}

// Column classes of the subspace solve for k coordinates: the device's
// row-owner form (k + 1 <= kEigLdsN) sums four, the form that spreads the
// rows over workgroups sixteen.
static inline int eig_classes(int k) { return k + 1 > kEigLdsN ? 16 : 4; }

// x.x, x.z and x.Hx of the set-up: ascending sums in the row-owner form, the
// 64-lane order (red64) in the form that spreads the rows over workgroups.
static inline void setup_sums(int cls, int k, const double* x, const double* z, const double* Hx, double& xTx,
                              double& xTz, double& xTHx) {
  if (cls == 16) {
    xTx = red64(k, [&](int i) { return x[i] * x[i]; });
    xTz = red64(k, [&](int i) { return x[i] * z[i]; });
    xTHx = red64(k, [&](int i) { return x[i] * Hx[i]; });
    return;
  }
  xTx = 0.0;
  xTz = 0.0;
  xTHx = 0.0;
  for (int i = 0; i < k; ++i) xTx = xTx + x[i] * x[i];
  for (int i = 0; i < k; ++i) xTz = xTz + x[i] * z[i];
  for (int i = 0; i < k; ++i) xTHx = xTHx + x[i] * Hx[i];
}

static inline bool dd_is_zero(dd a) { return a.hi == 0.0 && a.lo == 0.0; }
static inline dd dd_abs(dd a) { return dd_lt0(a) ? dd_neg(a) : a; }
static inline bool dd_ge(dd a, dd b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }

// Fixed pseudo-random value in [-1, 1) for index i (integer hash, exact).
static inline double lz_hash_unit(int i) {
  uint32_t h = (uint32_t)(i + 1) * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  h *= 0xC2B2AE3Du;
  h ^= h >> 16;
  return (double)(h >> 8) * 0x1p-23 - 1.0;
}

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

// Eigenvalues of tridiag(e, d, e) (order m) below s: sign changes of the
// leading minors p_0 = 1, p_i = (d_i - s) p_{i-1} - e_{i-1}^2 p_{i-2} (a
// zero minor counts as positive).  After every eighth step (i = 8, 16, ...)
// the pair (p_i, p_{i-1}) is rescaled by 2^-300 when max(|p_i|, |p_{i-1}|)
// exceeds 2^300, by 2^300 when it is nonzero and below 2^-300 (exact powers
// of two).
static int sturm_count(int m, const double* d, const double* e2, double s) {
  const double big = 0x1p300, small = 0x1p-300;
  double pm = 1.0, p = d[0] - s;
  int cnt = p < 0.0 ? 1 : 0;
  for (int i = 1; i < m; ++i) {
    const double t = (d[i] - s) * p, u = e2[i - 1] * pm;
    const double pn = t - u;
    cnt += ((pn < 0.0) != (p < 0.0)) ? 1 : 0;
    pm = p;
    p = pn;
    if ((i & 7) == 0) {
      const double a = std::fmax(std::fabs(p), std::fabs(pm));
      const double sc = a > big ? small : ((a < small && a != 0.0) ? big : 1.0);
      p = p * sc;
      pm = pm * sc;
    }
  }
  return cnt;
}

// Lowest eigenvalue of tridiag(e, d, e) by multisection with 128 points per
// round (the device evaluates them two per lane on one wave): the bracket
// [lo, hi) keeps count(lo) = 0 and count(hi) >= 1; a round counts at
// lo + (hi - lo) * ((j + 1) / 129), j = 0..127, and keeps the first cell
// whose right end counts an eigenvalue (the last cell when none does).  The
// start is the Gershgorin interval padded by 1e-6 of its width; when warm
// (a previous check of this solve), a first round counts at `hint` and at
// prev + |prev| 2^-30, which become lo (count 0) and hi (count >= 1) when
// they lie inside the bracket.  Stops when hi - lo <= 2^-50 max(|lo|, |hi|)
// + 2^-1000, when a round does not shrink the bracket, or after 40 rounds;
// returns lo.
static double ms_lowest(int m, const double* d, const double* e, bool warm, double hint, double prev) {
  if (m == 1) return d[0];
  std::vector<double> e2(m, 0.0);
  for (int i = 0; i + 1 < m; ++i) e2[i] = e[i] * e[i];
  double glo = HUGE_VAL, ghi = -HUGE_VAL;
  for (int i = 0; i < m; ++i) {
    double r = 0.0;
    if (i > 0) r = r + std::fabs(e[i - 1]);
    if (i < m - 1) r = r + std::fabs(e[i]);
    glo = std::fmin(glo, d[i] - r);
    ghi = std::fmax(ghi, d[i] + r);
  }
  const double pad = (ghi - glo) * 1e-6 + 1e-200;
  double lo = glo - pad, hi = ghi + pad;
  if (warm) {
    const double hu = prev + std::fabs(prev) * 0x1p-30;
    const int c0 = sturm_count(m, d, e2.data(), hint);
    const int c1 = sturm_count(m, d, e2.data(), hu);
    const double nlo = (hint > lo && hint < hi && c0 == 0) ? hint : lo;
    const double nhi = (hu > lo && hu < hi && c1 >= 1) ? hu : hi;
    if (nlo < nhi) {
      lo = nlo;
      hi = nhi;
    }
  }
  double sv[128];
  for (int round = 0; round < 40; ++round) {
    const double w = hi - lo;
    const double tol = std::fmax(std::fabs(lo), std::fabs(hi)) * 0x1p-50 + 0x1p-1000;
    if (!(w > tol)) break;
    int first = -1;
    for (int j = 0; j < 128; ++j) {
      sv[j] = lo + w * ((double)(j + 1) / 129.0);
      if (first < 0 && sturm_count(m, d, e2.data(), sv[j]) >= 1) first = j;
    }
    double nlo, nhi;
    if (first < 0) {
      nlo = sv[127];
      nhi = hi;
    } else {
      nhi = sv[first];
      nlo = first > 0 ? sv[first - 1] : lo;
    }
    if (!(nhi - nlo < w)) break;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// Eigenvector of tridiag(e, d, e) for eigenvalue lam (lam within rounding
// of the root, so each solve shrinks the other components by ~1e-12): two
// inverse-iteration solves from the all-ones vector on one factorisation,
// each followed by division by the largest magnitude.
static std::vector<double> tri_vector(int m, const double* d, const double* e, double lam) {
  std::vector<double> y(m, 1.0);
  TriLU F = tri_factor(m, d, e, lam);
  std::vector<double> r = tri_rcp(m, F);
  for (int it = 0; it < 2; ++it) {
    tri_solve_d(m, F, r.data(), y.data());
    double mx = 0.0;
    for (int i = 0; i < m; ++i) mx = std::max(mx, std::fabs(y[i]));
    for (int i = 0; i < m; ++i) y[i] = y[i] / mx;
  }
  return y;
}

// Lowest eigenpair of the symmetric n x n matrix A(r, c) by Lanczos from v0
// (unnormalised).  first_check: the smallest number of Lanczos vectors at
// which convergence is tested (then every step).
template <class Aop>
static EigPair lanczos_lowest(int n, Aop A, const std::vector<double>& v0, int first_check, double perturb,
                              int cls) {
  EigPair out;
  std::vector<double> Q((size_t)(n + 1) * n), w(n), h(n), alpha(n), beta(n), y;
  {
    const double s0 = std::sqrt(red64(n, [&](int i) { return v0[i] * v0[i]; }));
    double* q = Q.data();
    for (int i = 0; i < n; ++i) q[i] = (s0 > 0.0 ? v0[i] / s0 : 0.0) + perturb * lz_hash_unit(i);
    const double s1 = std::sqrt(red64(n, [&](int i) { return q[i] * q[i]; }));
    for (int i = 0; i < n; ++i) q[i] = q[i] / s1;
  }
  const int mcap = lz_capacity(n);
  double tn_done = 0.0;  // max row sum of the completed rows of T
  double theta = 0.0, hint = -HUGE_VAL, prev = 0.0, tn = 0.0;
  int m = 0;
  for (int j = 0; j < n; ++j) {
    const double* qj = &Q[(size_t)j * n];
    for (int r = 0; r < n; ++r) w[r] = dotc(cls, n, [&](int c) { return A(r, c); }, [&](int c) { return qj[c]; });
    if (j > 0)
      for (int r = 0; r < n; ++r) w[r] = std::fma(-beta[j - 1], Q[(size_t)(j - 1) * n + r], w[r]);
    alpha[j] = red64(n, [&](int i) { return qj[i] * w[i]; });
    for (int r = 0; r < n; ++r) w[r] = std::fma(-alpha[j], qj[r], w[r]);
    // Full reorthogonalisation: one classical Gram-Schmidt pass against
    // q_0..q_j (coefficients as class sums over the rows, then subtracted
    // in ascending order).
    for (int i = 0; i <= j; ++i)
      h[i] = dotc(cls, n, [&](int r) { return Q[(size_t)i * n + r]; }, [&](int r) { return w[r]; });
    for (int r = 0; r < n; ++r) {
      double acc = w[r];
      for (int i = 0; i <= j; ++i) acc = std::fma(-h[i], Q[(size_t)i * n + r], acc);
      w[r] = acc;
    }
    beta[j] = std::sqrt(red64(n, [&](int i) { return w[i] * w[i]; }));
    m = j + 1;
    if (j > 0) {
      double rj = std::fabs(alpha[j - 1]);
      if (j > 1) rj = rj + std::fabs(beta[j - 2]);
      rj = rj + std::fabs(beta[j - 1]);
      tn_done = std::max(tn_done, rj);
    }
    double last = std::fabs(alpha[j]);
    if (j > 0) last = last + std::fabs(beta[j - 1]);
    tn = std::max(tn_done, last);
    const bool breakdown = !(beta[j] > kLzTol * tn);
    if (m == n || breakdown || m == mcap || (m >= first_check && (m - first_check) % 2 == 0)) {
      theta = ms_lowest(m, alpha.data(), beta.data(), hint > -HUGE_VAL, hint, prev);
      y = tri_vector(m, alpha.data(), beta.data(), theta);
      double ss = 0.0;
      for (int i = 0; i < m; ++i) ss = ss + y[i] * y[i];
      const double rn = std::sqrt(ss);
      const double est = beta[j] * std::fabs(y[m - 1]);
      const bool conv = m == n || breakdown || est <= (kLzTol * tn) * rn;
      if (conv || m == mcap) break;
      // Some eigenvalue of A lies within est / |y| of theta, and the next
      // lowest Ritz value lies in [lambda_min(A), theta].
      hint = theta - est / rn;
      prev = theta;
    }
    double* qn = &Q[(size_t)(j + 1) * n];
    for (int r = 0; r < n; ++r) qn[r] = w[r] / beta[j];
  }
  std::vector<double> v(n);
  for (int r = 0; r < n; ++r) {
    double acc = 0.0;
    for (int i = 0; i < m; ++i) acc = std::fma(y[i], Q[(size_t)i * n + r], acc);
    v[r] = acc;
  }
  const dd nrm = dd_sqrt(red64_dd(n, [&](int i) { return two_prod(v[i], v[i]); }));
  out.q.assign(n, dd{0.0, 0.0});
  for (int i = 0; i < n; ++i) out.q[i] = dd_div(dd{v[i], 0.0}, nrm);
  int lead = 0;
  while (lead < n - 1 && dd_is_zero(out.q[lead])) ++lead;
  if (dd_lt0(out.q[lead]))
    for (int i = 0; i < n; ++i) out.q[i] = dd_neg(out.q[i]);
  out.lambda = theta;
  out.steps = m;
  out.tn = tn;
  out.alpha0 = alpha[0];
  return out;
}

// First convergence test of a solve, from the previous solve's count.
static inline int lz_first_check(int prev) { return std::max(4, prev - 1); }

struct EigStats {
  double max_residual = 0.0;
  double curr_max_abs_evec = 0.0;
  long sum_iterations = 0;
  long sum_inv_iter = 0;
  long curr_inv_iter = 0;
  double sum_diff_inv_iter = 0.0;
  double max_diff_inv_iter = 0.0;
  double curr_diff_inv_iter = 0.0;
  int lz_prev = 0;  // Lanczos vectors used by the previous solve
  long eig_unconverged = 0;  // solves accepted after the last restart without passing the residual test
};

enum { EIG_OK = 0, EIG_NONNEG = 1 };

// One subspace solve with the a-posteriori check: the Ritz vector q (hi
// parts ev, lo parts dl), Ae = A ev, the double-double Rayleigh quotient
// rq = ev.Ae + 2 dl.Ae and res2 = |Ae - rq ev|^2.  A solve whose residual
// exceeds kLzAccept |T|_inf restarts from its Ritz vector, checking from
// 4 vectors on; the next solve then checks from half the first point.
struct EigSolve {
  EigPair ep;
  std::vector<double> ev, dl, Ae;
  dd rq{0.0, 0.0};
  double res2 = 0.0;
  int attempts = 0;
  int steps_first = 0;
};

template <class Aop>
static EigSolve solve_lowest(int n, Aop A, std::vector<double> v0, EigStats& st, int cls) {
  EigSolve S;
  int first = lz_first_check(st.lz_prev);
  const int first0 = first;
  double pert = kLzPerturb;
  for (int attempt = 0; attempt < kLzAttempts; ++attempt) {
    S.ep = lanczos_lowest(n, A, v0, first, pert, cls);
    S.ev.assign(n, 0.0);
    S.dl.assign(n, 0.0);
    S.Ae.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      S.ev[i] = S.ep.q[i].hi;
      S.dl[i] = S.ep.q[i].lo;
    }
    for (int r = 0; r < n; ++r)
      S.Ae[r] = dotc(cls, n, [&](int c) { return A(r, c); }, [&](int c) { return S.ev[c]; });
    dd t1 = red64_dd(n, [&](int i) { return two_prod(S.ev[i], S.Ae[i]); });
    dd t2 = red64_dd(n, [&](int i) { return two_prod(S.dl[i], S.Ae[i]); });
    S.rq = dd_add(t1, dd_mul_d(t2, 2.0));
    const double lam = S.rq.hi;
    S.res2 = red64(n, [&](int i) {
      double t = S.Ae[i] + (-lam) * S.ev[i];
      return t * t;
    });
    S.attempts = attempt + 1;
    if (attempt == 0) S.steps_first = S.ep.steps;
    if (std::sqrt(S.res2) <= kLzAccept * S.ep.tn) break;
    if (attempt + 1 == kLzAttempts) {
      st.eig_unconverged += 1;
      break;
    }
    v0 = S.ev;
    first = 4;
    pert = 0.0;
  }
  st.lz_prev = S.attempts > 1 ? std::max(4, first0 / 2) : S.ep.steps;
  return S;
}

// The CDFCI subspace update (src/solver.h:333-609), canonical form.
// Hs: k x k row-major; only Hs[i*k+j] with j >= i is read (the reference's
// dsymv/dsyevr read the 'L' triangle of the column-major view, which is the
// upper triangle of the row-major array).  Returns dx; updates xx and sc.
static int eig_update(int k, const double* Hs, const double* x, const double* z, dd& xx, dd& sc,
                      dd xz, double* dx, EigStats& st) {
  auto H = [&](int r, int c) { return r <= c ? Hs[r * k + c] : Hs[c * k + r]; };
  const int cls = eig_classes(k);
  std::vector<double> Hx(k);
  for (int r = 0; r < k; ++r) Hx[r] = dotc(cls, k, [&](int c) { return H(r, c); }, [&](int c) { return x[c]; });
  double xTx, xTz, xTHx;
  setup_sums(cls, k, x, z, Hx.data(), xTx, xTz, xTHx);
  double yy = dd_add_d(xx, -xTx).hi;
  double y1 = std::sqrt(yy);

  if (std::fabs(yy) > 1e-14) {
    int n = k + 1;
    dd a00d = dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx);
    const double a00 = dd_div_d(a00d, yy).hi;
    std::vector<double> b(k);
    for (int i = 0; i < k; ++i) b[i] = (z[i] - Hx[i]) / y1;
    auto A = [&](int r, int c) -> double {
      if (r == 0) return c == 0 ? a00 : b[c - 1];
      if (c == 0) return b[r - 1];
      return H(r - 1, c - 1);
    };
    std::vector<double> v0(n);
    v0[0] = y1;
    for (int i = 0; i < k; ++i) v0[i + 1] = x[i];
    EigSolve S = solve_lowest(n, A, v0, st, cls);
    const std::vector<dd>& q = S.ep.q;
    const double lam = S.rq.hi;
    st.max_residual = std::max(st.max_residual, std::sqrt(S.res2));
    double mev = 0.0;
    for (int i = 1; i < n; ++i) mev = std::max(mev, std::fabs(S.ev[i]));
    st.curr_max_abs_evec = mev;
    dd lam_q = dd_neg(S.rq);
    if (!(lam_q.hi > 0.0)) return EIG_NONNEG;
    dd gam = dd_div_d(q[0], y1);
    dd sc_q = dd_mul(gam, dd_sqrt(lam_q));
    for (int i = 0; i < k; ++i) dx[i] = dd_add_d(dd_div(q[i + 1], gam), -x[i]).hi;
    double sc_d = S.ev[0] / y1 * std::sqrt(-lam);
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
  // First iteration: only the reference determinant carries weight; the
  // eigenpair of H_sub itself (src/solver.h:552-598), with the eigenvalue
  // taken as the double-double Rayleigh quotient of the normalised vector.
  // x is the reference determinant alone, which a symmetry can hide from
  // the lowest eigenvector: start from the fixed pseudo-random vector.
  std::vector<double> v0(k, 0.0);
  EigSolve S = solve_lowest(k, H, v0, st, cls);
  if (!(S.rq.hi < 0.0)) return EIG_NONNEG;
  double r = std::sqrt(-S.rq.hi);
  for (int i = 0; i < k; ++i) dx[i] = S.ev[i] * r - x[i];
  xx = dd_neg(S.rq);
  sc = dd_make(1.0);
  return EIG_OK;
}

}  // namespace oc
