// Subspace update of the CDFCI iteration (src/solver.h:333-609) and
// update_x (src/wavefunction.h:425-451) in one workgroup of 256 threads.
//
// The (k+1)-dimensional problem [alpha b^T; b H_sub] is solved for its
// lowest eigenpair by
//   1. Householder tridiagonalisation (dsytd2 conventions; the trailing
//      block is updated column-per-lane, which keeps it bitwise symmetric,
//      and the four waves take the columns c = w mod 4),
//   2. Laguerre's iteration on det(T - lambda I) from the Gershgorin lower
//      bound (monotone and cubically convergent for the smallest root),
//      safeguarded by a bracket,
//   3. three inverse-iteration solves on the tridiagonal matrix (double)
//      with one LU factorisation and reciprocal pivots,
//   4. the extended-precision refinement step of the reference
//      (src/solver.h:437-503) as one double-double inverse-iteration solve
//      with shift lambda - 1e-14, back-transformed in double-double.
// Row sums of matrix-vector products use four interleaved partial sums
// (c = j mod 4, ascending, combined as (S0 + S1) + (S2 + S3)); the other
// parallel sums use the 64-lane order (strided partials, then xor
// butterfly), evaluated redundantly by every wave so that all threads hold
// the same bits.  Sequential chains run on thread 0.

#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace mc {

constexpr int kEigLdsN = 90;  // A lives in LDS up to this dimension
constexpr int kNMax = kMaxK + 1;
constexpr int kEigThreads = kWG;  // four waves

template <class F>
__device__ __forceinline__ double red64(int n, F term) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc = acc + term(i);
  for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
  return acc;
}

template <class F>
__device__ __forceinline__ dd red64_dd(int n, F term) {
  const int lane = threadIdx.x & 63;
  dd acc{0.0, 0.0};
  for (int i = lane; i < n; i += 64) acc = dd_add(acc, term(i));
  for (int off = 32; off >= 1; off >>= 1) {
    dd o{__shfl_xor(acc.hi, off, 64), __shfl_xor(acc.lo, off, 64)};
    acc = dd_add(acc, o);
  }
  return acc;
}

// max |y_i| over i < n (exact, so any order gives the same bits).
__device__ __forceinline__ double wave_absmax(int n, const double* y) {
  const int lane = threadIdx.x & 63;
  double m = 0.0;
  for (int i = lane; i < n; i += 64) m = fmax(m, __builtin_fabs(y[i]));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  return m;
}

struct EigLds {
  double x[kMaxK], z[kMaxK], Hx[kMaxK], b0[kMaxK];
  double v[kNMax], p[kNMax], w[kNMax], col[kNMax];
  double d[kNMax], e[kNMax], e2[kNMax], tau[kNMax], y[kNMax];
  double part[4][kNMax];
  double dg[kNMax], du[kNMax], du2[kNMax], mul[kNMax], rdg[kNMax];
  int sw[kNMax];
  dd q[kNMax];
  double ev[kNMax], dl[kNMax], Ae[kNMax];
  double sc[16];
  int iv[8];
};

// Four-way partial row sums: thread (wave w, lane) accumulates the columns
// c = w mod 4 of rows lane, lane + 64, ...; part[w][r] receives S_w(r).
template <class F>
__device__ __forceinline__ void dot4_partials(EigLds& L, int rows, int cols, F term) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = lane; r < rows; r += 64) {
    double acc = 0.0;
#pragma unroll 4
    for (int c = w; c < cols; c += 4) acc = acc + term(r, c);
    L.part[w][r] = acc;
  }
}

__device__ __forceinline__ double dot4_combine(const EigLds& L, int r) {
  return (L.part[0][r] + L.part[1][r]) + (L.part[2][r] + L.part[3][r]);
}

// LU factorisation with partial pivoting of T - shift*I (dgttrf
// conventions), by one thread; factors land in dg/du/du2/mul/sw.  The
// elimination chain is carried in registers.
__device__ __forceinline__ void tri_factor(int n, const double* d, const double* e, double shift, double* dg,
                                           double* du, double* du2, double* mul, int* swp) {
  const double tiny = 1e-290;
  double cur_dg = d[0] - shift;
  double cur_du = n > 1 ? e[0] : 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    const double dl = e[i];
    double next_dg = d[i + 1] - shift;
    double next_du = (i + 2 < n) ? e[i + 1] : 0.0;
    if (__builtin_fabs(cur_dg) >= __builtin_fabs(dl)) {
      if (cur_dg == 0.0) cur_dg = tiny;
      const double m = dl / cur_dg;
      mul[i] = m;
      swp[i] = 0;
      dg[i] = cur_dg;
      du[i] = cur_du;
      if (i + 2 < n) du2[i] = 0.0;
      next_dg = next_dg - m * cur_du;
    } else {
      const double m = cur_dg / dl;
      mul[i] = m;
      swp[i] = 1;
      dg[i] = dl;
      du[i] = next_dg;
      const double ndg1 = cur_du - m * next_dg;
      if (i + 2 < n) {
        du2[i] = next_du;
        next_du = -m * next_du;
      }
      next_dg = ndg1;
    }
    cur_dg = next_dg;
    cur_du = next_du;
  }
  if (cur_dg == 0.0) cur_dg = tiny;
  dg[n - 1] = cur_dg;
}

__device__ __forceinline__ void tri_solve_d(int n, const double* rdg, const double* du, const double* du2,
                                            const double* mul, const int* swp, double* b) {
  double cur = b[0];
  for (int i = 0; i + 1 < n; ++i) {
    const double nb = b[i + 1];
    if (!swp[i]) {
      b[i] = cur;
      cur = nb - mul[i] * cur;
    } else {
      b[i] = nb;
      cur = cur - mul[i] * nb;
    }
  }
  double b1 = cur * rdg[n - 1];
  b[n - 1] = b1;
  if (n < 2) return;
  double b0 = (b[n - 2] - du[n - 2] * b1) * rdg[n - 2];
  b[n - 2] = b0;
  for (int i = n - 3; i >= 0; --i) {
    double t = b[i] - du[i] * b0;
    t = t - du2[i] * b1;
    const double nb = t * rdg[i];
    b[i] = nb;
    b1 = b0;
    b0 = nb;
  }
}

// Double factors, double-double right-hand side (the reference's double LU
// with binary128 vectors, src/solver.h:459-478).
__device__ __forceinline__ void tri_solve_dd(int n, const double* dg, const double* rdg, const double* du,
                                             const double* du2, const double* mul, const int* swp, dd* b) {
  dd cur = b[0];
  for (int i = 0; i + 1 < n; ++i) {
    const dd nb = b[i + 1];
    if (!swp[i]) {
      b[i] = cur;
      cur = dd_sub(nb, dd_mul_d(cur, mul[i]));
    } else {
      b[i] = nb;
      cur = dd_sub(cur, dd_mul_d(nb, mul[i]));
    }
  }
  dd b1 = dd_div_rcp(cur, dg[n - 1], rdg[n - 1]);
  b[n - 1] = b1;
  if (n < 2) return;
  dd b0 = dd_div_rcp(dd_sub(b[n - 2], dd_mul_d(b1, du[n - 2])), dg[n - 2], rdg[n - 2]);
  b[n - 2] = b0;
  for (int i = n - 3; i >= 0; --i) {
    dd t = dd_sub(b[i], dd_mul_d(b0, du[i]));
    t = dd_sub(t, dd_mul_d(b1, du2[i]));
    const dd nb = dd_div_rcp(t, dg[i], rdg[i]);
    b[i] = nb;
    b1 = b0;
    b0 = nb;
  }
}

// Leading-minor recurrence of T - lam*I, p_i = (d_i - lam) p_{i-1} -
// e_{i-1}^2 p_{i-2}, with first and second lam-derivatives, rescaled by
// exact powers of two whenever p_i leaves [2^-300, 2^300] (no division on
// the chain).  False as soon as some p_i is not positive; otherwise
// G = f'/f and Hh = (f'/f)^2 - f''/f for f = det(T - lam*I).
__device__ __forceinline__ bool minor_pass(int n, const double* d, const double* e2, double lam, double& G,
                                           double& Hh) {
  const double big = 0x1p300, small = 0x1p-300;
  double pm = 1.0, dpm = 0.0, ddpm = 0.0;
  double p = d[0] - lam, dp = -1.0, ddp = 0.0;
  if (!(p > 0.0)) return false;
  bool ok = true;
  for (int i = 1; i < n; ++i) {
    const double a = d[i] - lam, b = e2[i - 1];
    double pn = a * p - b * pm;
    double dpn = (a * dp - b * dpm) - p;
    double ddpn = (a * ddp - b * ddpm) - 2.0 * dp;
    if (!(pn > 0.0)) {
      ok = false;
      break;
    }
    if (pn > big) {
      pn *= small, dpn *= small, ddpn *= small, p *= small, dp *= small, ddp *= small;
    } else if (pn < small) {
      pn *= big, dpn *= big, ddpn *= big, p *= big, dp *= big, ddp *= big;
    }
    pm = p, dpm = dp, ddpm = ddp;
    p = pn, dp = dpn, ddp = ddpn;
  }
  if (!ok) return false;
  const double g = dp / p;
  G = g;
  Hh = g * g - ddp / p;
  return true;
}

// Lowest eigenvalue of tridiag(e, d, e): Laguerre from the Gershgorin lower
// bound; lam always passes minor_pass.  A failed candidate is within
// rounding of the root (the exact iteration never passes it): the point
// 1e-13 * (|cand| + 1) below it is tried once and, if it passes, is the
// answer; otherwise the failed points bound a bisection.
__device__ __forceinline__ double tri_lowest(int n, const double* d, const double* e, const double* e2) {
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    if (i > 0) r = r + __builtin_fabs(e[i - 1]);
    if (i < n - 1) r = r + __builtin_fabs(e[i]);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  const double pad = (hi - lo) * 1e-6 + 1e-200;
  double lam = lo - pad;
  hi = hi + pad;
  const double nn = (double)n;
  double G, Hh;
  if (!minor_pass(n, d, e2, lam, G, Hh)) return lam;
  for (int it = 0; it < 100; ++it) {
    double disc = (nn - 1.0) * (nn * Hh - G * G);
    if (!(disc > 0.0)) disc = 0.0;
    double cand = lam - nn / (G - __builtin_sqrt(disc));
    if (!(cand < hi)) cand = lam + 0.5 * (hi - lam);
    if (!(cand > lam) || !(cand < hi)) break;
    double G2, H2;
    if (minor_pass(n, d, e2, cand, G2, H2)) {
      lam = cand, G = G2, Hh = H2;
      continue;
    }
    hi = cand;
    const double t = cand - 1e-13 * (__builtin_fabs(cand) + 1.0);
    if (!(t > lam)) break;
    if (minor_pass(n, d, e2, t, G2, H2)) {
      lam = t;
      break;
    }
    hi = t;
  }
  return lam;
}

#define EIG_STAMP(i) \
  do {                 \
    if (threadIdx.x == 0 && prof) prof[i] = (long long)wall_clock64(); \
  } while (0)

// Lowest eigenpair of the symmetric n x n matrix in A (destroyed).  Leaves
// lambda in L.sc[0] and the normalised dd eigenvector in L.q.
__device__ __forceinline__ void lowest_eigpair(EigLds& L, double* A, int n, long long* prof) {
  const int tid = threadIdx.x;
  const int bs = blockDim.x;
  if (n == 1) {
    if (tid == 0) {
      L.sc[0] = A[0];
      L.q[0] = dd{1.0, 0.0};
    }
    __syncthreads();
    return;
  }
  for (int s = 0; s <= n - 3; ++s) {
    const int m = n - s - 1;
    for (int i = tid; i < m; i += bs) L.col[i] = A[s * n + s + 1 + i];
    __syncthreads();
    const double alpha = L.col[0];
    const double xnorm2 = red64(m - 1, [&](int i) {
      double t = L.col[1 + i];
      return t * t;
    });
    if (tid == 0) L.d[s] = A[s * n + s];
    if (xnorm2 == 0.0) {
      if (tid == 0) {
        L.tau[s] = 0.0;
        L.e[s] = alpha;
      }
      __syncthreads();
      continue;
    }
    const double beta = -__builtin_copysign(__builtin_sqrt(alpha * alpha + xnorm2), alpha);
    const double tt = (beta - alpha) / beta;
    const double scl = 1.0 / (alpha - beta);
    for (int i = tid; i < m; i += bs) {
      double vi = (i == 0) ? 1.0 : L.col[i] * scl;
      L.v[i] = vi;
      A[s * n + s + 1 + i] = vi;
    }
    if (tid == 0) {
      L.tau[s] = tt;
      L.e[s] = beta;
    }
    __syncthreads();
    const double* B = A + (size_t)(s + 1) * n + (s + 1);
    dot4_partials(L, m, m, [&](int r, int c) { return B[(size_t)c * n + r] * L.v[c]; });
    __syncthreads();
    for (int r = tid; r < m; r += bs) L.p[r] = tt * dot4_combine(L, r);
    __syncthreads();
    const double K = (-0.5 * tt) * red64(m, [&](int i) { return L.p[i] * L.v[i]; });
    for (int r = tid; r < m; r += bs) L.w[r] = L.p[r] + K * L.v[r];
    __syncthreads();
    {
      const int w = tid >> 6, lane = tid & 63;
      double* Bw = A + (size_t)(s + 1) * n + (s + 1);
      for (int r = lane; r < m; r += 64) {
        const double vr = L.v[r], wr = L.w[r];
#pragma unroll 4
        for (int c = w; c < m; c += 4) {
          double a = Bw[(size_t)c * n + r];
          Bw[(size_t)c * n + r] = a - (L.v[c] * wr + L.w[c] * vr);
        }
      }
    }
    __syncthreads();
  }
  EIG_STAMP(3);
  if (tid == 0) {
    L.d[n - 2] = A[(n - 2) * n + (n - 2)];
    L.d[n - 1] = A[(n - 1) * n + (n - 1)];
    L.e[n - 2] = A[(n - 1) * n + (n - 2)];
  }
  __syncthreads();
  for (int i = tid; i + 1 < n; i += bs) L.e2[i] = L.e[i] * L.e[i];
  __syncthreads();
  if (tid == 0) {
    const double lam = tri_lowest(n, L.d, L.e, L.e2);
    L.sc[0] = lam;
    tri_factor(n, L.d, L.e, lam, L.dg, L.du, L.du2, L.mul, L.sw);
  }
  __syncthreads();
  EIG_STAMP(4);
  for (int i = tid; i < n; i += bs) {
    L.rdg[i] = 1.0 / L.dg[i];
    L.y[i] = 1.0;
  }
  __syncthreads();
  for (int it = 0; it < 3; ++it) {
    if (tid == 0) tri_solve_d(n, L.rdg, L.du, L.du2, L.mul, L.sw, L.y);
    __syncthreads();
    const double mx = wave_absmax(n, L.y);
    __syncthreads();
    for (int i = tid; i < n; i += bs) L.y[i] = L.y[i] / mx;
    __syncthreads();
  }
  for (int i = tid; i < n; i += bs) L.q[i] = mkdd(L.y[i]);
  if (tid == 0) tri_factor(n, L.d, L.e, L.sc[0] - 1e-14, L.dg, L.du, L.du2, L.mul, L.sw);
  __syncthreads();
  for (int i = tid; i < n; i += bs) L.rdg[i] = 1.0 / L.dg[i];
  __syncthreads();
  if (tid == 0) tri_solve_dd(n, L.dg, L.rdg, L.du, L.du2, L.mul, L.sw, L.q);
  __syncthreads();
  EIG_STAMP(5);
  // Back-transform q <- H_0 ... H_{n-3} q in double-double.
  for (int s = n - 3; s >= 0; --s) {
    const double ts = L.tau[s];
    if (ts == 0.0) continue;
    const int m = n - s - 1;
    const double* vs = A + (size_t)s * n + s + 1;
    const dd r = red64_dd(m, [&](int i) { return dd_mul_d(L.q[s + 1 + i], vs[i]); });
    const dd t = dd_mul_d(r, ts);
    __syncthreads();
    for (int i = tid; i < m; i += bs) L.q[s + 1 + i] = dd_sub(L.q[s + 1 + i], dd_mul_d(t, vs[i]));
    __syncthreads();
  }
  EIG_STAMP(6);
  const dd nrm2 = red64_dd(n, [&](int i) { return dd_mul(L.q[i], L.q[i]); });
  const dd nrm = dd_sqrt(nrm2);
  __syncthreads();
  for (int i = tid; i < n; i += bs) L.q[i] = dd_div(L.q[i], nrm);
  __syncthreads();
  if (tid == 0) {
    int lead = 0;
    while (lead < n - 1 && dd_zero_p(L.q[lead])) ++lead;
    L.iv[1] = dd_neg_p(L.q[lead]) ? 1 : 0;
  }
  __syncthreads();
  if (L.iv[1])
    for (int i = tid; i < n; i += bs) L.q[i] = dd_neg(L.q[i]);
  __syncthreads();
}

// mode 0: solver iteration (reads picked x/z, writes dx, updates the state
// and the table: update_x).  mode 1: test hook (state holds the inputs; no
// table access).
__global__ __launch_bounds__(kWG) void k_eig(DevState* st, int k, int ldh, const double* Hs,
                                             const double* px, const double* pz, double* dx,
                                             double* Aglob, Slot* tab, const uint32_t* pslot, int mode) {
  __shared__ EigLds L;
  __shared__ double Alds[kEigLdsN * kEigLdsN];
  if (st->error) return;
  long long* prof = st->prof;
  EIG_STAMP(0);
  const int tid = threadIdx.x, bs = blockDim.x;
  for (int i = tid; i < k; i += bs) {
    L.x[i] = px[i];
    L.z[i] = pz[i];
  }
  __syncthreads();
  if (tid == 0) {
    double xTx = 0.0;
    for (int i = 0; i < k; ++i) xTx = xTx + L.x[i] * L.x[i];
    dd xx{st->xx_hi, st->xx_lo};
    const double yy = dd_add_d(xx, -xTx).hi;
    L.sc[4] = yy;
    L.sc[5] = __builtin_sqrt(yy);
  }
  __syncthreads();
  const double yy = L.sc[4], y1 = L.sc[5];
  const bool normal = __builtin_fabs(yy) > 1e-14;
  const int n = normal ? k + 1 : k;
  const int o = normal ? 1 : 0;
  double* A = (n <= kEigLdsN) ? Alds : Aglob;
  // H_sub into the trailing block (upper triangle of Hs mirrored).
  for (int idx = tid; idx < k * k; idx += bs) {
    const int r = idx / k, c = idx - r * k;
    A[(size_t)(r + o) * n + (c + o)] = r <= c ? Hs[r * ldh + c] : Hs[c * ldh + r];
  }
  __syncthreads();
  const double* Hb = A + (size_t)o * n + o;
  dot4_partials(L, k, k, [&](int r, int c) { return Hb[(size_t)c * n + r] * L.x[c]; });
  __syncthreads();
  for (int r = tid; r < k; r += bs) L.Hx[r] = dot4_combine(L, r);
  __syncthreads();
  if (normal) {
    if (tid == 0) {
      double xTz = 0.0, xTHx = 0.0;
      for (int i = 0; i < k; ++i) xTz = xTz + L.x[i] * L.z[i];
      for (int i = 0; i < k; ++i) xTHx = xTHx + L.x[i] * L.Hx[i];
      dd xz = fix_to_dd((((i128)st->xz_hi) << 64) | (i128)st->xz_lo);
      dd a00 = dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx);
      const double a = dd_div_d(a00, yy).hi;
      L.sc[6] = a;
      A[0] = a;
    }
    for (int i = tid; i < k; i += bs) {
      const double b = (L.z[i] - L.Hx[i]) / y1;
      L.b0[i] = b;
      A[i + 1] = b;
      A[(size_t)(i + 1) * n] = b;
    }
  }
  __syncthreads();
  const double a00 = L.sc[6];
  // Entries of the original matrix (A itself is overwritten below).
  auto Aorig = [&](int r, int c) -> double {
    if (!normal) return r <= c ? Hs[r * ldh + c] : Hs[c * ldh + r];
    if (r == 0 && c == 0) return a00;
    if (r == 0) return L.b0[c - 1];
    if (c == 0) return L.b0[r - 1];
    return (r - 1) <= (c - 1) ? Hs[(r - 1) * ldh + (c - 1)] : Hs[(c - 1) * ldh + (r - 1)];
  };
  EIG_STAMP(1);
  EIG_STAMP(2);
  // Separate inlined calls keep the LDS case on ds_* instructions.
  if (n <= kEigLdsN)
    lowest_eigpair(L, Alds, n, prof);
  else
    lowest_eigpair(L, Aglob, n, prof);
  EIG_STAMP(7);
  const double lam = L.sc[0];
  if (normal) {
    for (int i = tid; i < n; i += bs) {
      L.ev[i] = L.q[i].hi;
      L.dl[i] = L.q[i].lo;
    }
    __syncthreads();
    dot4_partials(L, n, n, [&](int r, int c) { return Aorig(r, c) * L.ev[c]; });
    __syncthreads();
    for (int r = tid; r < n; r += bs) L.Ae[r] = dot4_combine(L, r);
    __syncthreads();
    const double res2 = red64(n, [&](int i) {
      double t = L.Ae[i] + (-lam) * L.ev[i];
      return t * t;
    });
    const dd t1 = red64_dd(n, [&](int i) { return two_prod(L.ev[i], L.Ae[i]); });
    const dd t2 = red64_dd(n, [&](int i) { return two_prod(L.dl[i], L.Ae[i]); });
    const dd rq = dd_add(t1, dd_mul_d(t2, 2.0));
    const dd lam_q = dd_neg(rq);
    if (!(lam_q.hi > 0.0)) {
      if (tid == 0) st->error = ERR_NUMERIC;
      return;
    }
    const dd gam = dd_div_d(L.q[0], y1);
    for (int i = tid; i < k; i += bs) dx[i] = dd_add_d(dd_div(L.q[i + 1], gam), -L.x[i]).hi;
    if (tid == 0) {
      const dd sc_q = dd_mul(gam, dd_sqrt(lam_q));
      double mev = 0.0;
      for (int i = 1; i < n; ++i) mev = fmax(mev, __builtin_fabs(L.ev[i]));
      st->max_residual = fmax(st->max_residual, __builtin_sqrt(res2));
      st->curr_max_abs_evec = mev;
      const double sc_d = L.ev[0] / y1 * __builtin_sqrt(-lam);
      const dd diff = dd_add_d(dd_abs(sc_q), -__builtin_fabs(sc_d));
      const double ad = __builtin_fabs(diff.hi);
      st->sum_iterations += 1;
      st->sum_inv_iter += 1;
      st->curr_inv_iter = 1;
      st->sum_diff_inv += ad;
      st->max_diff_inv = fmax(st->max_diff_inv, ad);
      st->curr_diff_inv = ad;
      const dd xx = dd_div(dd_div(lam_q, sc_q), sc_q);
      st->sc_hi = sc_q.hi;
      st->sc_lo = sc_q.lo;
      st->xx_hi = xx.hi;
      st->xx_lo = xx.lo;
    }
  } else {
    if (!(lam < 0.0)) {
      if (tid == 0) st->error = ERR_NUMERIC;
      return;
    }
    const double r = __builtin_sqrt(-lam);
    for (int i = tid; i < k; i += bs) dx[i] = L.q[i].hi * r - L.x[i];
    if (tid == 0) {
      st->xx_hi = -lam;
      st->xx_lo = 0.0;
      st->sc_hi = 1.0;
      st->sc_lo = 0.0;
    }
  }
  EIG_STAMP(8);
  if (mode != 0) return;
  __syncthreads();
  // update_x (src/wavefunction.h:425-451): x_t += dx_t; x.z += z_t * dx_t.
  // Picked determinants are distinct, so every thread owns its slots; the
  // x.z increments are exact fixed-point integers (order-free sum).
  __shared__ long long nx_w[kWG / 64];
  __shared__ i128 xz_w[kWG / 64];
  long long nx = 0;
  i128 inc = 0;
  for (int t = tid; t < k; t += bs) {
    Slot* s = &tab[pslot[t]];
    const double xo = s->x;
    if (!(__builtin_fabs(xo) != 0.0)) ++nx;
    s->x = xo + dx[t];
    inc += to_fix(L.z[t] * dx[t]);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    nx += __shfl_xor(nx, off, 64);
    int64_t hi = (int64_t)(inc >> 64);
    uint64_t lo = (uint64_t)inc;
    int64_t ohi = __shfl_xor(hi, off, 64);
    uint64_t olo = (uint64_t)__shfl_xor((long long)lo, off, 64);
    inc += (((i128)ohi) << 64) | (i128)olo;
  }
  if ((tid & 63) == 0) {
    nx_w[tid >> 6] = nx;
    xz_w[tid >> 6] = inc;
  }
  __syncthreads();
  if (tid == 0) {
    i128 xz = (((i128)st->xz_hi) << 64) | (i128)st->xz_lo;
    long long n = 0;
    for (int w = 0; w < (bs >> 6); ++w) {
      xz += xz_w[w];
      n += nx_w[w];
    }
    st->xz_hi = (int64_t)(xz >> 64);
    st->xz_lo = (uint64_t)xz;
    st->x_size += n;
    st->k_cur = k;
  }
  EIG_STAMP(9);
}

void launch_eig(hipStream_t s, DevState* st, int k, int ldh, const double* Hs, const double* px,
                const double* pz, double* dx, double* Aglob, Slot* tab, const uint32_t* pslot, int mode) {
  // The four-way partial sums assume exactly four waves.
  hipLaunchKernelGGL(k_eig, dim3(1), dim3(kEigThreads), 0, s, st, k, ldh, Hs, px, pz, dx, Aglob, tab, pslot, mode);
}

}  // namespace mc
