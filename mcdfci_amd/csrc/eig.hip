// Subspace update of the CDFCI iteration (src/solver.h:333-609) and
// update_x (src/wavefunction.h:425-451).
//
// The (k+1)-dimensional problem A = [alpha b^T; b H_sub] is solved for its
// lowest eigenpair by Lanczos from the current state (sqrt(yy), x_I),
// normalised and perturbed by a fixed 2^-16 pseudo-random vector:
//   w = A q_j - beta_{j-1} q_{j-1};  alpha_j = q_j.w;  w -= alpha_j q_j;
//   w -= sum_i (q_i.w) q_i;  beta_j = |w|;  q_{j+1} = w / beta_j.
// Row sums of A v and the Gram-Schmidt coefficients are interleaved partial
// sums over column (row) classes: four, c = g mod 4 (ascending fused
// multiply-add chains from 0.0), combined (S0 + S1) + (S2 + S3), in the
// row-owner form, sixteen combined pairwise over four levels in the form
// that spreads the rows over workgroups; dot products are 64 strided
// per-lane partials (rows l, l + 64, ... added to 0.0 in order) reduced by
// an xor butterfly; w - beta q, w - alpha q, the Gram-Schmidt subtraction
// (ascending i) and the Ritz vector are fused multiply-adds.  The oracle
// restates both forms (dot4, dot16):
//  * n <= kEigLdsN (k_eig1, below): A in LDS, one wave per 64 rows, each
//    lane owning one row;
//  * larger n (k_eig): the rows of A spread over ceil(n / kRRows)
//    workgroups (kRRows rows each in LDS, from the symmetric H_sub that
//    k_submatrix stores in both triangles); every product A v is formed by
//    each workgroup for its rows and all-gathered, and every workgroup runs
//    the same solve on eight waves (waves 0-3 hold every row of the step,
//    so the dot products and the next Lanczos vector are known to each of
//    them; they share the Gram-Schmidt update and the Ritz vector by rows
//    through LDS; the others join the Gram-Schmidt coefficients);
//    workgroup 0 writes the results.
// After the alpha step one classical Gram-Schmidt pass against all previous
// Lanczos vectors (held in LDS; q_i.w as four or sixteen interleaved sums
// over the rows, as above) keeps them orthogonal, so a convergence test made
// late costs steps but not accuracy.  The LDS holds lz_capacity(n) vectors
// (12288 doubles shared with A up to n = kEigLdsN, 8192 above); a solve
// reaching that many stops and restarts from its Ritz vector.  At scheduled steps (from
// one less than the previous solve's count, then every second step) wave 0
// finds the lowest eigenvalue of the Lanczos matrix T_m by multisection on
// Sturm counts (two counts per lane, 128 points per round, warm-started from
// the previous check) and its eigenvector by two inverse-iteration solves
// (one wave for m <= 64, thread 0 above); the solve stops when
// beta_j |y_m| <= 1e-15 |T_m|_inf |y|, at breakdown or at m = n.  The Ritz
// vector is normalised in double-double; the solve is accepted when
// |A v - rq v| <= 8e-15 |T_m|_inf (rq the Rayleigh quotient) and restarted
// from v otherwise (at most five times).  The update follows
// src/solver.h:510-538.  The first iteration (only the reference
// determinant carries weight) starts from the pseudo-random vector alone.

#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"
#include "lanes.h"

namespace mc {

constexpr int kEigLdsN = 90;    // A lives in LDS up to this dimension
constexpr int kQBudget = 12288;  // LDS doubles for A (n <= kEigLdsN) and the Lanczos basis
constexpr int kNMax = kMaxK + 1;
constexpr int kRows = (kNMax + 63) / 64;  // rows per lane
constexpr int kEigThreads = 512;          // k_eig: eight waves per workgroup
constexpr double kLzTol = 1e-15;
constexpr double kLzAccept = 8e-15;  // |A v - rq v| <= kLzAccept |T|_inf accepts a solve
constexpr double kLzPerturb = 0x1p-16;
constexpr int kLzAttempts = 6;

// ---------------------------------------------------------------- exchange between workgroups
// Above n = kEigLdsN the rows of A are spread over the workgroups of k_eig
// (kRRows per block, in LDS), every block runs the same solve, and each
// product A v is all-gathered: every block publishes its row sums and reads
// everybody's.  A word crosses workgroups as a relaxed device-scope atomic
// carrying its round's tag in its upper half ((tag << 32) | one half of a
// double); the reader polls each word until its tag arrives, so no fence
// orders the words (a device-scope fence writes back and invalidates the
// XCD's L2).  Exchange buffer (kXchgWords 64-bit words, zeroed when
// allocated): [0] the launch counter (block 0 advances it at the end), then
// two slots of two words per row, slot p for the rounds r with r & 1 = p.
// Round r of the launch with counter c carries tag c * 1024 + r + 2, so a
// word left from an earlier round never matches; a slot is rewritten two
// rounds later only after every block has read it (a block writes round
// r + 2 after reading all of round r + 1, which every other block wrote after
// reading all of round r).
constexpr int kRRows = 12;        // rows of A per workgroup (wave 0: 12 rows x 4 column classes)
constexpr int kRLd = kMaxK + 4;   // LDS row stride, 4 mod 32 doubles: the wave's rows x classes
                                  // spread over all banks
constexpr int kQBudgetRep = 8192; // LDS doubles of the Lanczos basis above n = kEigLdsN
constexpr int kXHead = 64;
constexpr int kXRep = 544;        // 2 * kNMax, padded
constexpr int kXPoll = 1 << 22;   // polls before a word counts as lost
static_assert(kXHead + 2 * kXRep <= kXchgWords, "exchange buffer layout");
static_assert(2 * kNMax <= kXRep, "exchange slots");

__device__ __forceinline__ uint32_t xtag(uint32_t counter, int round) { return counter * 1024u + (uint32_t)round + 2u; }

__device__ __forceinline__ void xput(uint64_t* p, double v, uint32_t tag) {
  const uint64_t b = (uint64_t)__double_as_longlong(v), t = (uint64_t)tag << 32;
  __hip_atomic_store(p, t | (b & 0xffffffffu), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The double at p once both halves carry tag (bounded: bad = 1 when a half
// does not arrive within kXPoll polls).
__device__ __forceinline__ double xget(uint64_t* p, uint32_t tag, int& bad) {
  uint64_t lo, hi;
  int n = 0;
  do {
    lo = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hi = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } while (((uint32_t)(lo >> 32) != tag || (uint32_t)(hi >> 32) != tag) && ++n < kXPoll);
  if (n >= kXPoll) bad = 1;
  return __longlong_as_double((long long)((lo & 0xffffffffu) | (hi << 32)));
}

// Lanczos vectors the LDS holds for dimension n; a solve that reaches this
// many stops and restarts from its Ritz vector.
__device__ __forceinline__ int lz_capacity(int n) {
  const int avail = n <= kEigLdsN ? kQBudget - n * n : kQBudgetRep;
  const int c = avail / n;
  return c < n ? c : n;
}

// Sum over the wave by the xor butterfly at offsets 32, 16, ..., 1 (every
// lane ends with the same value).
__device__ __forceinline__ double bfly_sum(double v) {
  v = v + lane_xor<32>(v);
  v = v + lane_xor<16>(v);
  v = v + lane_xor<8>(v);
  v = v + lane_xor<4>(v);
  v = v + lane_xor<2>(v);
  v = v + lane_xor<1>(v);
  return v;
}

template <int M>
__device__ __forceinline__ dd bfly_step_dd(dd v) {
  return dd_add(v, dd{lane_xor<M>(v.hi), lane_xor<M>(v.lo)});
}

__device__ __forceinline__ dd bfly_sum_dd(dd v) {
  v = bfly_step_dd<32>(v);
  v = bfly_step_dd<16>(v);
  v = bfly_step_dd<8>(v);
  v = bfly_step_dd<4>(v);
  v = bfly_step_dd<2>(v);
  v = bfly_step_dd<1>(v);
  return v;
}

template <class F>
__device__ __forceinline__ double red64(int n, F term) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc = acc + term(i);
  return bfly_sum(acc);
}

template <class F>
__device__ __forceinline__ dd red64_dd(int n, F term) {
  const int lane = threadIdx.x & 63;
  dd acc{0.0, 0.0};
  for (int i = lane; i < n; i += 64) acc = dd_add(acc, term(i));
  return bfly_sum_dd(acc);
}

// Fixed pseudo-random value in [-1, 1) for index i (integer hash, exact).
__device__ __forceinline__ double lz_hash_unit(int i) {
  uint32_t h = (uint32_t)(i + 1) * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  h *= 0xC2B2AE3Du;
  h ^= h >> 16;
  return (double)(h >> 8) * 0x1p-23 - 1.0;
}

struct EigLds {
  double x[kMaxK], z[kMaxK], Hx[kMaxK], b0[kMaxK];
  double v0[kNMax];
  double part[2][4][kNMax];
  double alpha[kNMax], beta[kNMax], e2[kNMax], hbuf[kNMax], wv[kNMax];
  double dg[kNMax], du[kNMax], du2[kNMax], mul[kNMax], rdg[kNMax], y[kNMax];
  int sw[kNMax];
  dd q[kNMax];
  double ev[kNMax], dl[kNMax], Ae[kNMax];
  double red[2][2][64];  // cross-wave partial sums of the row-owner form
  double sc[16];
  int iv[8];
};

// The bordered matrix A = [a00 b^T; b H] (or H itself in the first
// iteration) spread over the workgroups of k_eig: block g holds rows
// r0 .. r0 + nr - 1 in LDS (row stride kRLd).  Every product A v is formed
// by each block for its rows and all-gathered (rep_mv), so every block holds
// every row sum and runs the same solve on the same numbers.
struct MatRep {
  static constexpr bool kSums = true;
  uint64_t* xb;
  uint32_t counter;
  int* round;          // LDS: the next exchange round (the block's threads agree on it)
  int* bad;            // LDS: an exchanged word was lost
  const double* rows;  // LDS: this block's rows
  int r0, nr;
};

// The chain of one of sixteen column classes, r = r0, r0 + 16, ... < n in
// ascending r, loads issued sixteen pairs at a time; and the sum of the
// sixteen class sums held by lanes 16 i + g (g = 0..15, all lanes active),
// combined pairwise over four levels, (((S0 + S1) + (S2 + S3)) + ...): each
// level adds the partner lane's value, and a + b = b + a exactly, so every
// lane of the sixteen gets the same bits.
__device__ __forceinline__ double class_chain16(const double* a, const double* b, int r0, int n) {
  double S = 0.0;
  int r = r0;
  for (; r + 240 < n; r += 256) {
    double x[16], y[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      x[u] = a[r + 16 * u];
      y[u] = b[r + 16 * u];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) S = __builtin_fma(x[u], y[u], S);
  }
  for (; r + 48 < n; r += 64) {
    double x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = a[r + 16 * u];
      y[u] = b[r + 16 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) S = __builtin_fma(x[u], y[u], S);
  }
  for (; r < n; r += 16) S = __builtin_fma(a[r], b[r], S);
  return S;
}

__device__ __forceinline__ double tree16(double S) {
  const double s1 = S + lane_xor<1>(S);
  const double s2 = s1 + lane_xor<2>(s1);
  const double s3 = s2 + lane_xor<4>(s2);
  return s3 + lane_xor<8>(s3);
}

// out[r] = sum_c rows(r, c + sft) v[c], c < nv, for every row r < nout of the
// matrix (sft = 1: the H part of the bordered rows, whose H row r - 1 the
// block's row r holds from column 1; the output index is then r - 1), each
// row in the canonical order of this form: sixteen fused multiply-add chains
// over the column classes c mod 16, combined pairwise (tree16).  Lane 16 i +
// g of waves 0-2 forms class g of the block's row i; the row sums are
// exchanged through the reply slots of the exchange buffer.  Every thread of every block calls
// it with the same arguments; it starts and ends with a barrier.
__device__ void rep_mv(const MatRep& D, int nout, int nv, const double* v, int sft, double* out) {
  const int tid = threadIdx.x, bs = blockDim.x;
  __syncthreads();
  if (*D.bad) {
    // an earlier exchange failed: no more rounds (the solve then breaks
    // down at once and the kernel reports the failure)
    for (int r = tid; r < nout; r += bs) out[r] = 0.0;
    __syncthreads();
    return;
  }
  const int rnd = *D.round;
  const uint32_t tag = xtag(D.counter, rnd);
  uint64_t* rw = D.xb + kXHead + (rnd & 1) * kXRep;
  if (tid < 16 * kRRows) {
    const int i = tid >> 4, g = tid & 15;
    const bool own = i < D.nr && D.r0 + i >= sft;
    double S = 0.0;
    if (own) S = class_chain16(D.rows + i * kRLd + sft, v, g, nv);
    const double sum = tree16(S);
    if (g == 0 && own) xput(rw + 2 * (D.r0 + i - sft), sum, tag);
  }
  int bad = 0;
  for (int r = tid; r < nout; r += bs) out[r] = xget(rw + 2 * r, tag, bad);
  if (bad) *D.bad = 1;
  __syncthreads();
  if (tid == 0) *D.round = rnd + 1;
}

// A condition of the one-thread tridiagonal code as a wave-uniform value
// (lane 0 is the only active lane there), so that it branches on the
// scalar unit instead of masking.
__device__ __forceinline__ bool uni(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }

// LU factorisation with partial pivoting of T - shift*I (dgttrf
// conventions), by one thread; factors land in dg/du/du2/mul/sw.  The
// elimination chain is carried in registers.
__device__ __forceinline__ void tri_factor(int n, const double* d, const double* e, double shift, double* dg,
                                           double* du, double* du2, double* mul, int* swp) {
  // Thread 0 alone; the next entries are loaded one step ahead so that the
  // LDS latency overlaps the elimination chain.
  const double tiny = 1e-290;
  double cur_dg = d[0] - shift;
  double cur_du = n > 1 ? e[0] : 0.0;
  double e_i = n > 1 ? e[0] : 0.0, d_n = n > 1 ? d[1] : 0.0, e_n = n > 2 ? e[1] : 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    const double dl = e_i;
    double next_dg = d_n - shift;
    double next_du = (i + 2 < n) ? e_n : 0.0;
    e_i = e_n;
    if (i + 2 < n) d_n = d[i + 2];
    if (i + 3 < n) e_n = e[i + 2];
    if (uni(__builtin_fabs(cur_dg) >= __builtin_fabs(dl))) {
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
  double nb_n = n > 1 ? b[1] : 0.0, mul_n = n > 1 ? mul[0] : 0.0;
  int sw_n = n > 1 ? swp[0] : 0;
  for (int i = 0; i + 1 < n; ++i) {
    const double nb = nb_n, mi = mul_n;
    const int si = sw_n;
    if (i + 2 < n) {
      nb_n = b[i + 2];
      mul_n = mul[i + 1];
      sw_n = swp[i + 1];
    }
    if (!uni(si != 0)) {
      b[i] = cur;
      cur = nb - mi * cur;
    } else {
      b[i] = nb;
      cur = cur - mi * nb;
    }
  }
  double b1 = cur * rdg[n - 1];
  b[n - 1] = b1;
  if (n < 2) return;
  double b0 = (b[n - 2] - du[n - 2] * b1) * rdg[n - 2];
  b[n - 2] = b0;
  double bi_n = n > 2 ? b[n - 3] : 0.0, du_n = n > 2 ? du[n - 3] : 0.0, du2_n = n > 2 ? du2[n - 3] : 0.0,
         r_n = n > 2 ? rdg[n - 3] : 0.0;
  for (int i = n - 3; i >= 0; --i) {
    const double bi = bi_n, dui = du_n, du2i = du2_n, ri = r_n;
    if (i > 0) {
      bi_n = b[i - 1];
      du_n = du[i - 1];
      du2_n = du2[i - 1];
      r_n = rdg[i - 1];
    }
    double t = bi - dui * b0;
    t = t - du2i * b1;
    const double nb = t * ri;
    b[i] = nb;
    b1 = b0;
    b0 = nb;
  }
}

// Broadcast of lane `l` (wave-uniform) of a double held across the wave.
__device__ __forceinline__ double rl_double(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}

// f(a_i, b_i) for i = 0..n-1 in ascending order, a_i and b_i broadcast by
// readlane from the lane holding them (lane i mod 64 of a[i / 64]).
template <class F>
__device__ __forceinline__ void bcast_each2(const double (&a)[2], const double (&b)[2], int n, F f) {
  const int n0 = n < 64 ? n : 64;
  for (int i = 0; i < n0; ++i) f(rl_double(a[0], i), rl_double(b[0], i));
  for (int i = 64; i < n; ++i) f(rl_double(a[1], i - 64), rl_double(b[1], i - 64));
}

// Eigenvalues of tridiag(e, d, e) below s0 and below s1: the sign changes
// of the leading minors p_0 = 1, p_i = (d_i - s) p_{i-1} - e_{i-1}^2 p_{i-2}
// (a zero minor counts as positive), the two recurrences interleaved.  After
// every eighth step (i = 8, 16, ...) the pair (p_i, p_{i-1}) is rescaled by
// 2^-300 when max(|p_i|, |p_{i-1}|) > 2^300, by 2^300 when it is nonzero
// and below 2^-300 (exact powers of two: signs are unchanged, so the sign of
// the previous minor is carried as a flag).  d and e2 (e^2) are in LDS,
// read as broadcasts eight steps at a time; the steps run in blocks of
// eight ending at the rescaling points.
__device__ __forceinline__ void sturm_count2(int m, const double* d, const double* e2, double s0, double s1,
                                             int& c0, int& c1) {
  const double big = 0x1p300, small = 0x1p-300;
  const double d0 = d[0];
  double pmA = 1.0, pA = d0 - s0, pmB = 1.0, pB = d0 - s1;
  bool negA = pA < 0.0, negB = pB < 0.0;
  int cA = negA ? 1 : 0, cB = negB ? 1 : 0;
  auto step = [&](double di, double ei) {
    const double tA = (di - s0) * pA, uA = ei * pmA;
    const double tB = (di - s1) * pB, uB = ei * pmB;
    const double nA = tA - uA, nB = tB - uB;
    const bool ngA = nA < 0.0, ngB = nB < 0.0;
    cA += ngA != negA ? 1 : 0;
    cB += ngB != negB ? 1 : 0;
    negA = ngA;
    negB = ngB;
    pmA = pA;
    pA = nA;
    pmB = pB;
    pB = nB;
  };
  int i0 = 1;
  for (; i0 + 8 <= m; i0 += 8) {  // steps i0 .. i0 + 7, the last one i = 0 mod 8
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i0 + u];
      ev[u] = e2[i0 - 1 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) step(dv[u], ev[u]);
    const double aA = fmax(__builtin_fabs(pA), __builtin_fabs(pmA));
    const double aB = fmax(__builtin_fabs(pB), __builtin_fabs(pmB));
    const double sA = aA > big ? small : ((aA < small && aA != 0.0) ? big : 1.0);
    const double sB = aB > big ? small : ((aB < small && aB != 0.0) ? big : 1.0);
    pA = pA * sA;
    pmA = pmA * sA;
    pB = pB * sB;
    pmB = pmB * sB;
  }
  if (i0 < m) {  // at most seven steps, none at a rescaling point
    double dv[7], ev[7];
#pragma unroll
    for (int u = 0; u < 7; ++u) {
      const int i = i0 + u < m ? i0 + u : m - 1;
      dv[u] = d[i];
      ev[u] = e2[i - 1];
    }
#pragma unroll
    for (int u = 0; u < 7; ++u)
      if (i0 + u < m) step(dv[u], ev[u]);
  }
  c0 = cA;
  c1 = cB;
}

// Lowest eigenvalue of the Lanczos matrix tridiag(e, d, e) of order m by
// multisection, one wave (every lane active): the bracket [lo, hi) keeps
// count(lo) = 0 and count(hi) >= 1; each round counts at the 128 points
// lo + (hi - lo) * ((j + 1) / 129), j = 0..127 (lane l takes j = l and
// j = l + 64), and the bracket shrinks to the first cell whose right end
// counts an eigenvalue (the last cell when none does).  The start is the
// Gershgorin interval padded by 1e-6 of its width, narrowed when warm (a
// previous check of this solve) by a first round that counts at `hint`
// (becomes lo when it counts 0) and at prev + |prev| 2^-30 (becomes hi when
// it counts an eigenvalue), each only when inside the bracket.  It stops
// when hi - lo <= 2^-50 max(|lo|, |hi|) + 2^-1000, when a round no longer
// shrinks the bracket, or after 40 rounds, and returns lo.  Every lane
// returns the same value.
__device__ double wave_lowest(int m, const double* d, const double* e, const double* e2g, bool warm, double hint,
                              double prev) {
  constexpr int NB = (kNMax + 63) / 64;
  const int lane = threadIdx.x & 63;
  double dv[NB];
  double glo = __builtin_huge_val(), ghi = -__builtin_huge_val();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int i = 64 * b + lane;
    dv[b] = i < m ? d[i] : 0.0;
    if (i < m) {
      double r = 0.0;
      if (i > 0) r = r + __builtin_fabs(e[i - 1]);
      if (i < m - 1) r = r + __builtin_fabs(e[i]);
      glo = fmin(glo, dv[b] - r);
      ghi = fmax(ghi, dv[b] + r);
    }
  }
  if (m == 1) return rl_double(dv[0], 0);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, off, 64));
    ghi = fmax(ghi, __shfl_xor(ghi, off, 64));
  }
  const double pad = (ghi - glo) * 1e-6 + 1e-200;
  double lo = glo - pad, hi = ghi + pad;
  if (warm) {
    const double hu = prev + __builtin_fabs(prev) * 0x1p-30;
    int c0, c1;
    sturm_count2(m, d, e2g, hint, hu, c0, c1);
    const double nlo = (hint > lo && hint < hi && c0 == 0) ? hint : lo;
    const double nhi = (hu > lo && hu < hi && c1 >= 1) ? hu : hi;
    if (nlo < nhi) {
      lo = nlo;
      hi = nhi;
    }
  }
  const double fa = (double)(lane + 1) / 129.0, fb = (double)(lane + 65) / 129.0;
  for (int round = 0; round < 40; ++round) {
    const double w = hi - lo;
    const double tol = fmax(__builtin_fabs(lo), __builtin_fabs(hi)) * 0x1p-50 + 0x1p-1000;
    if (!(w > tol)) break;
    const double sa = lo + w * fa, sb = lo + w * fb;
    int ca, cb;
    sturm_count2(m, d, e2g, sa, sb, ca, cb);
    const unsigned long long ba = __ballot(ca >= 1), bb = __ballot(cb >= 1);
    double nlo, nhi;
    if (ba != 0) {
      const int l = __ffsll((long long)ba) - 1;
      nhi = rl_double(sa, l);
      nlo = l > 0 ? rl_double(sa, l - 1) : lo;
    } else if (bb != 0) {
      const int l = __ffsll((long long)bb) - 1;
      nhi = rl_double(sb, l);
      nlo = l > 0 ? rl_double(sb, l - 1) : rl_double(sa, 63);
    } else {
      nlo = rl_double(sb, 63);
      nhi = hi;
    }
    if (!(nhi - nlo < w)) break;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// Eigenvector of the Lanczos matrix T_m for eigenvalue lam (thread 0; lam
// within rounding of the root): two inverse-iteration solves from the
// all-ones vector on one factorisation, each followed by division by the
// largest magnitude.  Leaves y in L.y.
__device__ void tri_vector(EigLds& L, int m, double lam) {
  tri_factor(m, L.alpha, L.beta, lam, L.dg, L.du, L.du2, L.mul, L.sw);
  for (int i = 0; i < m; ++i) {
    L.rdg[i] = 1.0 / L.dg[i];
    L.y[i] = 1.0;
  }
  for (int it = 0; it < 2; ++it) {
    tri_solve_d(m, L.rdg, L.du, L.du2, L.mul, L.sw, L.y);
    double mx = 0.0;
    for (int i = 0; i < m; ++i) mx = fmax(mx, __builtin_fabs(L.y[i]));
    for (int i = 0; i < m; ++i) L.y[i] = L.y[i] / mx;
  }
}

// v with the entry of lane l (wave-uniform) replaced by x (x is the same
// in every lane).
__device__ __forceinline__ double wl_double(double v, double x, int l) { return (int)(threadIdx.x & 63) == l ? x : v; }

// tri_vector on one wave for m <= 64, with the same factorisation, solves
// and scalings in the same order: the vectors are held one entry per lane
// (entry i in lane i), the serial recurrences read their entries by
// readlane and write results by writelane, and the reciprocals of the
// pivots, the maxima and the divisions, independent per entry, run across
// the lanes.  Leaves y in L.y; ss = sum_i y_i^2 (ascending), ylast = y_{m-1}.
__device__ void tri_vector_w(EigLds& L, int m, double lam, double& ss, double& ylast) {
  const int lane = threadIdx.x & 63;
  const double D = lane < m ? L.alpha[lane] : 0.0;
  const double E = lane + 1 < m ? L.beta[lane] : 0.0;
  double DG = 0.0, DU = 0.0, DU2 = 0.0, MUL = 0.0;
  int SW = 0;
  const double tiny = 1e-290;
  double cur_dg = rl_double(D, 0) - lam;
  double cur_du = m > 1 ? rl_double(E, 0) : 0.0;
  for (int i = 0; i + 1 < m; ++i) {
    const double dl = rl_double(E, i);
    double next_dg = rl_double(D, i + 1) - lam;
    double next_du = i + 2 < m ? rl_double(E, i + 1) : 0.0;
    if (uni(__builtin_fabs(cur_dg) >= __builtin_fabs(dl))) {
      if (cur_dg == 0.0) cur_dg = tiny;
      const double mm = dl / cur_dg;
      MUL = wl_double(MUL, mm, i);
      DG = wl_double(DG, cur_dg, i);
      DU = wl_double(DU, cur_du, i);
      next_dg = next_dg - mm * cur_du;
    } else {
      const double mm = cur_dg / dl;
      MUL = wl_double(MUL, mm, i);
      SW = (int)(threadIdx.x & 63) == i ? 1 : SW;
      DG = wl_double(DG, dl, i);
      DU = wl_double(DU, next_dg, i);
      const double ndg1 = cur_du - mm * next_dg;
      if (i + 2 < m) {
        DU2 = wl_double(DU2, next_du, i);
        next_du = -mm * next_du;
      }
      next_dg = ndg1;
    }
    cur_dg = next_dg;
    cur_du = next_du;
  }
  if (cur_dg == 0.0) cur_dg = tiny;
  DG = wl_double(DG, cur_dg, m - 1);
  const double RDG = lane < m ? 1.0 / DG : 0.0;
  double Y = 1.0;
  for (int it = 0; it < 2; ++it) {
    double cur = rl_double(Y, 0);
    for (int i = 0; i + 1 < m; ++i) {
      const double nb = rl_double(Y, i + 1), mi = rl_double(MUL, i);
      if (!uni(__builtin_amdgcn_readlane(SW, i) != 0)) {
        Y = wl_double(Y, cur, i);
        cur = nb - mi * cur;
      } else {
        Y = wl_double(Y, nb, i);
        cur = cur - mi * nb;
      }
    }
    double b1 = cur * rl_double(RDG, m - 1);
    Y = wl_double(Y, b1, m - 1);
    if (m >= 2) {
      double b0 = (rl_double(Y, m - 2) - rl_double(DU, m - 2) * b1) * rl_double(RDG, m - 2);
      Y = wl_double(Y, b0, m - 2);
      for (int i = m - 3; i >= 0; --i) {
        double t = rl_double(Y, i) - rl_double(DU, i) * b0;
        t = t - rl_double(DU2, i) * b1;
        const double nb = t * rl_double(RDG, i);
        Y = wl_double(Y, nb, i);
        b1 = b0;
        b0 = nb;
      }
    }
    double mx = lane < m ? __builtin_fabs(Y) : 0.0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    Y = Y / mx;
  }
  if (lane < m) L.y[lane] = Y;
  double acc = 0.0;
  for (int i = 0; i < m; ++i) {
    const double yi = rl_double(Y, i);
    acc = acc + yi * yi;
  }
  ss = acc;
  ylast = rl_double(Y, m - 1);
}

// Lowest eigenpair of the n x n matrix A by Lanczos from L.v0 (normalised,
// plus perturb times the fixed pseudo-random vector).  Q (LDS) holds the
// Lanczos vectors, row i at Q + i * n, at most lz_capacity(n) of them.
// Leaves the eigenvalue in L.sc[0], |T_m|_inf in L.sc[1], the normalised
// double-double Ritz vector (first nonzero entry positive) in L.q and the
// number of Lanczos vectors in L.iv[2].
template <class Mat>
__device__ void lanczos_lowest(EigLds& L, const Mat& A, int n, int first_check, double perturb, double* Q) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  // Waves 0-3 hold every row and form the step (the Gram-Schmidt update
  // and the Ritz vector split by rows between them); the others (if any)
  // join the products A q_j and the Gram-Schmidt coefficients only.
  const bool cw = w < 4;
  const int mcap = lz_capacity(n);
  double QC[kRows], QP[kRows], W[kRows];
  if (cw) {
    double p0 = 0.0;
#pragma unroll
    for (int t = 0; t < kRows; ++t) {
      const int r = lane + 64 * t;
      if (r < n) p0 = p0 + L.v0[r] * L.v0[r];
    }
    const double s0 = __builtin_sqrt(bfly_sum(p0));
    double p1 = 0.0;
#pragma unroll
    for (int t = 0; t < kRows; ++t) {
      const int r = lane + 64 * t;
      QP[t] = 0.0;
      QC[t] = 0.0;
      if (r < n) {
        const double q = (s0 > 0.0 ? L.v0[r] / s0 : 0.0) + perturb * lz_hash_unit(r);
        QC[t] = q;
        p1 = p1 + q * q;
      }
    }
    const double s1 = __builtin_sqrt(bfly_sum(p1));
#pragma unroll
    for (int t = 0; t < kRows; ++t) {
      const int r = lane + 64 * t;
      if (r < n) {
        QC[t] = QC[t] / s1;
        if ((t & 3) == w) Q[r] = QC[t];
      }
    }
  }
  __syncthreads();  // q_0 visible to every wave
  double aprev = 0.0, bprev = 0.0, bprev2 = 0.0, tn_done = 0.0, tn = 0.0;
  double hint = -__builtin_huge_val(), prev = 0.0;  // warm start of the next check (every thread)
  int m = 0;
  for (int j = 0; j < n; ++j) {
    rep_mv(A, n, n, Q + (size_t)j * n, 0, L.part[j & 1][0]);
    double alpha = 0.0;
    if (cw) {
      double(*P)[kNMax] = L.part[j & 1];
      double pa = 0.0;
#pragma unroll
      for (int t = 0; t < kRows; ++t) {
        const int r = lane + 64 * t;
        W[t] = 0.0;
        if (r < n) {
          double s = P[0][r];
          if (j > 0) s = __builtin_fma(-bprev, QP[t], s);
          W[t] = s;
          pa = pa + QC[t] * s;
        }
      }
      alpha = bfly_sum(pa);
#pragma unroll
      for (int t = 0; t < kRows; ++t) {
        const int r = lane + 64 * t;
        W[t] = __builtin_fma(-alpha, QC[t], W[t]);
        if (r < n && w == 0) L.wv[r] = W[t];
      }
    }
    __syncthreads();
    // Full reorthogonalisation, one classical Gram-Schmidt pass:
    // h_i = q_i.w as sixteen interleaved sums S_g over the rows r = g mod 16
    // (ascending), combined pairwise (tree16); then waves 0-3 subtract
    // sum_i h_i q_i in ascending i.  Thread 16 i + g forms S_g.
    for (int i0 = 0; i0 <= j; i0 += (int)blockDim.x >> 4) {
      const int i = i0 + (tid >> 4), g = tid & 15;
      double S = 0.0;
      if (i <= j) S = class_chain16(Q + (size_t)i * n, L.wv, g, n);
      const double h = tree16(S);
      if (g == 0 && i <= j) L.hbuf[i] = h;
    }
    __syncthreads();
    // w -= sum_i h_i q_i, each row's chain in ascending i: wave w forms rows
    // lane + 64 t of t = w and t = w + 4 (interleaved; rows at or past n
    // read row 0 and are discarded) and publishes them in L.wv, free since
    // the coefficients were formed; then every row wave reads all rows back.
    constexpr int kOwn = (kRows + 3) / 4;
    if (cw) {
      double acc[kOwn];
      int rc[kOwn];
#pragma unroll
      for (int s = 0; s < kOwn; ++s) {
        const int r = lane + 64 * (w + 4 * s);
        rc[s] = r < n ? r : 0;
        acc[s] = r < n ? L.wv[r] : 0.0;
      }
      int i = 0;
      for (; i + 7 <= j; i += 8) {
        double ha[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) ha[u] = L.hbuf[i + u];
#pragma unroll
        for (int s = 0; s < kOwn; ++s) {
          double qa[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) qa[u] = Q[(size_t)(i + u) * n + rc[s]];
#pragma unroll
          for (int u = 0; u < 8; ++u) acc[s] = __builtin_fma(-ha[u], qa[u], acc[s]);
        }
      }
      for (; i <= j; ++i) {
        const double h = L.hbuf[i];
#pragma unroll
        for (int s = 0; s < kOwn; ++s) acc[s] = __builtin_fma(-h, Q[(size_t)i * n + rc[s]], acc[s]);
      }
#pragma unroll
      for (int s = 0; s < kOwn; ++s) {
        const int r = lane + 64 * (w + 4 * s);
        if (r < n) L.wv[r] = acc[s];
      }
    }
    __syncthreads();
    if (cw) {
      double pb = 0.0;
#pragma unroll
      for (int t = 0; t < kRows; ++t) {
        const int r = lane + 64 * t;
        if (r < n) {
          W[t] = L.wv[r];
          pb = pb + W[t] * W[t];
        }
      }
      const double beta = __builtin_sqrt(bfly_sum(pb));
      if (j > 0) {
        double rj = __builtin_fabs(aprev);
        if (j > 1) rj = rj + __builtin_fabs(bprev2);
        rj = rj + __builtin_fabs(bprev);
        tn_done = fmax(tn_done, rj);
      }
      double last = __builtin_fabs(alpha);
      if (j > 0) last = last + __builtin_fabs(bprev);
      tn = fmax(tn_done, last);
      if (tid == 0) {
        const int mm = j + 1;
        const bool breakdown = !(beta > kLzTol * tn);
        L.alpha[j] = alpha;
        L.beta[j] = beta;
        L.e2[j] = beta * beta;
        L.iv[3] = (mm == n || breakdown || mm == mcap || (mm >= first_check && ((mm - first_check) & 1) == 0)) ? 1 : 0;
        L.iv[4] = breakdown ? 1 : 0;
      }
      aprev = alpha;
      bprev2 = bprev;
      bprev = beta;
    }
    m = j + 1;
    __syncthreads();  // this step's alpha, beta and the check decision
    if (L.iv[3]) {
      double theta = 0.0;
      if (w == 0) theta = wave_lowest(m, L.alpha, L.beta, L.e2, hint > -__builtin_huge_val(), hint, prev);
      // eigenvector of T_m: wave 0 for m <= 64, thread 0 above (the same
      // arithmetic)
      double ss = 0.0, ylast = 0.0;
      if (m <= 64) {
        if (w == 0) tri_vector_w(L, m, theta, ss, ylast);
      } else if (tid == 0) {
        tri_vector(L, m, theta);
        for (int i = 0; i < m; ++i) ss = ss + L.y[i] * L.y[i];
        ylast = L.y[m - 1];
      }
      if (tid == 0) {
        const double beta = L.beta[j];
        const bool breakdown = L.iv[4] != 0;
        const double rn = __builtin_sqrt(ss);
        const double est = beta * __builtin_fabs(ylast);
        const bool cv = m == n || breakdown || est <= (kLzTol * tn) * rn;
        L.iv[0] = (cv || m == mcap) ? 1 : 0;
        L.sc[0] = theta;
        // Some eigenvalue of A lies within est / |y| of theta, and the next
        // lowest Ritz value lies in [lambda_min(A), theta].
        L.sc[2] = theta - est / rn;
        L.sc[3] = theta;
      }
      __syncthreads();
      hint = L.sc[2];
      prev = L.sc[3];
      if (L.iv[0]) break;
    }
    if (cw) {
      const double beta = bprev;
#pragma unroll
      for (int t = 0; t < kRows; ++t) {
        const int r = lane + 64 * t;
        QP[t] = QC[t];
        if (r < n) {
          const double q = W[t] / beta;
          QC[t] = q;
          if ((t & 3) == w) Q[(size_t)(j + 1) * n + r] = q;
        }
      }
    }
    __syncthreads();  // q_{j+1} visible to every wave
  }
  // Ritz vector v = sum_i y_i q_i, each row's chain in ascending i, the rows
  // spread over waves 0-3 as in the Gram-Schmidt update (through L.wv);
  // wave 0 normalises it in double-double.
  constexpr int kOwnR = (kRows + 3) / 4;
  if (cw) {
    double V[kOwnR];
    int rc[kOwnR];
#pragma unroll
    for (int s = 0; s < kOwnR; ++s) {
      const int r = lane + 64 * (w + 4 * s);
      V[s] = 0.0;
      rc[s] = r < n ? r : 0;
    }
    int i = 0;
    for (; i + 7 < m; i += 8) {
      double ya[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) ya[u] = L.y[i + u];
#pragma unroll
      for (int s = 0; s < kOwnR; ++s) {
        double qa[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) qa[u] = Q[(size_t)(i + u) * n + rc[s]];
#pragma unroll
        for (int u = 0; u < 8; ++u) V[s] = __builtin_fma(ya[u], qa[u], V[s]);
      }
    }
    for (; i < m; ++i) {
      const double y = L.y[i];
#pragma unroll
      for (int s = 0; s < kOwnR; ++s) V[s] = __builtin_fma(y, Q[(size_t)i * n + rc[s]], V[s]);
    }
#pragma unroll
    for (int s = 0; s < kOwnR; ++s) {
      const int r = lane + 64 * (w + 4 * s);
      if (r < n) L.wv[r] = V[s];
    }
  }
  __syncthreads();
  if (w == 0) {
    double V[kRows];
    dd pn{0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kRows; ++t) {
      const int r = lane + 64 * t;
      V[t] = 0.0;
      if (r < n) {
        V[t] = L.wv[r];
        pn = dd_add(pn, two_prod(V[t], V[t]));
      }
    }
    const dd nrm = dd_sqrt(bfly_sum_dd(pn));
#pragma unroll
    for (int t = 0; t < kRows; ++t) {
      const int r = lane + 64 * t;
      if (r < n) L.q[r] = dd_div(mkdd(V[t]), nrm);
    }
  }
  __syncthreads();
  if (tid == 0) {
    int lead = 0;
    while (lead < n - 1 && dd_zero_p(L.q[lead])) ++lead;
    L.iv[1] = dd_neg_p(L.q[lead]) ? 1 : 0;
    L.iv[2] = m;
    L.sc[1] = tn;
  }
  __syncthreads();
  if (L.iv[1])
    for (int i = tid; i < n; i += blockDim.x) L.q[i] = dd_neg(L.q[i]);
  __syncthreads();
}

// ---------------------------------------------------------------- row-owner form
// For n <= kEigLdsN the same solve runs with one row per lane: wave w of
// nw = ceil(n / 64) (one or two) owns rows 64 w + lane and keeps, for its
// row, the four interleaved partial sums S_g (columns c = g mod 4,
// ascending) that the class form spreads over its waves, so every row
// sum, dot product and Gram-Schmidt coefficient is formed in the same order
// and the results are bit-identical.  A dot product adds, for lane l, the
// term of row l to 0.0 and then the term of row l + 64, and reduces by the
// xor butterfly: the class form's strided order.  One wave issues at
// most one instruction every fourth cycle, so a second wave for rows 64..89
// runs on another SIMD instead of lengthening the first wave's stream.
constexpr int kR1 = 2;  // entries per lane of the setup vectors (k < 128)

// Row r of A v: (S0 + S1) + (S2 + S3), S_g the ascending sum over c = g
// mod 4 of A(r, c) v_c.  A is symmetric (k_submatrix writes both triangles
// from one value, the border is copied), so row r is read along its
// contiguous storage a[r * lda + c], sixteen columns per batch; v in LDS.
__device__ __forceinline__ double matvec_row(const double* a, int lda, int n, int r, const double* v) {
  const double* rowp = a + (r < n ? r : n - 1) * lda;
  double S[4] = {0.0, 0.0, 0.0, 0.0};
  int c0 = 0;
  for (; c0 + 16 <= n; c0 += 16) {
    double vv[16], aa[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      vv[u] = v[c0 + u];
      aa[u] = rowp[c0 + u];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) S[u & 3] = __builtin_fma(aa[u], vv[u], S[u & 3]);
  }
  if (c0 + 8 <= n) {
    double vv[8], aa[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      vv[u] = v[c0 + u];
      aa[u] = rowp[c0 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) S[u & 3] = __builtin_fma(aa[u], vv[u], S[u & 3]);
    c0 += 8;
  }
  if (c0 + 4 <= n) {
    double vv[4], aa[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      vv[u] = v[c0 + u];
      aa[u] = rowp[c0 + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) S[u] = __builtin_fma(aa[u], vv[u], S[u]);
    c0 += 4;
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (c0 + u < n) S[u] = __builtin_fma(rowp[c0 + u], v[c0 + u], S[u]);
  return (S[0] + S[1]) + (S[2] + S[3]);
}

// Sum over all rows of a per-row term p (0.0 where the lane's row does not
// exist): (0.0 + p(row l)) + p(row l + 64) per lane l, then the butterfly;
// every wave returns the same value.  buf alternates between consecutive
// calls so that a barrier always separates reuse of one buffer.
__device__ __forceinline__ double rows_sum(EigLds& L, double p, int nw, int buf) {
  double s = 0.0 + p;
  if (nw == 2) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    L.red[buf][w][lane] = p;
    __syncthreads();
    s = (0.0 + L.red[buf][0][lane]) + L.red[buf][1][lane];
  }
  return bfly_sum(s);
}

// lanczos_lowest for A held in LDS (n x n, lda n), row-owner form.
__device__ void lanczos1(EigLds& L, const double* A, int n, int first_check, double perturb, double* Q) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6, bs = blockDim.x;
  const int r = lane + 64 * w;
  const bool own = r < n;
  const bool wave_rows = 64 * w < n;  // wave-uniform: this wave owns a row
  const int mcap = lz_capacity(n);
  double QC = 0.0, QP = 0.0, W = 0.0;
  {
    const double v = own ? L.v0[r] : 0.0;
    const double s0 = __builtin_sqrt(rows_sum(L, own ? v * v : 0.0, nw, 0));
    double q = 0.0;
    if (own) q = (s0 > 0.0 ? v / s0 : 0.0) + perturb * lz_hash_unit(r);
    const double s1 = __builtin_sqrt(rows_sum(L, own ? q * q : 0.0, nw, 1));
    if (own) {
      QC = q / s1;
      Q[r] = QC;
    }
  }
  __syncthreads();
  double aprev = 0.0, bprev = 0.0, bprev2 = 0.0, tn_done = 0.0, tn = 0.0;
  double hint = -__builtin_huge_val(), prev = 0.0;
  int m = 0;
  for (int j = 0; j < n; ++j) {
    const double av = matvec_row(A, n, n, r, Q + j * n);
    double sv = 0.0;
    if (own) {
      sv = av;
      if (j > 0) sv = __builtin_fma(-bprev, QP, sv);
    }
    const double alpha = rows_sum(L, own ? QC * sv : 0.0, nw, 0);
    W = own ? __builtin_fma(-alpha, QC, sv) : 0.0;
    if (own) L.wv[r] = W;
    __syncthreads();
    // Gram-Schmidt coefficients h_i = q_i.w: lane 4i' + g forms S_g over
    // the rows r = g mod 4 (16 coefficients per pass, passes alternating
    // between the waves), combined as (S0 + S1) + (S2 + S3).
    {
      const int g = lane & 3;
      for (int i0 = 16 * w; i0 <= j; i0 += 16 * nw) {
        const int i = i0 + (lane >> 2);
        const double* qi = Q + (i <= j ? i : j) * n + g;
        const double* wp = L.wv + g;
        double S = 0.0;
        int b = 0;
        for (; b + 32 <= n; b += 32) {
          double qa[8], wa[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            qa[u] = qi[b + 4 * u];
            wa[u] = wp[b + 4 * u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) S = __builtin_fma(qa[u], wa[u], S);
        }
        for (; b < n; b += 4)
          if (b + g < n) S = __builtin_fma(qi[b], wp[b], S);
        const double s01 = S + lane_xor<1>(S);
        const double h = s01 + lane_xor<2>(s01);
        if (g == 0 && i <= j) L.hbuf[i] = h;
      }
    }
    __syncthreads();
    // w -= sum_i h_i q_i, ascending i, eight terms per batch
    if (wave_rows) {
      const double* qp = Q + (own ? r : n - 1);
      double acc = W;
      int i = 0;
      for (; i + 8 <= j + 1; i += 8) {
        double qa[8], ha[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          qa[u] = qp[(i + u) * n];
          ha[u] = L.hbuf[i + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = __builtin_fma(-ha[u], qa[u], acc);
      }
      for (; i <= j; ++i) acc = __builtin_fma(-L.hbuf[i], qp[i * n], acc);
      if (own) W = acc;
    }
    const double beta = __builtin_sqrt(rows_sum(L, own ? W * W : 0.0, nw, 1));
    if (j > 0) {
      double rj = __builtin_fabs(aprev);
      if (j > 1) rj = rj + __builtin_fabs(bprev2);
      rj = rj + __builtin_fabs(bprev);
      tn_done = fmax(tn_done, rj);
    }
    double last = __builtin_fabs(alpha);
    if (j > 0) last = last + __builtin_fabs(bprev);
    tn = fmax(tn_done, last);
    m = j + 1;
    if (tid == 0) {
      L.alpha[j] = alpha;
      L.beta[j] = beta;
      L.e2[j] = beta * beta;
    }
    const bool breakdown = !(beta > kLzTol * tn);
    if (m == n || breakdown || m == mcap || (m >= first_check && ((m - first_check) & 1) == 0)) {
      __syncthreads();
      const double theta = wave_lowest(m, L.alpha, L.beta, L.e2, hint > -__builtin_huge_val(), hint, prev);
      // eigenvector of T_m and the convergence test: wave 0 for m <= 64,
      // thread 0 above (the same arithmetic)
      double ss = 0.0, ylast = 0.0;
      if (m <= 64) {
        if (w == 0) tri_vector_w(L, m, theta, ss, ylast);
      } else if (tid == 0) {
        tri_vector(L, m, theta);
        for (int i = 0; i < m; ++i) ss = ss + L.y[i] * L.y[i];
        ylast = L.y[m - 1];
      }
      if (tid == 0) {
        const double rn = __builtin_sqrt(ss);
        const double est = beta * __builtin_fabs(ylast);
        const bool cv = m == n || breakdown || est <= (kLzTol * tn) * rn;
        L.iv[0] = (cv || m == mcap) ? 1 : 0;
        L.sc[0] = theta;
        L.sc[2] = theta - est / rn;
        L.sc[3] = theta;
      }
      __syncthreads();
      hint = L.sc[2];
      prev = L.sc[3];
      if (L.iv[0]) break;
    }
    QP = QC;
    if (own) {
      QC = W / beta;
      Q[(j + 1) * n + r] = QC;
    }
    __syncthreads();
    aprev = alpha;
    bprev2 = bprev;
    bprev = beta;
  }
  // Ritz vector v = sum_i y_i q_i, normalised in double-double: per lane l,
  // dd(0) + (row l)^2 + (row l + 64)^2, then the butterfly.
  double V = 0.0;
  dd tp{0.0, 0.0};
  if (wave_rows) {
    const double* qp = Q + (own ? r : n - 1);
    double acc = 0.0;
    int i = 0;
    for (; i + 8 <= m; i += 8) {
      double qa[8], ya[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        qa[u] = qp[(i + u) * n];
        ya[u] = L.y[i + u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = __builtin_fma(ya[u], qa[u], acc);
    }
    for (; i < m; ++i) acc = __builtin_fma(L.y[i], qp[i * n], acc);
    if (own) {
      V = acc;
      tp = two_prod(acc, acc);
    }
  }
  dd pn{0.0, 0.0};
  if (nw == 2) {
    L.red[0][w][lane] = tp.hi;
    L.red[1][w][lane] = tp.lo;
    __syncthreads();
    if (lane < n) pn = dd_add(pn, dd{L.red[0][0][lane], L.red[1][0][lane]});
    if (lane + 64 < n) pn = dd_add(pn, dd{L.red[0][1][lane], L.red[1][1][lane]});
  } else if (own) {
    pn = dd_add(pn, tp);
  }
  const dd nrm = dd_sqrt(bfly_sum_dd(pn));
  if (own) L.q[r] = dd_div(mkdd(V), nrm);
  __syncthreads();
  if (tid == 0) {
    int lead = 0;
    while (lead < n - 1 && dd_zero_p(L.q[lead])) ++lead;
    L.iv[1] = dd_neg_p(L.q[lead]) ? 1 : 0;
    L.iv[2] = m;
    L.sc[1] = tn;
  }
  __syncthreads();
  if (L.iv[1])
    for (int i = tid; i < n; i += bs) L.q[i] = dd_neg(L.q[i]);
  __syncthreads();
}

// The update of src/solver.h:510-538 from the accepted Ritz vector (L.q,
// L.ev) and its double-double Rayleigh quotient rq, then update_x: shared by
// the eight-wave and the row-owner solvers (block size bs, any multiple of
// 64 up to kEigThreads).
__device__ __forceinline__ void eig_finish(DevState* st, EigLds& L, int k, int n, bool normal, double y1, dd rq,
                                           double res2, double* dx, double* tabx, int slot_stride,
                                           const uint32_t* pslot, int mode) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const double lam = rq.hi;
  if (normal) {
    const dd lam_q = dd_neg(rq);
    if (!(lam_q.hi > 0.0)) {
      if (tid == 0) st->error = ERR_NUMERIC;
      return;
    }
    const dd gam = dd_div_d(L.q[0], y1);
    for (int i = tid; i < k; i += bs) dx[i] = dd_add_d(dd_div(L.q[i + 1], gam), -L.x[i]).hi;
    // max |ev_i|, i >= 1, on wave 0 (a maximum does not depend on the order)
    double mev = 0.0;
    if (tid < 64) {
      for (int i = 1 + tid; i < n; i += 64) mev = fmax(mev, __builtin_fabs(L.ev[i]));
      for (int off = 32; off >= 1; off >>= 1) mev = fmax(mev, __shfl_xor(mev, off, 64));
    }
    if (tid == 0) {
      const dd sc_q = dd_mul(gam, dd_sqrt(lam_q));
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
    if (!(rq.hi < 0.0)) {
      if (tid == 0) st->error = ERR_NUMERIC;
      return;
    }
    const double r = __builtin_sqrt(-rq.hi);
    for (int i = tid; i < k; i += bs) dx[i] = L.ev[i] * r - L.x[i];
    if (tid == 0) {
      st->xx_hi = -rq.hi;
      st->xx_lo = -rq.lo;
      st->sc_hi = 1.0;
      st->sc_lo = 0.0;
    }
  }
  if (mode != 0) return;
  __syncthreads();
  // update_x (src/wavefunction.h:425-451): x_t += dx_t; x.z += z_t * dx_t.
  // Picked determinants are distinct, so every thread owns its slots; the
  // x.z increments are exact fixed-point integers (order-free sum).
  __shared__ long long nx_w[kEigThreads / 64];
  __shared__ i128 xz_w[kEigThreads / 64];
  long long nx = 0;
  i128 inc = 0;
  for (int t = tid; t < k; t += bs) {
    double* xs = tabx + (size_t)pslot[t] * slot_stride;
    const double xo = *xs;
    if (!(__builtin_fabs(xo) != 0.0)) ++nx;
    *xs = xo + dx[t];
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
}

// mode 0: solver iteration (reads picked x/z, writes dx, updates the state
// and the table: update_x through tabx, the x of slot 0, and slot_stride,
// the slot size in doubles).  mode 1: test hook (state holds the inputs; no
// table access).  Hs holds both triangles of H_sub.
__global__ __launch_bounds__(kEigThreads) void k_eig(DevState* st, int k, int ldh, const double* Hs,
                                             const double* px, const double* pz, double* dx, double* tabx,
                                             int slot_stride, const uint32_t* pslot, int mode, uint64_t* xb) {
  __shared__ EigLds L;
  __shared__ double Qlds[kQBudgetRep];        // the Lanczos basis
  __shared__ double Arows[kRRows * kRLd];     // this block's rows of A
  if (st->error) return;
  const int tid = threadIdx.x, bs = blockDim.x;
  const bool lead = blockIdx.x == 0;  // writes the results
  const uint32_t xcount = (uint32_t)__hip_atomic_load(xb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (tid == 0) {
    L.iv[6] = 0;  // exchange round
    L.iv[7] = 0;  // a word was lost
  }
  for (int i = tid; i < k; i += bs) {
    L.x[i] = px[i];
    L.z[i] = pz[i];
  }
  __syncthreads();
  // x.x on wave 0 and x.z on wave 1, in the 64-lane order (red64)
  const int wave = tid >> 6;
  if (wave == 0) {
    const double xTx = red64(k, [&](int i) { return L.x[i] * L.x[i]; });
    if (tid == 0) {
      dd xx{st->xx_hi, st->xx_lo};
      const double yy = dd_add_d(xx, -xTx).hi;
      L.sc[4] = yy;
      L.sc[5] = __builtin_sqrt(yy);
    }
  } else if (wave == 1) {
    const double xTz = red64(k, [&](int i) { return L.x[i] * L.z[i]; });
    if (tid == 64) L.sc[7] = xTz;
  }
  __syncthreads();
  const double yy = L.sc[4], y1 = L.sc[5];
  const bool normal = __builtin_fabs(yy) > 1e-14;
  const int n = normal ? k + 1 : k;
  const int o = normal ? 1 : 0;
  // This block's rows of A: H_sub rows (its columns; k_submatrix stores both
  // triangles) from column o; the border column is filled after H x.
  const int r0 = blockIdx.x * kRRows;
  const int nr = n - r0 < 0 ? 0 : (n - r0 < kRRows ? n - r0 : kRRows);
  for (int idx = tid; idx < nr * k; idx += bs) {
    const int i = idx / k, c = idx - i * k, r = r0 + i;
    if (r >= o) Arows[i * kRLd + c + o] = Hs[(size_t)(r - o) * ldh + c];
  }
  const MatRep D{xb, xcount, &L.iv[6], &L.iv[7], Arows, r0, nr};
  if (normal) {
    // H x, all-gathered (the H part of the rows)
    rep_mv(D, k, k, L.x, 1, L.Hx);
    if (wave == 0) {
      const double xTz = L.sc[7];
      const double xTHx = red64(k, [&](int i) { return L.x[i] * L.Hx[i]; });
      if (tid == 0) {
        dd xz = fix_to_dd((((i128)st->xz_hi) << 64) | (i128)st->xz_lo);
        dd a00 = dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx);
        L.sc[6] = dd_div_d(a00, yy).hi;
        L.v0[0] = y1;
      }
    }
    for (int i = tid; i < k; i += bs) {
      L.b0[i] = (L.z[i] - L.Hx[i]) / y1;
      L.v0[i + 1] = L.x[i];
    }
    __syncthreads();
    for (int i = tid; i < nr; i += bs) {
      const int r = r0 + i;
      Arows[i * kRLd] = r == 0 ? L.sc[6] : L.b0[r - 1];
    }
    if (r0 == 0)
      for (int c = 1 + tid; c < n; c += bs) Arows[c] = L.b0[c - 1];
  } else {
    // x is the reference determinant alone, which a symmetry can hide from
    // the lowest eigenvector: start from the fixed pseudo-random vector.
    for (int i = tid; i < k; i += bs) L.v0[i] = 0.0;
  }
  __syncthreads();
  const int first0 = st->lz_prev - 1 > 4 ? st->lz_prev - 1 : 4;
  // Solve, then accept when |A v - rq v| <= kLzAccept |T|_inf; otherwise
  // restart from the Ritz vector (checks from 4 vectors on).
  int first = first0, attempts = 0;
  double perturb = kLzPerturb;
  dd rq{0.0, 0.0};
  double res2 = 0.0;
  for (;;) {
    lanczos_lowest(L, D, n, first, perturb, Qlds);
    for (int i = tid; i < n; i += bs) {
      L.ev[i] = L.q[i].hi;
      L.dl[i] = L.q[i].lo;
    }
    rep_mv(D, n, n, L.ev, 0, L.Ae);
    const dd t1 = red64_dd(n, [&](int i) { return two_prod(L.ev[i], L.Ae[i]); });
    const dd t2 = red64_dd(n, [&](int i) { return two_prod(L.dl[i], L.Ae[i]); });
    rq = dd_add(t1, dd_mul_d(t2, 2.0));
    const double lrq = rq.hi;
    res2 = red64(n, [&](int i) {
      double t = L.Ae[i] + (-lrq) * L.ev[i];
      return t * t;
    });
    ++attempts;
    if (__builtin_sqrt(res2) <= kLzAccept * L.sc[1]) break;
    if (attempts == kLzAttempts) {
      // accepted without passing the residual test: counted, reported
      if (lead && tid == 0) st->eig_unconverged += 1;
      break;
    }
    for (int i = tid; i < n; i += bs) L.v0[i] = L.ev[i];
    __syncthreads();
    first = 4;
    perturb = 0.0;
  }
  __syncthreads();
  if (L.iv[7] && tid == 0) st->error = ERR_SYNC;
  if (!lead || L.iv[7]) return;
  // advance the launch counter (every block read it before its first round)
  if (tid == 0) {
    __hip_atomic_store(xb, (uint64_t)(xcount + 1u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st->lz_prev = attempts > 1 ? (first0 / 2 > 4 ? first0 / 2 : 4) : L.iv[2];
  }
  eig_finish(st, L, k, n, normal, y1, rq, res2, dx, tabx, slot_stride, pslot, mode);
}

// k_eig in the row-owner form for n <= kEigLdsN (k + 1 <= 90), one wave
// per 64 rows: the same setup, solve, restart rule and update in the same
// arithmetic order.
__global__ __launch_bounds__(128) void k_eig1(DevState* st, int k, int ldh, const double* Hs, const double* px,
                                              const double* pz, double* dx, double* tabx, int slot_stride,
                                              const uint32_t* pslot, int mode) {
  __shared__ EigLds L;
  __shared__ double Qlds[kQBudget];  // A, then the Lanczos basis
  if (st->error) return;
  const int tid = threadIdx.x, bs = blockDim.x, lane = tid & 63, w = tid >> 6;
  // Every wave holds x, z of entries lane and lane + 64 (k <= 89).
  double xr[kR1], zr[kR1];
#pragma unroll
  for (int t = 0; t < kR1; ++t) {
    const int i = lane + 64 * t;
    xr[t] = i < k ? px[i] : 0.0;
    zr[t] = i < k ? pz[i] : 0.0;
    if (w == 0 && i < k) {
      L.x[i] = xr[t];
      L.z[i] = zr[t];
    }
  }
  // H_sub into LDS at its place A(r + 1, c + 1) of the bordered matrix
  // (lda k + 1), sixteen loads in flight per lane.
  {
    constexpr int U = 16;
    const int kk = k * k, na = k + 1;
    for (int i0 = tid; i0 < kk; i0 += U * bs) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = i0 + u * bs < kk ? i0 + u * bs : kk - 1;
        const int c = idx / k, r = idx - c * k;
        v[u] = Hs[(size_t)c * ldh + r];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = i0 + u * bs;
        if (idx < kk) {
          const int c = idx / k, r = idx - c * k;
          Qlds[(c + 1) * na + r + 1] = v[u];
        }
      }
    }
  }
  // x^T x and x^T z as ascending sums, the two chains interleaved (every
  // wave, entries broadcast by readlane)
  double xTx = 0.0, xTz = 0.0;
  bcast_each2(xr, zr, k, [&](double xi, double zi) {
    xTx = xTx + xi * xi;
    xTz = xTz + xi * zi;
  });
  const double yy = dd_add_d(dd{st->xx_hi, st->xx_lo}, -xTx).hi;
  const double y1 = __builtin_sqrt(yy);
  const bool normal = __builtin_fabs(yy) > 1e-14;
  const int n = normal ? k + 1 : k;
  __syncthreads();
  if (normal) {
    // H x from the LDS copy, then the border alpha = A(0, 0), b = A(1:, 0)
    const int r = lane + 64 * w;
    const double hxr = matvec_row(Qlds + n + 1, n, k, r, L.x);
    if (r < k) L.Hx[r] = hxr;
    __syncthreads();
    double hx[kR1];
#pragma unroll
    for (int t = 0; t < kR1; ++t) hx[t] = lane + 64 * t < k ? L.Hx[lane + 64 * t] : 0.0;
    double xTHx = 0.0;
    bcast_each2(xr, hx, k, [&](double xi, double hi) { xTHx = xTHx + xi * hi; });
    const dd xz = fix_to_dd((((i128)st->xz_hi) << 64) | (i128)st->xz_lo);
    const dd a00 = dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx);
    const double a0 = dd_div_d(a00, yy).hi;
    if (r < k) {
      const double b = ((w == 0 ? zr[0] : zr[1]) - hxr) / y1;
      Qlds[r + 1] = b;
      Qlds[(r + 1) * n] = b;
      L.v0[r + 1] = w == 0 ? xr[0] : xr[1];
    }
    if (tid == 0) {
      Qlds[0] = a0;
      L.v0[0] = y1;
    }
  } else {
    // x is the reference determinant alone (first iteration): A = H_sub,
    // lda k, reloaded from Hs; start from the pseudo-random vector.
    constexpr int U = 8;
    for (int i0 = tid; i0 < k * k; i0 += U * bs) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = i0 + u * bs < k * k ? i0 + u * bs : k * k - 1;
        const int c = idx / k, r = idx - c * k;
        v[u] = Hs[(size_t)c * ldh + r];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * bs < k * k) Qlds[i0 + u * bs] = v[u];
    }
    for (int i = tid; i < k; i += bs) L.v0[i] = 0.0;
  }
  __syncthreads();
  const int first0 = st->lz_prev - 1 > 4 ? st->lz_prev - 1 : 4;
  double* const Qb = Qlds + n * n;
  int first = first0, attempts = 0;
  double perturb = kLzPerturb;
  dd rq{0.0, 0.0};
  double res2 = 0.0;
  for (;;) {
    lanczos1(L, Qlds, n, first, perturb, Qb);
    for (int i = tid; i < n; i += bs) {
      L.ev[i] = L.q[i].hi;
      L.dl[i] = L.q[i].lo;
    }
    __syncthreads();
    {
      const int r = lane + 64 * w;
      const double ae = matvec_row(Qlds, n, n, r, L.ev);
      if (r < n) L.Ae[r] = ae;
    }
    __syncthreads();
    const dd t1 = red64_dd(n, [&](int i) { return two_prod(L.ev[i], L.Ae[i]); });
    const dd t2 = red64_dd(n, [&](int i) { return two_prod(L.dl[i], L.Ae[i]); });
    rq = dd_add(t1, dd_mul_d(t2, 2.0));
    const double lrq = rq.hi;
    res2 = red64(n, [&](int i) {
      double t = L.Ae[i] + (-lrq) * L.ev[i];
      return t * t;
    });
    ++attempts;
    if (__builtin_sqrt(res2) <= kLzAccept * L.sc[1]) break;
    if (attempts == kLzAttempts) {
      if (tid == 0) st->eig_unconverged += 1;
      break;
    }
    __syncthreads();
    for (int i = tid; i < n; i += bs) L.v0[i] = L.ev[i];
    __syncthreads();
    first = 4;
    perturb = 0.0;
  }
  if (tid == 0) st->lz_prev = attempts > 1 ? (first0 / 2 > 4 ? first0 / 2 : 4) : L.iv[2];
  eig_finish(st, L, k, n, normal, y1, rq, res2, dx, tabx, slot_stride, pslot, mode);
}

// Line-search update of the one picked coordinate (src/solver.h:254-314):
// the root of g(t) = t^3 + p1 t + q, p1 = xx - x^2 + h, q = z - h x, that
// the reference's closed form selects -- the outer root on the side
// opposite to q -- by Newton's iteration from outside every root
// (|t| <= max(1, |p1| + |q|)), then the reference's Newton polish of
// dx = t - x.  The state keeps scale = 1 and xx += dx (2x + dx), which the
// reference omits for this update.  One thread; update_x as in k_eig.
__global__ void k_ls(DevState* st, const double* Hs, const double* px, const double* pz, double* dx,
                     double* tabx, int slot_stride, const uint32_t* pslot) {
  if (threadIdx.x != 0 || st->error) return;
  const double h = Hs[0], x = px[0], z = pz[0];
  dd xx{st->xx_hi, st->xx_lo};
  const dd x2 = two_prod(x, x);
  const double p1 = dd_add_d(dd_add(xx, dd{-x2.hi, -x2.lo}), h).hi;
  const double q = z + (-h) * x;
  const double sd = q < 0.0 ? 1.0 : -1.0;
  double t = sd * fmax(1.0, __builtin_fabs(p1) + __builtin_fabs(q));
  for (int it = 0; it < 400; ++it) {
    const double g = (t * t + p1) * t + q;
    const double gp = 3.0 * (t * t) + p1;
    const double tn = t - g / gp;
    if (!(sd * tn < sd * t)) break;
    t = tn;
  }
  auto newton = [&](double d) {
    return d - (d * (d * (d + 3 * x) + (3 * x * x + p1)) + p1 * x + q + x * x * x) /
                   (d * (3 * d + 6 * x) + 3 * x * x + p1);
  };
  double d = t - x;
  double dn = newton(d);
  for (int iter = 0; __builtin_fabs((dn - d) / d) > 1e-12 && iter < 10; ++iter) {
    d = dn;
    dn = newton(d);
  }
  xx = dd_add(xx, dd_mul_d(two_sum(2.0 * x, d), d));
  st->xx_hi = xx.hi;
  st->xx_lo = xx.lo;
  st->sc_hi = 1.0;
  st->sc_lo = 0.0;
  dx[0] = d;
  double* xs = tabx + (size_t)pslot[0] * slot_stride;
  const double xo = *xs;
  if (!(__builtin_fabs(xo) != 0.0)) st->x_size += 1;
  *xs = xo + d;
  i128 xz = (((i128)st->xz_hi) << 64) | (i128)st->xz_lo;
  xz += to_fix(z * d);
  st->xz_hi = (int64_t)(xz >> 64);
  st->xz_lo = (uint64_t)xz;
  st->k_cur = 1;
}

void launch_ls(hipStream_t s, DevState* st, const double* Hs, const double* px, const double* pz, double* dx,
               double* tabx, int slot_stride, const uint32_t* pslot) {
  hipLaunchKernelGGL(k_ls, dim3(1), dim3(64), 0, s, st, Hs, px, pz, dx, tabx, slot_stride, pslot);
}

void launch_eig(hipStream_t s, DevState* st, int k, int ldh, const double* Hs, const double* px,
                const double* pz, double* dx, double* tabx, int slot_stride, const uint32_t* pslot, int mode,
                uint64_t* xb) {
  // Up to n = k + 1 = kEigLdsN the row-owner form, one wave per 64 rows;
  // above, the rows of A spread over ceil((k + 1) / kRRows) workgroups that
  // each run the solve on eight waves.
  if (k + 1 <= kEigLdsN)
    hipLaunchKernelGGL(k_eig1, dim3(1), dim3(k + 1 <= 64 ? 64 : 128), 0, s, st, k, ldh, Hs, px, pz, dx, tabx,
                       slot_stride, pslot, mode);
  else
    hipLaunchKernelGGL(k_eig, dim3((k + 1 + kRRows - 1) / kRRows), dim3(kEigThreads), 0, s, st, k, ldh, Hs, px,
                       pz, dx, tabx, slot_stride, pslot, mode, xb);
}

}  // namespace mc
