// Device-side data layout shared by the kernels and the host engine.
//
// HBM layout (one context):
//   table      : 2^p slots, open addressing with bucketed linear probing.
//                W = 1 (up to 31 spatial orbitals): {key, x, zA, zB} (32 B,
//                4 slots per 128 B line), key 0 = empty, bit 63 marks a slot
//                inserted during the current iteration (cleared in pass 2).
//                W = 2 (up to 63): {k0, k1, x, zA, zB} (40 B), claim, mark
//                and empty conventions of det.h.  z = (zA * 2^53 + zB) * 2^-64
//                exactly (num.h: tz_join).
//   integrals  : dense spatial (pq|rs) [M^4] and h [M^2] in fp64 and fp32,
//                single-excitation sums S_same/S_opp [M^3], diagonal terms
//                per spin-orbital; all read through L2/MALL.
//   iteration  : picked dets (two words each)/slots/x/z, dx, H_sub (k x k),
//                per-column visibility bitmaps, per-chunk top-M staging lists
//                and exact partial sums, the candidate lists.
#pragma once

#include <cstdint>

#include "det.h"
#include "num.h"

namespace mc {

constexpr int kWG = 256;               // threads per workgroup
constexpr int kEPT = 4;                // column elements per thread
constexpr int kChunk = kWG * kEPT;     // column elements per workgroup
constexpr int kSortCap = 4096;         // entries sorted by one workgroup
// Iteration counters (new keys, updates, column elements) are accumulated in
// kCtrStripes stripes, one 128-B line apart, so that the scatter's per-
// workgroup atomics spread over L2 channels instead of queueing on one line.
constexpr int kCtrStripes = 64;
constexpr int kCtrStride = 16;  // int64 per stripe
// Workgroups of the whole-table sweeps (init, dump, compaction), which loop
// over the slots: a dispatch's grid is limited to 2^32 work-items, the table
// may hold 2^32 slots.
constexpr uint64_t kSlotBlocks = 1 << 20;
constexpr int kMaxK = 256;             // coordinates per iteration
constexpr int kMaxOrb = 31;            // spatial orbitals of one-word determinants (bit 63 spare)
constexpr int kMaxOrb2 = 63;           // spatial orbitals of two-word determinants (bits 126, 127 spare)

template <int W>
struct SlotW;
template <>
struct SlotW<1> {
  uint64_t key;
  double x;
  int64_t za;
  int64_t zb;
};
template <>
struct SlotW<2> {
  uint64_t k0, k1;
  double x;
  int64_t za;
  int64_t zb;
};
using Slot = SlotW<1>;

// Integral tables and column enumeration layout for one (n_alpha, n_beta)
// sector.  Column element index e:
//   0                      diagonal
//   1 + [0, n_sa)          alpha singles   (occ o, virt u) -> o*nva + u
//   off_sb + [0, n_sb)     beta singles
//   off_daa + [0, n_daa)   alpha-alpha doubles (occ pair, virt pair)
//   off_dbb + [0, n_dbb)   beta-beta doubles
//   off_dab + [0, n_dab)   alpha-beta doubles ((oa*nb + ob)*nva + va)*nvb + vb
struct DevHam {
  int M, ns, na, nb, nva, nvb, ne;
  int cpa_o, cpb_o, cpa_v, cpb_v;  // pair counts C(na,2), C(nb,2), C(nva,2), C(nvb,2)
  int n_sa, n_sb, n_daa, n_dbb, n_dab;
  int off_sb, off_daa, off_dbb, off_dab, n_enum;
  int mixed;
  double core, thr;
  const double* eri;
  const float* eri32;
  const double* h1;
  const float* h132;
  const double* s_same;  // [p][r][k] = (pr|kk) - (pk|kr)
  const double* s_opp;   // [p][r][k] = (pr|kk)
  const float* s_same32;
  const float* s_opp32;
  const uint8_t* single_ok;  // [p][r]
  const double* d1;          // [spin-orbital] h_ii
  const double* d2;          // [i*ns + j] <ij||ij>, i < j
  // Spin-block strides (elements) for unrestricted integrals, 0 when one
  // block serves both spins: h1/h132/single_ok of spin s start at s*h_so,
  // S_same/S_opp of spin s at s*s_so; eri/eri32 hold the alpha-alpha block,
  // the beta-beta block at e_bb and the alpha-beta block at e_ab.
  size_t h_so, s_so, e_bb, e_ab;
};

// Device-resident solver state (scalars and counters).
struct DevState {
  double xx_hi, xx_lo, sc_hi, sc_lo;
  int64_t xz_hi;   // x.z as a fixed-point 128-bit value
  uint64_t xz_lo;
  long long x_size, z_size, updates, cols_size, iterations;
  int k_cur;       // columns of the running update
  int k_prev;      // columns whose candidates feed the next pick
  int error;
  unsigned int ticket;
  double z_thr;
  double max_size;
  double dx_max;
  double max_residual, curr_max_abs_evec, sum_diff_inv, max_diff_inv, curr_diff_inv;
  long long sum_iterations, sum_inv_iter, curr_inv_iter;
  int lz_prev;     // Lanczos vectors used by the previous subspace solve
  long long eig_unconverged;  // subspace solves accepted after the last restart without passing the residual test
  unsigned int znorm_cnt;     // slots listed by pass 2 for folding B into A (k_znorm)
  int znorm_full;             // the list overflowed: k_znorm sweeps the whole table
  int znorm_bits;             // fold when |B| > 2^znorm_bits (61; lower only to test the fold)
};

constexpr unsigned int kZNormCap = 1 << 16;  // slots the fold list holds per iteration

enum { ERR_NONE = 0, ERR_TABLE_FULL = 3, ERR_NUMERIC = 5, ERR_CANDIDATES = 6, ERR_PROBE = 8, ERR_CAP = 9,
       ERR_SYNC = 10 };

struct StageEnt {
  uint64_t val;   // bits of |x*xss+z| (non-negative double)
  uint64_t det;   // determinant word 0
  uint64_t det1;  // determinant word 1 (0 for one-word determinants)
  uint32_t slot;  // table slot, or 0xffffffff when absent
  uint32_t pad;
};

struct CandEnt {
  uint64_t det;
  uint64_t det1;
  uint32_t slot;
  uint32_t pad;
};

// A refetched candidate of the pick: key = bits of |x*xss + z|.
struct PickEnt {
  uint64_t key;
  uint64_t det;
  uint64_t det1;
  uint32_t slot;
  uint32_t pad;
};

}  // namespace mc
