// CDFCI device kernels for gfx950 (determinants of one or two 64-bit words,
// template parameter W).
//
// One iteration of the reference loop (src/solver.h:701-759) is six launches:
//   k_pick       gcd_grad pick: refetch candidates, sort, take k distinct
//                (src/solver.h:118-224)
//   k_submatrix  H_sub by Slater-Condon rules (src/solver.h:637-670,
//                src/hamiltonian.h:252-328)
//   k_eig        (k+1)-dim subspace update + update_x (src/solver.h:333-609,
//                src/wavefunction.h:425-451)
//   k_scatter    column generation (src/hamiltonian.h:336-418) fused with the
//                z += H dx scatter-add and insert gate (src/wavefunction.h:499-528)
//   k_pass2      post-scatter values: new_z, x.z, per-chunk top-M
//                (src/wavefunction.h:540-551)
//   k_merge      per-column top-M, z reinsert, scalar bookkeeping
//                (src/wavefunction.h:554-604)
// Columns are never materialised: every workgroup regenerates its chunk of
// column elements from the determinant and the L2-resident integral tables.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "device.h"
#include "kernels.h"
#include "sort.h"
#include "column.h"

namespace mc {

// ------------------------------------------------------------------ helpers
// Relaxed device-coherent load: no cache invalidation, no ordering.  Every
// cross-workgroup fact the scatter needs (present / inserted this iteration)
// lives in the key word itself, so no acquire/release fences are required.
__device__ __forceinline__ uint64_t rload(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// z += v (2^-64 units): one atomic on B for |v| < 2^53, else A and B.
template <class S>
__device__ __forceinline__ void zadd(S* s, i128 v) {
  constexpr i128 kOne = ((i128)1) << kZShift;
  if (v > -kOne && v < kOne) {
    if (v) __hip_atomic_fetch_add(&s->zb, (int64_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  int64_t a, b;
  tz_split(v, a, b);
  if (a) __hip_atomic_fetch_add(&s->za, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (b) __hip_atomic_fetch_add(&s->zb, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double dbits(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t bitsd(double d) { return (uint64_t)__double_as_longlong(d); }

// Gradient magnitude |x*s + z| with s = xx*scale^2 in double-double
// (src/solver.h:124,211; src/wavefunction.h:475,545).
__device__ __forceinline__ double grad_key(dd s, double x, double z) {
  dd r = dd_add_d(dd_mul_d(s, x), z);
  return __builtin_fabs(r.hi);
}

__device__ __forceinline__ dd load_xss(const DevState* st) {
  dd xx{st->xx_hi, st->xx_lo}, sc{st->sc_hi, st->sc_lo};
  return dd_mul(dd_mul(xx, sc), sc);
}

// Single excitation value h_pr + sum_{k in occ} <ik||ak> (src/hamiltonian.h:343-347),
// summed over the occupied spin-orbitals in ascending order; the tables of
// the moving electron's spin.
template <bool MIXED>
__device__ __forceinline__ double single_value(const DevHam& H, const uint8_t* occ, int ne, int p,
                                               int r, int spin) {
  const int M = H.M;
  const size_t base = ((size_t)p * M + r) * M + spin * H.s_so;
  const size_t hi = spin * H.h_so + p * M + r;
  if (MIXED) {
    float v = H.h132[hi];
    for (int x = 0; x < ne; ++x) {
      int k = occ[x];
      float t = ((k & 1) == spin) ? H.s_same32[base + (k >> 1)] : H.s_opp32[base + (k >> 1)];
      v = v + t;
    }
    return (double)v;
  } else {
    double v = H.h1[hi];
    for (int x = 0; x < ne; ++x) {
      int k = occ[x];
      double t = ((k & 1) == spin) ? H.s_same[base + (k >> 1)] : H.s_opp[base + (k >> 1)];
      v = v + t;
    }
    return v;
  }
}

// (p_a r_a|q_b s_b): the alpha-beta block.
template <bool MIXED>
__device__ __forceinline__ double eri_val(const DevHam& H, int p, int r, int q, int s) {
  const int M = H.M;
  size_t i = (((size_t)p * M + r) * M + q) * M + s + H.e_ab;
  return MIXED ? (double)H.eri32[i] : H.eri[i];
}

// <pq||rs> of one spin: (pr|qs) - (ps|qr) in that spin's block.
template <bool MIXED>
__device__ __forceinline__ double same_spin_val(const DevHam& H, int spin, int p, int r, int q, int s) {
  const int M = H.M;
  const size_t o = spin ? H.e_bb : 0;
  size_t i1 = (((size_t)p * M + r) * M + q) * M + s + o;
  size_t i2 = (((size_t)p * M + s) * M + q) * M + r + o;
  if (MIXED) return (double)(H.eri32[i1] - H.eri32[i2]);
  return H.eri[i1] - H.eri[i2];
}

// Element e >= 1 of the column of C.det.  Returns false when the element is
// not part of the column (threshold filter of src/hamiltonian.h:138-161,206).
// Signs: i -> a applied first, then j -> b on the intermediate determinant,
// each the parity of the occupied spin-orbitals between the two positions
// (src/hamiltonian.h:349-356,395-403).
template <bool MIXED, int W>
__device__ __forceinline__ bool gen_elem(const DevHam& H, const ColCtxW<W>& C, int e, D2& out, double& V) {
  const D2 D = C.det;
  int i, j, a, b;
  double v;
  if (e < H.off_daa) {
    int f = e - 1;
    int spin, o, u;
    if (f < H.n_sa) {
      spin = 0;
      o = f / H.nva;
      u = f - o * H.nva;
    } else {
      spin = 1;
      f -= H.n_sa;
      o = f / H.nvb;
      u = f - o * H.nvb;
    }
    int p = spin ? C.ob[o] : C.oa[o];
    int r = spin ? C.vb[u] : C.va[u];
    if (!H.single_ok[spin * H.h_so + p * H.M + r]) return false;
    v = single_value<MIXED>(H, C.occ, C.ne, p, r, spin);
    i = 2 * p + spin;
    a = 2 * r + spin;
    D2 d1 = D;
    dflip<W>(d1, i);
    if (dparity<W>(d1, i, a)) v = -v;
    dflip<W>(d1, a);
    out = d1;
    V = v;
    return true;
  }
  if (e < H.off_dab) {
    int spin, f, cv;
    if (e < H.off_dbb) {
      spin = 0;
      f = e - H.off_daa;
      cv = H.cpa_v;
    } else {
      spin = 1;
      f = e - H.off_dbb;
      cv = H.cpb_v;
    }
    int op = f / cv;
    int vp = f - op * cv;
    uint16_t po = spin ? C.pob[op] : C.poa[op];
    uint16_t pv = spin ? C.pvb[vp] : C.pva[vp];
    const uint8_t* ol = spin ? C.ob : C.oa;
    const uint8_t* vl = spin ? C.vb : C.va;
    int p = ol[po & 0xff], q = ol[po >> 8], r = vl[pv & 0xff], s = vl[pv >> 8];
    v = same_spin_val<MIXED>(H, spin, p, r, q, s);
    if (!(__builtin_fabs(v) > H.thr)) return false;
    i = 2 * p + spin;
    j = 2 * q + spin;
    a = 2 * r + spin;
    b = 2 * s + spin;
  } else {
    int f = e - H.off_dab;
    int vbi = f % H.nvb;
    int t = f / H.nvb;
    int vai = t % H.nva;
    t = t / H.nva;
    int obi = t % H.nb;
    int oai = t / H.nb;
    int p = C.oa[oai], q = C.ob[obi], r = C.va[vai], s = C.vb[vbi];
    v = eri_val<MIXED>(H, p, r, q, s);
    if (!(__builtin_fabs(v) > H.thr)) return false;
    if (p <= q) {
      i = 2 * p;
      j = 2 * q + 1;
      a = 2 * r;
      b = 2 * s + 1;
    } else {
      i = 2 * q + 1;
      j = 2 * p;
      a = 2 * s + 1;
      b = 2 * r;
    }
  }
  D2 d1 = D;
  dflip<W>(d1, i);
  int s1 = dparity<W>(d1, i, a);
  dflip<W>(d1, a);
  dflip<W>(d1, j);
  int s2 = dparity<W>(d1, j, b);
  dflip<W>(d1, b);
  if (s1 ^ s2) v = -v;
  out = d1;
  V = v;
  return true;
}

// ------------------------------------------------------------------ diagonal / off-diagonal
// get_off_diagonal (src/hamiltonian.h:274-328), fp64.
template <int W>
__device__ double off_diagonal_of(const DevHam& H, const D2& D1, const D2& D2b) {
  const D2 x = dxor(D1, D2b);
  const int pc = dpopc<W>(x);
  if (pc == 2) {
    int i = dffs<W>(dand(x, D1));
    int a = dffs<W>(dand(x, D2b));
    int spin = i & 1;
    int p = i >> 1, r = a >> 1;
    if (!H.single_ok[spin * H.h_so + p * H.M + r]) return 0.0;
    const size_t base = ((size_t)p * H.M + r) * H.M + spin * H.s_so;
    double v = H.h1[spin * H.h_so + p * H.M + r];
    D2 o = D1;
    for (int k = dffs<W>(o); k >= 0; k = dffs<W>(o)) {
      dpop_low<W>(o);
      double t = ((k & 1) == spin) ? H.s_same[base + (k >> 1)] : H.s_opp[base + (k >> 1)];
      v = v + t;
    }
    D2 d1 = D1;
    dflip<W>(d1, i);
    if (dparity<W>(d1, i, a)) v = -v;
    return v;
  }
  if (pc == 4) {
    D2 from = dand(x, D1), to = dand(x, D2b);
    int i = dffs<W>(from);
    dpop_low<W>(from);
    int j = dffs<W>(from);
    int a0 = dffs<W>(to);
    dpop_low<W>(to);
    int a1 = dffs<W>(to);
    int a, b;
    double v;
    if ((i & 1) == (j & 1)) {
      a = a0;
      b = a1;
      v = same_spin_val<false>(H, i & 1, i >> 1, a >> 1, j >> 1, b >> 1);
    } else {
      if ((a0 & 1) == (i & 1)) {
        a = a0;
        b = a1;
      } else {
        a = a1;
        b = a0;
      }
      // (i a|j b) with i, a of one spin and j, b of the other: the
      // alpha-beta block with the alpha pair first
      v = ((i & 1) == 0 || H.e_ab == 0) ? eri_val<false>(H, i >> 1, a >> 1, j >> 1, b >> 1)
                                        : eri_val<false>(H, j >> 1, b >> 1, i >> 1, a >> 1);
    }
    if ((i & 1) != (a & 1) || (j & 1) != (b & 1)) return 0.0;
    if (!(__builtin_fabs(v) > H.thr)) return 0.0;
    D2 d1 = D1;
    dflip<W>(d1, i);
    int s1 = dparity<W>(d1, i, a);
    dflip<W>(d1, a);
    dflip<W>(d1, j);
    int s2 = dparity<W>(d1, j, b);
    if (s1 ^ s2) v = -v;
    return v;
  }
  return 0.0;
}

// ------------------------------------------------------------------ k_submatrix
// One workgroup per row i computes H[i][j] for j >= i and stores it in both
// triangles (the subspace solver reads columns of the symmetric matrix).
// The diagonal keeps get_diagonal's sequential order (src/hamiltonian.h:
// 257-272): its terms h_aa, <ab||ab> (b > a) for occupied a ascending are
// gathered into LDS by all threads, then summed by one.  Two-word sectors
// hold at most 94 electrons (63 of one spin, 31 of the other).
template <int W>
__global__ __launch_bounds__(kWG) void k_submatrix(DevHam H, DevState* st, const D2* dets, int k,
                                                   double* Hs, int ldh) {
  constexpr int kNe = W == 1 ? 2 * kMaxOrb : 94;
  __shared__ double terms[kNe + kNe * (kNe - 1) / 2];
  __shared__ uint8_t occ[kNe];
  __shared__ int nocc;
  if (st && st->error) return;
  const int i = blockIdx.x;
  if (st && i == 0 && threadIdx.x == 0) {  // the previous merge has folded the z words pass 2 listed
    st->znorm_cnt = 0;
    st->znorm_full = 0;
  }
  const D2 Di = dets[i];
  if (threadIdx.x == 0) {
    int ne = 0;
    D2 r = Di;
    for (int s = dffs<W>(r); s >= 0; s = dffs<W>(r)) {
      dpop_low<W>(r);
      occ[ne++] = (uint8_t)s;
    }
    nocc = ne;
  }
  __syncthreads();
  const int ne = nocc;
  const int nterms = ne + ne * (ne - 1) / 2;
  for (int p = threadIdx.x; p < nterms; p += blockDim.x) {
    // row a holds ne - a terms: h_{occ[a]}, then <occ[a] occ[b]||...> for b > a
    int a = 0, off = 0;
    while (p >= off + (ne - a)) {
      off += ne - a;
      ++a;
    }
    const int q = p - off;
    terms[p] = q == 0 ? H.d1[occ[a]] : H.d2[occ[a] * H.ns + occ[a + q]];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int p = 0; p < nterms; ++p) v += terms[p];
    v += H.core;
    Hs[i * ldh + i] = v;
  }
  for (int j = i + 1 + threadIdx.x; j < k; j += blockDim.x) {
    const double v = off_diagonal_of<W>(H, Di, dets[j]);
    Hs[i * ldh + j] = v;
    Hs[j * ldh + i] = v;
  }
}

template <int N>
struct LaneCount {
  static constexpr int value = N;
};

// ------------------------------------------------------------------ wave / block reductions
__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ i128 wave_sum_i128(i128 v) {
  for (int off = 32; off >= 1; off >>= 1) {
    int64_t hi = (int64_t)(v >> 64);
    uint64_t lo = (uint64_t)v;
    int64_t ohi = __shfl_xor(hi, off, 64);
    uint64_t olo = (uint64_t)__shfl_xor((long long)lo, off, 64);
    v += (((i128)ohi) << 64) | (i128)olo;
  }
  return v;
}

// Exclusive prefix sum over the workgroup (blockDim <= 1024, multiple of
// 64); wtot needs blockDim/64 + 1 ints of LDS.  Returns this thread's
// prefix and the total in `total`.
__device__ __forceinline__ int block_excl_scan(int v, int* wtot, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wtot[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < nw; ++i) {
      int t = wtot[i];
      wtot[i] = acc;
      acc += t;
    }
    wtot[nw] = acc;
  }
  __syncthreads();
  total = wtot[nw];
  const int r = x - v + wtot[w];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ k_scatter
// Column generation (src/hamiltonian.h:336-418) fused with the scatter-add
// z += H dx and the insert gate (src/wavefunction.h:499-528).  Each
// workgroup takes EPT * kWG consecutive column elements; the visibility
// bitmap is indexed by element, so EPT is free of the pass-2 chunking.
template <bool MIXED, int W, int EPT>
__global__ __launch_bounds__(kWG) void k_scatter(DevHam H, DevState* st, const D2* pdet, const double* dx,
                                                 const double* Hs, int ldh, SlotW<W>* tab, uint64_t mask,
                                                 uint64_t* vis, int vis_words, long long* ctr, uint32_t* eslot,
                                                 int estride) {
  __shared__ ColCtxW<W> C;
  if (st->error) return;
  const int t = blockIdx.y;
  const int chunk = blockIdx.x;
  setup_ctx<W>(C, H, pdet[t]);
  const double dxt = dx[t];
  const dd sc{st->sc_hi, st->sc_lo};
  const double zthr = st->z_thr;
  const double diag = Hs[t * ldh + t];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long n_new = 0, n_upd = 0, n_valid = 0;
  bool fail = false;
  for (int r = 0; r < EPT; ++r) {
    const int e = chunk * (kWG * EPT) + r * kWG + threadIdx.x;
    bool visible = false;
    if (e < H.n_enum) {
      D2 D{0, 0};
      double V = 0.0;
      bool valid;
      if (e == 0) {
        D = C.det;
        V = diag;
        valid = true;
      } else {
        valid = gen_elem<MIXED, W>(H, C, e, D, V);
      }
      if (valid) {
        ++n_valid;
        const double dz = V * dxt;
        const bool gate = dd_abs_gt(dd_mul_d(sc, dz), zthr);
        SlotW<W>* hit = nullptr;
        bool mine = false, marked = false;
        uint64_t b = dhash<W>(D) & mask & ~3ULL;
        bool done = false;
        uint64_t stop = ~0ULL;  // slot where the probe ended (hit, or the empty slot)
        if constexpr (W == 1) {
          for (uint64_t n = 0; n <= mask && !done; n += 4) {
            for (int jj = 0; jj < 4 && !done; ++jj) {
              SlotW<W>* s = &tab[b + jj];
              uint64_t k = rload(&s->key);
              if ((k & kKeyMask) == D.lo) {
                hit = s;
                marked = (k & kMark) != 0;
                done = true;
              } else if (k == 0) {
                if (!gate) {
                  done = true;
                  stop = b + jj;
                } else {
                  uint64_t expect = 0;
                  bool ok = __hip_atomic_compare_exchange_strong(&s->key, &expect, D.lo | kMark, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  if (ok) {
                    hit = s;
                    mine = true;
                    done = true;
                  } else if ((expect & kKeyMask) == D.lo) {
                    hit = s;
                    done = true;
                  }
                }
              }
            }
            b = (b + 4) & mask;
          }
        } else {
          // One slot per trip of a flat loop.  A slot claimed (word 1) but
          // not yet published (word 0) is read again on the next trip
          // rather than waited on inside the trip: the claiming lane may sit
          // in the same wavefront, and its publishing store then completes
          // within the current trip.
          uint64_t n = 0;
          int jj = 0, spins = 0;
          while (!done) {
            SlotW<W>* s = &tab[b + jj];
            uint64_t k1 = rload(&s->k1);
            if (k1 == 0) {
              if (!gate) {
                done = true;
                stop = b + jj;
              } else {
                uint64_t expect = 0;
                if (__hip_atomic_compare_exchange_strong(&s->k1, &expect, D.hi | kOcc | kMark, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                  __hip_atomic_store(&s->k0, D.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  hit = s;
                  mine = true;
                  done = true;
                }
                k1 = expect;
              }
            }
            if (!done && (k1 & kHiMask) == D.hi) {
              const uint64_t k0 = rload(&s->k0);
              if (k0 == kSent) {
                if (++spins > (1 << 22)) break;
                continue;
              }
              if (k0 == D.lo) {
                hit = s;
                marked = (k1 & kMark) != 0;
                done = true;
              }
            }
            if (!done && ++jj == 4) {
              jj = 0;
              b = (b + 4) & mask;
              n += 4;
              if (n > mask) break;
            }
          }
        }
        if (!done) fail = true;
        if (hit) stop = (uint64_t)(hit - tab);
        eslot[(size_t)t * estride + e] = (uint32_t)(stop & 0xffffffffULL);
        if (hit) {
          // Gate-passing elements always add; the others only into
          // determinants that were present before this iteration.
          visible = gate || !marked;
          if (mine) ++n_new;
          if (visible) {
            zadd(hit, to_fix(dz));
            ++n_upd;
          }
        }
      }
    }
    unsigned long long ball = __ballot(visible);
    const int word = chunk * (kWG * EPT / 64) + r * (kWG / 64) + wave;
    if (lane == 0 && word < vis_words) vis[(size_t)t * vis_words + word] = ball;
  }
  n_new = wave_sum_ll(n_new);
  n_upd = wave_sum_ll(n_upd);
  n_valid = wave_sum_ll(n_valid);
  __shared__ long long wsum[kWG / 64][3];
  if (lane == 0) {
    wsum[wave][0] = n_new;
    wsum[wave][1] = n_upd;
    wsum[wave][2] = n_valid;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    long long v = 0;
    for (int w = 0; w < kWG / 64; ++w) v += wsum[w][threadIdx.x];
    const int stripe = (int)((blockIdx.y * gridDim.x + blockIdx.x) % kCtrStripes);
    if (v) atomicAdd((unsigned long long*)&ctr[stripe * kCtrStride + threadIdx.x], (unsigned long long)v);
  }
  if (fail) st->error = ERR_PROBE;
}

// ------------------------------------------------------------------ k_pass2
// Values after the scatter: per-element |x*xss+z| (0 when the element does
// not see its determinant), exact partial sums of H*x and dz*x, and the
// chunk's top-M list sorted by (value, determinant) descending.
// SEL 0: full sort of the chunk; 1: top L <= 64 by the per-wave tournament;
// 2: top 256 (kept entries 65..256).
// At one word the registers are capped for five waves per SIMD (more random
// slot reads in flight; a handful of spills): k = 256 pass 2 358 -> 345 us;
// six waves spill 20 registers and run slower.
template <bool MIXED, int W, int SEL>
__global__ __launch_bounds__(kWG, W == 1 ? 5 : 1) void k_pass2(DevHam H, DevState* st, const D2* pdet, const double* dx,
                                               const double* Hs, int ldh, SlotW<W>* tab, uint64_t mask,
                                               const uint64_t* vis, int vis_words, int M, int nchunks,
                                               StageEnt* stage, int* stage_cnt, int64_t* part,
                                               const uint32_t* eslot, uint32_t* znorm_list) {
  __shared__ ColCtxW<W> C;
  // Selection scratch: the top-L tournament needs 512 entries, a full
  // sort the whole chunk (less LDS per workgroup, more workgroups per CU).
  constexpr int kSel = SEL == 1 ? 8 * 64 : (SEL == 2 ? 512 : kChunk);
  __shared__ uint64_t ka[kSel];
  __shared__ uint64_t kb[kSel];
  __shared__ uint64_t kc[W == 2 ? kSel : 1];
  __shared__ uint32_t pl[kSel];
  __shared__ i128 red[kWG / 64][2];
  if (st->error) return;
  // Columns and chunks in the reverse of the scatter's order: the lines the
  // scatter touched last are the likeliest still in the last-level cache.
  const int t = gridDim.y - 1 - blockIdx.y;
  const int chunk = gridDim.x - 1 - blockIdx.x;
  setup_ctx<W>(C, H, pdet[t]);
  const double dxt = dx[t];
  const dd s = load_xss(st);
  const double diag = Hs[t * ldh + t];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  i128 nz = 0, xt = 0;
  uint64_t A[kEPT], B[kEPT], B1[kEPT];
  uint32_t P[kEPT];
  int nvalid = 0;
  // Generate the thread's elements, then read every element's slot at the
  // point where the scatter's probe ended before resolving any of them, so
  // that the thread's table reads are in flight together.
  D2 Dv[kEPT];
  double Vv[kEPT];
  bool ok[kEPT];
  uint32_t r0[kEPT];
  uint64_t vw[kEPT];
#pragma unroll
  for (int r = 0; r < kEPT; ++r) {
    const int e = chunk * kChunk + r * kWG + threadIdx.x;
    Dv[r] = D2{0, 0};
    Vv[r] = 0.0;
    ok[r] = false;
    r0[r] = 0xffffffffu;
    vw[r] = 0;
    if (e < H.n_enum) {
      if (e == 0) {
        Dv[r] = C.det;
        Vv[r] = diag;
        ok[r] = true;
      } else {
        ok[r] = gen_elem<MIXED, W>(H, C, e, Dv[r], Vv[r]);
      }
      if (ok[r]) {
        r0[r] = eslot[(size_t)t * nchunks * kChunk + e];
        vw[r] = vis[(size_t)t * vis_words + (e >> 6)];
      }
    }
  }
  SlotW<W> sv[kEPT];
#pragma unroll
  for (int r = 0; r < kEPT; ++r) {
    if (r0[r] != 0xffffffffu) {
      sv[r] = tab[r0[r]];
    } else {
      if constexpr (W == 1) sv[r].key = 0;
      else sv[r].k0 = kSent, sv[r].k1 = 0;
      sv[r].x = 0.0;
      sv[r].za = 0;
      sv[r].zb = 0;
    }
  }
#pragma unroll
  for (int r = 0; r < kEPT; ++r) {
    const int e = chunk * kChunk + r * kWG + threadIdx.x;
    uint64_t keyv = 0;
    D2 keyd{0, 0};
    uint32_t slot = 0xffffffffu;
    if (ok[r]) {
      const D2 D = Dv[r];
      const double V = Vv[r];
      const double dz = V * dxt;
      // the slot read above: D's slot, empty (D absent), or another
      // determinant (then the probe continues after it)
      long long si;
      bool here, empty;
      if constexpr (W == 1) {
        here = (sv[r].key & kKeyMask) == D.lo;
        empty = sv[r].key == 0;
      } else {
        here = sv[r].k1 != 0 && (sv[r].k1 & kHiMask) == D.hi && sv[r].k0 == D.lo;
        empty = sv[r].k1 == 0;
      }
      if (r0[r] == 0xffffffffu) si = find_slot<W>(tab, mask, D);
      else if (here) si = r0[r];
      else if (empty) si = -1;
      else si = find_slot_at<W>(tab, mask, D, (uint32_t)((r0[r] + 1ULL) & mask));
      double x = 0.0, val = 0.0;
      if (si >= 0) {
        SlotW<W>* sl = &tab[si];
        SlotW<W> cur = (si == (long long)r0[r] && r0[r] != 0xffffffffu) ? sv[r] : *sl;
        // clear the inserted-this-iteration mark
        if constexpr (W == 1) {
          if (cur.key & kMark) sl->key = D.lo;
        } else {
          if (cur.k1 & kMark) sl->k1 = D.hi | kOcc;
        }
        x = cur.x;
        if ((vw[r] >> (e & 63)) & 1) {
          double z = fix_to_d(tz_join(cur.za, cur.zb));
          val = grad_key(s, x, z);
        }
        // B near its range: list the slot for k_znorm (duplicates are harmless)
        const int64_t lim = (int64_t)1 << st->znorm_bits;
        if (cur.zb > lim || cur.zb < -lim) {
          const unsigned int pos = atomicAdd(&st->znorm_cnt, 1u);
          if (pos < kZNormCap) znorm_list[pos] = (uint32_t)si;
          else st->znorm_full = 1;
        }
        slot = (uint32_t)si;
      }
      nz += to_fix(V * x);
      xt += to_fix(dz * x);
      keyv = bitsd(val);
      keyd = D;
      ++nvalid;
    }
    A[r] = keyv;
    B[r] = keyd.lo;
    B1[r] = keyd.hi;
    P[r] = slot;
  }
  // Exact chunk sums.
  nz = wave_sum_i128(nz);
  xt = wave_sum_i128(xt);
  if (lane == 0) {
    red[wave][0] = nz;
    red[wave][1] = xt;
  }
  // Valid entries have a nonzero determinant and sort before the (0, 0)
  // fillers.  Up to 64 kept entries: per-wave sorts and a top-L merge
  // tournament (the top L under a total order, so the same entries as a full
  // sort's first L); above, one sort of the whole chunk.
  const int mw = M < kChunk ? M : kChunk;
  int L = 1;
  while (L < mw) L <<= 1;
  uint64_t ta = 0, tb = 0, tc = 0;
  uint32_t tp = 0;
  if constexpr (SEL == 1)
    block_top_desc3<kEPT, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc}, L, ta, tb, tc, tp);
  else if constexpr (SEL == 2)
    block_top256_desc3<W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc});
  else
    block_sort_desc3<kEPT, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc});
  __shared__ int wtot[kWG / 64 + 1];
  int total_valid;
  (void)block_excl_scan(nvalid, wtot, total_valid);
  if (threadIdx.x == 0) {
    i128 a = 0, b = 0;
    for (int w = 0; w < kWG / 64; ++w) {
      a += red[w][0];
      b += red[w][1];
    }
    int64_t* o = part + ((size_t)t * nchunks + chunk) * 4;
    o[0] = (int64_t)(a >> 64);
    o[1] = (int64_t)(uint64_t)a;
    o[2] = (int64_t)(b >> 64);
    o[3] = (int64_t)(uint64_t)b;
  }
  StageEnt* out = stage + ((size_t)t * nchunks + chunk) * mw;
  const int keep = total_valid < mw ? total_valid : mw;
  if constexpr (SEL == 1) {
    if ((int)threadIdx.x < keep) {
      StageEnt& o = out[threadIdx.x];
      o.val = ta;
      o.det = tb;
      o.det1 = tc;
      o.slot = tp;
      o.pad = 0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < kEPT; ++r) {
      const int i = threadIdx.x * kEPT + r;
      if (i < keep) {
        out[i].val = A[r];
        out[i].det = B[r];
        out[i].det1 = B1[r];
        out[i].slot = P[r];
        out[i].pad = 0;
      }
    }
  }
  if (threadIdx.x == 0) stage_cnt[t * nchunks + chunk] = keep;
}

// Tournament in the staging buffer (chunk lists of at most 256 entries):
// list i absorbs list i + s at stride s, each pair merged by one wave into
// its top L (L = the power of two >= mw) and written back truncated to mw
// entries (an entry below position mw of a merged list is below the
// column's top M).  List 0 ends as the first mw entries of a full sort;
// they are returned in blocked registers (entry threadIdx.x * E + r).
template <int W, int E, int T>
__device__ void column_tournament(StageEnt* lists, int nchunks, int mw, const int* offs, int* lcnt,
                                uint64_t (&A)[E], uint64_t (&B)[E], uint64_t (&B1)[E], uint32_t (&P)[E]) {
  for (int c = threadIdx.x; c < nchunks; c += T) lcnt[c] = offs[c + 1] - offs[c];
  __syncthreads();
  int L = 1;
  while (L < mw) L <<= 1;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  auto merge_pair = [&](auto e2tag, int i, int j) {
    constexpr int E2 = decltype(e2tag)::value;
    uint64_t a[E2], b[E2], c[E2];
    uint32_t pp[E2];
    const int ni = lcnt[i], nj = lcnt[j];
    const StageEnt* X = lists + (size_t)i * mw;
    const StageEnt* Y = lists + (size_t)j * mw;
#pragma unroll
    for (int r = 0; r < E2; ++r) {
      const int q = lane * E2 + r;
      uint64_t xa = 0, xb = 0, xc = 0, ya = 0, yb = 0, yc = 0;
      uint32_t xp = 0xffffffffu, yp = 0xffffffffu;
      if (q < L && q < ni) {
        xa = X[q].val, xb = X[q].det, xc = X[q].det1, xp = X[q].slot;
      }
      if (q < L && L - 1 - q < nj) {
        const StageEnt& e = Y[L - 1 - q];
        ya = e.val, yb = e.det, yc = e.det1, yp = e.slot;
      }
      const bool ty = key_gt3<W == 2>(ya, yb, yc, xa, xb, xc);
      a[r] = ty ? ya : xa;
      b[r] = ty ? yb : xb;
      c[r] = ty ? yc : xc;
      pp[r] = ty ? yp : xp;
    }
    wave_bitonic_merge<E2, W == 2>(a, b, c, pp, L);
    const int keep = ni + nj < mw ? ni + nj : mw;
#pragma unroll
    for (int r = 0; r < E2; ++r) {
      const int q = lane * E2 + r;
      if (q < keep) {
        StageEnt& o = lists[(size_t)i * mw + q];
        o.val = a[r];
        o.det = b[r];
        o.det1 = c[r];
        o.slot = pp[r];
        o.pad = 0;
      }
    }
    if (lane == 0) lcnt[i] = keep;
  };
  for (int sd = 1; sd < nchunks; sd <<= 1) {
    for (int i = 2 * sd * wv; i + sd < nchunks; i += 2 * sd * (T / 64)) {
      if (L <= 64) merge_pair(LaneCount<1>{}, i, i + sd);
      else if (L == 128) merge_pair(LaneCount<2>{}, i, i + sd);
      else merge_pair(LaneCount<4>{}, i, i + sd);
    }
    __syncthreads();
  }
  const int n0 = lcnt[0];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int g = threadIdx.x * E + r;
    A[r] = 0;
    B[r] = 0;
    B1[r] = 0;
    P[r] = 0xffffffffu;
    if (g < n0) {
      const StageEnt& e = lists[g];
      A[r] = e.val;
      B[r] = e.det;
      B1[r] = e.det1;
      P[r] = e.slot;
    }
  }
}

// Two-level selection for columns of many chunk lists (mw <= 256, at most
// T*E/kHead lists): the first kHead entries of every list (its head) are
// reduced in LDS by the in-place pairwise tournament of column_tournament
// (list i absorbing list i + s, one wave per pair, each pair kept whole up to
// mw entries), giving S, the top mw of all heads.  A list's later entries
// (its tail) all rank below its last head entry, so the column's top mw lies
// in S and the tails of the lists whose last head entry is not below S's
// last (all tails when S holds fewer than mw entries): the tails of up to
// T*E/L - 1 such lists (L = the power of two >= mw) are loaded into LDS
// beside S and reduced with it by the same tournament.  List 0 ends as the
// first mw entries of a full sort of the column's entries, returned in
// blocked registers (entry threadIdx.x*E + r).  With more such lists (or
// more than tails_cap) it returns false without output and the caller runs
// column_tournament, which reads the staging buffer only.
constexpr int kHead = 64;
template <int W, int E, int T>
__device__ bool column_heads(const StageEnt* lists, int nchunks, int mw, const int* offs, int* lcnt,
                             const SortScratch& x, int tails_cap, uint64_t (&A)[E], uint64_t (&B)[E],
                             uint64_t (&B1)[E], uint32_t (&P)[E]) {
  constexpr bool C3 = W == 2;
  constexpr int kLists = T * E / kHead;
  __shared__ uint64_t lha[kLists], lhb[kLists], lhc[kLists];
  __shared__ int need;
  __shared__ int tail_of[kLists];
  for (int g = threadIdx.x; g < nchunks * kHead; g += T) {
    const int c = g / kHead, j = g - c * kHead;
    uint64_t a = 0, b = 0, b1 = 0;
    uint32_t p = 0xffffffffu;
    if (j < offs[c + 1] - offs[c]) {
      const StageEnt& e = lists[(size_t)c * mw + j];
      a = e.val;
      b = e.det;
      b1 = e.det1;
      p = e.slot;
    }
    x.a[g] = a;
    x.b[g] = b;
    x.p[g] = p;
    if constexpr (C3) x.c[g] = b1;
    if (j == kHead - 1) {
      lha[c] = a;
      lhb[c] = b;
      lhc[c] = b1;
    }
  }
  for (int c = threadIdx.x; c < nchunks; c += T) {
    const int cnt = offs[c + 1] - offs[c];
    lcnt[c] = cnt < kHead ? cnt : kHead;
  }
  if (threadIdx.x == 0) need = 0;
  __syncthreads();
  int L = 1;
  while (L < mw) L <<= 1;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // The lists at xo (count lcnt[i]) and yo (count lcnt[j]): the top Lk =
  // 64*E2 of both (Lk at least either count), written back from xo truncated
  // to mw entries; lcnt[i] becomes the kept count.
  auto merge_pair = [&](auto e2tag, int xo, int yo, int i, int j) {
    constexpr int E2 = decltype(e2tag)::value;
    constexpr int Lk = 64 * E2;
    uint64_t a[E2], b[E2], c[E2];
    uint32_t pp[E2];
    const int ni = lcnt[i], nj = lcnt[j];
#pragma unroll
    for (int r = 0; r < E2; ++r) {
      const int q = lane * E2 + r;
      uint64_t xa = 0, xb = 0, xc = 0, ya = 0, yb = 0, yc = 0;
      uint32_t xp = 0xffffffffu, yp = 0xffffffffu;
      if (q < ni) {
        xa = x.a[xo + q], xb = x.b[xo + q], xp = x.p[xo + q];
        if constexpr (C3) xc = x.c[xo + q];
      }
      if (Lk - 1 - q < nj) {
        const int yq = yo + Lk - 1 - q;
        ya = x.a[yq], yb = x.b[yq], yp = x.p[yq];
        if constexpr (C3) yc = x.c[yq];
      }
      const bool ty = key_gt3<C3>(ya, yb, yc, xa, xb, xc);
      a[r] = ty ? ya : xa;
      b[r] = ty ? yb : xb;
      c[r] = ty ? yc : xc;
      pp[r] = ty ? yp : xp;
    }
    wave_bitonic_merge<E2, C3>(a, b, c, pp, Lk);
    const int keep = ni + nj < mw ? ni + nj : mw;
#pragma unroll
    for (int r = 0; r < E2; ++r) {
      const int q = lane * E2 + r;
      if (q < keep) {
        x.a[xo + q] = a[r];
        x.b[xo + q] = b[r];
        x.p[xo + q] = pp[r];
        if constexpr (C3) x.c[xo + q] = c[r];
      }
    }
    if (lane == 0) lcnt[i] = keep;
  };
  auto merge_lk = [&](int Lk, int xo, int yo, int i, int j) {
    if (Lk <= 64) merge_pair(LaneCount<1>{}, xo, yo, i, j);
    else if (Lk == 128) merge_pair(LaneCount<2>{}, xo, yo, i, j);
    else merge_pair(LaneCount<4>{}, xo, yo, i, j);
  };
  // Heads: list c at c*kHead; a list at stride sd holds at most
  // min(sd*kHead, mw) entries, so Lk = min(2*sd*kHead, L) covers either
  // count and every kept entry (in the pair's own regions).
  for (int sd = 1; sd < nchunks; sd <<= 1) {
    const int Lk = 2 * sd * kHead < L ? 2 * sd * kHead : L;
    for (int i = 2 * sd * wv; i + sd < nchunks; i += 2 * sd * (T / 64)) merge_lk(Lk, i * kHead, (i + sd) * kHead, i, i + sd);
    __syncthreads();
  }
  const int n0 = lcnt[0];
  for (int c = threadIdx.x; c < nchunks; c += T) {
    if (offs[c + 1] - offs[c] > kHead &&
        (n0 < mw || !key_gt3<C3>(x.a[mw - 1], x.b[mw - 1], C3 ? x.c[mw - 1] : 0, lha[c], lhb[c], lhc[c]))) {
      const int r = atomicAdd(&need, 1);
      if (r < kLists) tail_of[r] = c;
    }
  }
  __syncthreads();
  const int m = need;
  if (m > T * E / L - 1 || m > tails_cap) return false;
  if (m > 0) {
    // S stays at [0, n0); tail r (list tail_of[r], entries kHead..count-1,
    // at most mw - kHead < L) goes to region r + 1 of L entries.
    for (int g = threadIdx.x; g < m * L; g += T) {
      const int r = g / L, j = g - r * L;
      const int c = tail_of[r];
      uint64_t a = 0, b = 0, b1 = 0;
      uint32_t p = 0xffffffffu;
      if (j < offs[c + 1] - offs[c] - kHead) {
        const StageEnt& e = lists[(size_t)c * mw + kHead + j];
        a = e.val;
        b = e.det;
        b1 = e.det1;
        p = e.slot;
      }
      const int o = (r + 1) * L + j;
      x.a[o] = a;
      x.b[o] = b;
      x.p[o] = p;
      if constexpr (C3) x.c[o] = b1;
    }
    // lcnt is free again: region counts (lcnt[0] = n0 already)
    if ((int)threadIdx.x < m) {
      const int c = tail_of[threadIdx.x];
      lcnt[threadIdx.x + 1] = offs[c + 1] - offs[c] - kHead;
    }
    __syncthreads();
    for (int sd = 1; sd < m + 1; sd <<= 1) {
      for (int i = 2 * sd * wv; i + sd < m + 1; i += 2 * sd * (T / 64)) merge_lk(L, i * L, (i + sd) * L, i, i + sd);
      __syncthreads();
    }
  }
  const int nf = lcnt[0];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int g = threadIdx.x * E + r;
    const bool in = g < nf;
    A[r] = in ? x.a[g] : 0;
    B[r] = in ? x.b[g] : 0;
    B1[r] = (C3 && in) ? x.c[g] : 0;
    P[r] = in ? x.p[g] : 0xffffffffu;
  }
  return true;
}

// ------------------------------------------------------------------ k_merge
// One workgroup per column: merge chunk lists into the column's top-M, keep
// the determinants present in the table (the refetch of
// src/wavefunction.h:567-575), reinsert z_t := new_z_t, and let the last
// workgroup update x.z and the counters.
// STREAM = false: all chunk lists of the column are sorted at once (at most
// T*E entries).  STREAM = true: the window of T*E = 2*Mp entries (Mp >= M,
// a power of two) holds the running top Mp in its first half; each further
// chunk list is loaded reversed into the second half and one bitonic merge
// level keeps the top Mp (chunk lists are sorted descending by pass 2).
template <int E, int T, bool STREAM, int W>
__global__ __launch_bounds__(T) void k_merge(DevState* st, const uint32_t* pslot, SlotW<W>* tab, int M,
                                               int nchunks, StageEnt* stage,
                                               const int* stage_cnt, const int64_t* part,
                                               int64_t* col_xz, CandEnt* cand, int* cand_cnt,
                                               const double* dx, int count_iteration, long long* ctr,
                                               uint64_t nslots, const uint32_t* znorm_list, int merge_tails) {
  __shared__ uint64_t ka[T * E];
  __shared__ uint64_t kb[T * E];
  __shared__ uint64_t kc[W == 2 ? T * E : 1];
  __shared__ uint32_t pl[T * E];
  __shared__ int offs[1025];
  __shared__ int total;
  __shared__ bool last;
  if (st->error) return;
  const int t = blockIdx.x;
  const int k = gridDim.x;
  const int mw = M < kChunk ? M : kChunk;
  {
    // Fold z word B into A (B -= c 2^53, A += c, c = floor(B / 2^53)) for
    // the slots pass 2 listed with |B| > 2^znorm_bits, or, when the list
    // overflowed, for every slot of the table.  Nothing else touches z here
    // except this kernel's own reinsert of the picked coordinates, whose
    // slots are skipped (the reinsert writes their z in folded form).  A slot
    // listed twice is folded once: the CAS succeeds for one observed B, and
    // the B re-read after it is small.  k_submatrix empties the list.
    const unsigned int zcnt = st->znorm_cnt;
    if (zcnt) {
      const bool full = st->znorm_full != 0;
      const uint64_t n = full ? nslots : (zcnt < kZNormCap ? zcnt : kZNormCap);
      const int64_t lim = (int64_t)1 << st->znorm_bits;
      const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
      for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t si = full ? i : (uint64_t)znorm_list[i];
        SlotW<W>* sl = &tab[si];
        int64_t b = __hip_atomic_load(&sl->zb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(b > lim || b < -lim)) continue;
        bool picked = false;
        for (int j = 0; j < k; ++j) picked = picked || pslot[j] == (uint32_t)si;
        if (picked) continue;
        while (b > lim || b < -lim) {
          const int64_t c = b >> kZShift;  // floor
          const int64_t nb = b - (c << kZShift);
          if (__hip_atomic_compare_exchange_strong(&sl->zb, &b, nb, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_fetch_add(&sl->za, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
    }
  }
  __shared__ int wtot[T / 64 + 1];
  {
    // nchunks <= 1024 <= 4 * T: up to four chunk counts per thread, in order.
    int v[4], sum = 0;
    for (int j = 0; j < 4; ++j) {
      const int c = threadIdx.x * 4 + j;
      v[j] = c < nchunks ? stage_cnt[t * nchunks + c] : 0;
      sum += v[j];
    }
    int tot;
    int pre = block_excl_scan(sum, wtot, tot);
    for (int j = 0; j < 4; ++j) {
      const int c = threadIdx.x * 4 + j;
      if (c <= nchunks) offs[c] = pre;
      pre += v[j];
    }
    if (threadIdx.x == 0) total = tot;
  }
  __syncthreads();
  const int n = total;
  uint64_t A[E], B[E], B1[E];
  uint32_t P[E];
  const SortScratch scr{ka, kb, pl, kc};
  if constexpr (!STREAM) {
    if (n > T * E) {
      if (threadIdx.x == 0) st->error = ERR_CAP;
      return;
    }
    int L = 1;
    while (L < mw) L <<= 1;
    if (L <= 64 && nchunks * L <= T * E) {
      // Tournament: the chunk lists (sorted, padded to L with fillers) are
      // reduced pairwise in place, list i absorbing list i + s at stride s,
      // each pair by one wave keeping the top L.  List 0 ends as the top L of
      // all entries: the first L entries of a full sort.
      for (int g = threadIdx.x; g < nchunks * L; g += T) {
        const int c = g / L, j = g - c * L;
        uint64_t a = 0, b = 0, b1 = 0;
        uint32_t p = 0xffffffffu;
        if (j < offs[c + 1] - offs[c]) {
          const StageEnt& e = stage[((size_t)t * nchunks + c) * mw + j];
          a = e.val;
          b = e.det;
          b1 = e.det1;
          p = e.slot;
        }
        ka[g] = a;
        kb[g] = b;
        pl[g] = p;
        if constexpr (W == 2) kc[g] = b1;
      }
      __syncthreads();
      const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
      for (int s = 1; s < nchunks; s <<= 1) {
        for (int i = 2 * s * wv; i + s < nchunks; i += 2 * s * (T / 64)) {
          uint64_t a, b, b1;
          uint32_t p;
          wave_merge_top<W == 2>(scr, i * L, (i + s) * L, L, a, b, b1, p);
          if (lane < L) {
            ka[i * L + lane] = a;
            kb[i * L + lane] = b;
            pl[i * L + lane] = p;
            if constexpr (W == 2) kc[i * L + lane] = b1;
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const int g = threadIdx.x * E + r;
        const bool in = g < L;
        A[r] = in ? ka[g] : 0;
        B[r] = in ? kb[g] : 0;
        B1[r] = (W == 2 && in) ? kc[g] : 0;
        P[r] = in ? pl[g] : 0xffffffffu;
      }
    } else if (mw <= 256) {
      __shared__ int lcnt[1024];
      column_tournament<W, E, T>(stage + (size_t)t * nchunks * mw, nchunks, mw, offs, lcnt, A, B, B1, P);
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const int g = threadIdx.x * E + r;
        A[r] = 0;
        B[r] = 0;
        B1[r] = 0;
        P[r] = 0xffffffffu;
        if (g < n) {
          int lo = 0, hi = nchunks;
          while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (offs[mid] <= g) lo = mid;
            else hi = mid;
          }
          const StageEnt& e = stage[((size_t)t * nchunks + lo) * mw + (g - offs[lo])];
          A[r] = e.val;
          B[r] = e.det;
          B1[r] = e.det1;
          P[r] = e.slot;
        }
      }
      block_sort_desc3<E, W == 2>(A, B, B1, P, scr);
    }
  } else if (mw <= 256) {
    __shared__ int lcnt[1024];
    bool done = false;
    if constexpr (T * E >= 2048) {
      if (nchunks * kHead <= T * E)
        done = column_heads<W, E, T>(stage + (size_t)t * nchunks * mw, nchunks, mw, offs, lcnt, scr, merge_tails,
                                     A, B, B1, P);
    }
    if (!done)
      column_tournament<W, E, T>(stage + (size_t)t * nchunks * mw, nchunks, mw, offs, lcnt, A, B, B1, P);
  } else {
    const int half = T * E / 2;  // >= mw
    auto fetch = [&](int chunk, int j, uint64_t& a, uint64_t& b, uint64_t& b1, uint32_t& p) {
      a = 0;
      b = 0;
      b1 = 0;
      p = 0xffffffffu;
      if (chunk < nchunks && j < offs[chunk + 1] - offs[chunk]) {
        const StageEnt& e = stage[((size_t)t * nchunks + chunk) * mw + j];
        a = e.val;
        b = e.det;
        b1 = e.det1;
        p = e.slot;
      }
    };
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const int i = threadIdx.x * E + r;
      if (i < half) fetch(0, i, A[r], B[r], B1[r], P[r]);
      else fetch(1, 2 * half - 1 - i, A[r], B[r], B1[r], P[r]);
    }
    block_sort_desc3<E, W == 2>(A, B, B1, P, scr, T * E);
    for (int chunk = 2; chunk < nchunks; ++chunk) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const int i = threadIdx.x * E + r;
        if (i >= half) fetch(chunk, 2 * half - 1 - i, A[r], B[r], B1[r], P[r]);
      }
      block_sort_desc3<E, W == 2>(A, B, B1, P, scr, T * E);
    }
  }
  // First M entries; keep those present in the table (order preserved).
  const int m = n < M ? n : M;
  {
    int c = 0;
#pragma unroll
    for (int r = 0; r < E; ++r) c += (threadIdx.x * E + r < m && P[r] != 0xffffffffu) ? 1 : 0;
    int kept;
    int pos = block_excl_scan(c, wtot, kept);
    if (threadIdx.x == 0) cand_cnt[t] = kept;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if (threadIdx.x * E + r < m && P[r] != 0xffffffffu) {
        cand[(size_t)t * M + pos].det = B[r];
        cand[(size_t)t * M + pos].det1 = B1[r];
        cand[(size_t)t * M + pos].slot = P[r];
        cand[(size_t)t * M + pos].pad = 0;
        ++pos;
      }
    }
  }
  // Exact per-column sums (order-free) and the reinsert of z_t := new_z_t.
  __shared__ i128 sums[T / 64][2];
  {
    i128 nz = 0, xt = 0;
    for (int ch = threadIdx.x; ch < nchunks; ch += T) {
      const int64_t* o = part + ((size_t)t * nchunks + ch) * 4;
      nz += (((i128)o[0]) << 64) | (i128)(uint64_t)o[1];
      xt += (((i128)o[2]) << 64) | (i128)(uint64_t)o[3];
    }
    nz = wave_sum_i128(nz);
    xt = wave_sum_i128(xt);
    if ((threadIdx.x & 63) == 0) {
      sums[threadIdx.x >> 6][0] = nz;
      sums[threadIdx.x >> 6][1] = xt;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    i128 nz = 0, xt = 0;
    for (int w = 0; w < T / 64; ++w) {
      nz += sums[w][0];
      xt += sums[w][1];
    }
    SlotW<W>* sp = &tab[pslot[t]];
    int64_t za, zb;
    tz_split(nz, za, zb);
    sp->za = za;
    sp->zb = zb;
    col_xz[2 * t] = (int64_t)(xt >> 64);
    col_xz[2 * t + 1] = (int64_t)(uint64_t)xt;
    __threadfence();
    unsigned int tk = atomicAdd(&st->ticket, 1u);
    last = (tk == (unsigned int)(k - 1));
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // Iteration counters: wave 0 sums the stripes and clears them; the
  // columns' x.z increments (exact integers) and max |dx| over k <= 256
  // columns, four per lane of wave 0 (order-free sum and maximum).
  // Every load is issued before the first sum (one memory round trip), the
  // state thread 0 updates included.
  long long c_new = 0, c_upd = 0, c_cols = 0;
  i128 xzc = 0;
  double dm = 0.0;
  int64_t s_xh = 0;
  uint64_t s_xl = 0;
  long long s_zs = 0, s_it = 0;
  double s_max = 0.0;
  if (threadIdx.x < 64) {
    long long* cs = ctr + threadIdx.x * kCtrStride;
    c_new = cs[0];
    c_upd = cs[1];
    c_cols = cs[2];
    int64_t xh[kMaxK / 64], xl[kMaxK / 64];
    double dd[kMaxK / 64];
#pragma unroll
    for (int j = 0; j < kMaxK / 64; ++j) {
      const int c = threadIdx.x + 64 * j;
      xh[j] = 0;
      xl[j] = 0;
      dd[j] = 0.0;
      if (c < k) {
        xh[j] = col_xz[2 * c];
        xl[j] = col_xz[2 * c + 1];
        dd[j] = dx[c];
      }
    }
    if (threadIdx.x == 0) {
      s_xh = st->xz_hi;
      s_xl = st->xz_lo;
      s_zs = st->z_size;
      s_it = st->iterations;
      s_max = st->max_size;
    }
    cs[0] = 0;
    cs[1] = 0;
    cs[2] = 0;
#pragma unroll
    for (int j = 0; j < kMaxK / 64; ++j) {
      xzc += (((i128)xh[j]) << 64) | (i128)(uint64_t)xl[j];
      dm = fmax(dm, __builtin_fabs(dd[j]));
    }
    c_new = wave_sum_ll(c_new);
    c_upd = wave_sum_ll(c_upd);
    c_cols = wave_sum_ll(c_cols);
    xzc = wave_sum_i128(xzc);
    for (int off = 32; off >= 1; off >>= 1) dm = fmax(dm, __shfl_xor(dm, off, 64));
  }
  if (threadIdx.x != 0) return;
  i128 xz = (((i128)s_xh) << 64) | (i128)s_xl;
  xz += xzc;
  st->xz_hi = (int64_t)(xz >> 64);
  st->xz_lo = (uint64_t)xz;
  st->z_size = s_zs + c_new;
  st->updates = c_upd;
  st->cols_size = c_cols;
  st->ticket = 0;
  st->k_prev = k;
  st->dx_max = dm;
  if (count_iteration) st->iterations = s_it + 1;
  if ((double)(s_zs + c_new) > 0.79 * s_max) st->error = ERR_TABLE_FULL;
}

// ------------------------------------------------------------------ k_pick
// gcd_grad pick (src/solver.h:118-224): refetch every candidate's (x, z),
// sort by (|x*s+z| desc, determinant desc), keep the first k distinct.
template <int E, int W>
__global__ __launch_bounds__(1024) void k_pick(DevState* st, int k, int M, const CandEnt* cand,
                                               const int* cand_cnt, const SlotW<W>* tab, D2* pdet,
                                               uint32_t* pslot, double* px, double* pz) {
  __shared__ uint64_t ka[1024 * E];
  __shared__ uint64_t kb[1024 * E];
  __shared__ uint64_t kc[W == 2 ? 1024 * E : 1];
  __shared__ uint32_t pl[1024 * E];
  __shared__ int offs[kMaxK + 2];
  if (st->error) return;
  __shared__ int wtot[17];
  const int kp = st->k_prev;
  {
    // kp <= 256: one column count per thread (the block has at least 64
    // threads; a 64- or 128-thread block scans four counts per thread).
    int v[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = threadIdx.x * 4 + j;
      v[j] = c < kp ? cand_cnt[c] : 0;
      sum += v[j];
    }
    int tot;
    int pre = block_excl_scan(sum, wtot, tot);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = threadIdx.x * 4 + j;
      if (c <= kp) offs[c] = pre;
      pre += v[j];
    }
    if (threadIdx.x == 0) offs[kp] = tot;
  }
  __syncthreads();
  const int n = offs[kp];
  const int cap = blockDim.x * E;
  if (n > cap) {
    if (threadIdx.x == 0) st->error = ERR_CAP;
    return;
  }
  const dd s = load_xss(st);
  // Refetch: candidate g belongs to the column c with offs[c] <= g < offs[c+1].
  uint64_t A[E], B[E], B1[E];
  uint32_t P[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int g = threadIdx.x * E + r;
    A[r] = 0;
    B[r] = 0;
    B1[r] = 0;
    P[r] = 0xffffffffu;
    if (g < n) {
      int lo = 0, hi = kp;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (offs[mid] <= g) lo = mid;
        else hi = mid;
      }
      const CandEnt ce = cand[(size_t)lo * M + (g - offs[lo])];
      const SlotW<W>& sl = tab[ce.slot];
      double z = fix_to_d(tz_join(sl.za, sl.zb));
      A[r] = bitsd(grad_key(s, sl.x, z));
      B[r] = ce.det;
      B1[r] = ce.det1;
      P[r] = ce.slot;
    }
  }
  block_sort_desc3<E, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc});
  __syncthreads();
#pragma unroll
  for (int r = 0; r < E; ++r) {
    kb[threadIdx.x * E + r] = B[r];
    if constexpr (W == 2) kc[threadIdx.x * E + r] = B1[r];
  }
  __syncthreads();
  // Adjacent duplicates (equal determinants sort together) count once.
  int cnt = 0;
  bool first[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    bool dup = i > 0 && kb[i - 1] == B[r];
    if constexpr (W == 2) dup = dup && kc[i - 1] == B1[r];
    first[r] = i < n && !dup;
    cnt += first[r] ? 1 : 0;
  }
  int distinct;
  int pos = block_excl_scan(cnt, wtot, distinct);
  if (threadIdx.x == 0 && distinct < k) st->error = ERR_CANDIDATES;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    if (!first[r] || pos >= k) continue;
    pdet[pos] = D2{B[r], B1[r]};
    pslot[pos] = P[r];
    const SlotW<W>& sl = tab[P[r]];
    px[pos] = sl.x;
    pz[pos] = fix_to_d(tz_join(sl.za, sl.zb));
    ++pos;
  }
}

// ------------------------------------------------------------------ pick for k*M > 4096
// The first k distinct candidates of the full (key desc, det desc) order,
// found without one global sort: every column list is sorted and cut to
// L = min(M, k) entries (an entry below position k of its own list has k
// distinct determinants above it), then lists are merged pairwise, adjacent
// duplicates dropped and the result cut to k distinct entries, until one
// list remains.  Same result as sorting all candidates at once.  Lists are
// stored k entries apart.
template <int E, int T, int W>
__global__ __launch_bounds__(T) void k_pick_cols(DevState* st, int M, int L, int stride, const CandEnt* cand,
                                                 const int* cand_cnt, const SlotW<W>* tab, PickEnt* out,
                                                 int* out_cnt) {
  __shared__ uint64_t ka[T * E];
  __shared__ uint64_t kb[T * E];
  __shared__ uint64_t kc[W == 2 ? T * E : 1];
  __shared__ uint32_t pl[T * E];
  if (st->error) return;
  const int t = blockIdx.x;
  const int n = t < st->k_prev ? min(cand_cnt[t], M) : 0;
  const dd s = load_xss(st);
  uint64_t A[E], B[E], B1[E];
  uint32_t P[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int g = threadIdx.x * E + r;
    A[r] = 0;
    B[r] = 0;
    B1[r] = 0;
    P[r] = 0xffffffffu;
    if (g < n) {
      const CandEnt ce = cand[(size_t)t * M + g];
      const SlotW<W>& sl = tab[ce.slot];
      const double z = fix_to_d(tz_join(sl.za, sl.zb));
      A[r] = bitsd(grad_key(s, sl.x, z));
      B[r] = ce.det;
      B1[r] = ce.det1;
      P[r] = ce.slot;
    }
  }
  block_sort_desc3<E, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc});
  const int keep = n < L ? n : L;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    if (i < keep) out[(size_t)t * stride + i] = PickEnt{A[r], B[r], B1[r], P[r], 0};
  }
  if (threadIdx.x == 0) out_cnt[t] = keep;
}

// The column sorts fused with the tree's first level: workgroup j refetches
// the candidates of columns C j .. C j + C - 1 into the C parts of its window
// (T * E / C >= M entries each), sorts the window and keeps the first K
// distinct entries -- what a tree node makes of those columns' sorted lists
// cut to min(M, k): an entry below position K of its own column has K
// distinct entries above it.  Output as k_pick_merge's (lists K apart).
template <int E, int T, int W, int C>
__global__ __launch_bounds__(T) void k_pick_cols_merge(DevState* st, int M, int K, const CandEnt* cand,
                                                       const int* cand_cnt, const SlotW<W>* tab, PickEnt* out,
                                                       int* out_cnt) {
  __shared__ uint64_t ka[T * E];
  __shared__ uint64_t kb[T * E];
  __shared__ uint64_t kc[W == 2 ? T * E : 1];
  __shared__ uint32_t pl[T * E];
  __shared__ int wtot[T / 64 + 1];
  if (st->error) return;
  const int j = blockIdx.x;
  constexpr int Q = T * E / C;
  const int kp = st->k_prev;
  const dd s = load_xss(st);
  uint64_t A[E], B[E], B1[E];
  uint32_t P[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    const int c = i / Q, pos = i - c * Q;
    const int t = C * j + c;
    const int n = (t < K && t < kp) ? min(cand_cnt[t], M) : 0;
    A[r] = 0;
    B[r] = 0;
    B1[r] = 0;
    P[r] = 0xffffffffu;
    if (pos < n) {
      const CandEnt ce = cand[(size_t)t * M + pos];
      const SlotW<W>& sl = tab[ce.slot];
      const double z = fix_to_d(tz_join(sl.za, sl.zb));
      A[r] = bitsd(grad_key(s, sl.x, z));
      B[r] = ce.det;
      B1[r] = ce.det1;
      P[r] = ce.slot;
    }
  }
  block_sort_desc3<E, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc});
  __syncthreads();
#pragma unroll
  for (int r = 0; r < E; ++r) {
    kb[threadIdx.x * E + r] = B[r];
    if constexpr (W == 2) kc[threadIdx.x * E + r] = B1[r];
  }
  __syncthreads();
  bool first[E];
  int cnt = 0;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    // Fillers are (0, 0, 0); a real determinant is never zero.
    bool dup = i > 0 && kb[i - 1] == B[r];
    if constexpr (W == 2) dup = dup && kc[i - 1] == B1[r];
    first[r] = (B[r] | B1[r]) != 0 && !dup;
    cnt += first[r] ? 1 : 0;
  }
  int distinct;
  int pos = block_excl_scan(cnt, wtot, distinct);
#pragma unroll
  for (int r = 0; r < E; ++r) {
    if (!first[r]) continue;
    if (pos < K) out[(size_t)j * K + pos] = PickEnt{A[r], B[r], B1[r], P[r], 0};
    ++pos;
  }
  if (threadIdx.x == 0) out_cnt[j] = distinct < K ? distinct : K;
}

// One node of the pick's merge tree: the first K distinct entries of the
// union of F sorted lists (input lists F j .. F j + F - 1).  The window of T * E
// entries holds list q in its q-th part of T * E / F (>= K) entries, reversed
// for odd q, so that each pair of parts is bitonic: the sort's levels from
// T * E / (F / 2) up merge them.  Any fan-in gives the same first K distinct.
// The root (emit = 1) writes the picked coordinates as k_pick_emit does.
template <int E, int T, int W, int F>
__global__ __launch_bounds__(T) void k_pick_merge(DevState* st, int K, const PickEnt* in, const int* in_cnt,
                                                  int nin, PickEnt* out, int* out_cnt, int emit,
                                                  const SlotW<W>* tab, D2* pdet, uint32_t* pslot, double* px,
                                                  double* pz) {
  __shared__ uint64_t ka[T * E];
  __shared__ uint64_t kb[T * E];
  __shared__ uint64_t kc[W == 2 ? T * E : 1];
  __shared__ uint32_t pl[T * E];
  __shared__ int wtot[T / 64 + 1];
  if (st->error) return;
  const int j = blockIdx.x;
  constexpr int Q = T * E / F;  // >= K
  uint64_t A[E], B[E], B1[E];
  uint32_t P[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    const int q = i / Q, pos = i - q * Q;
    const int li = F * j + q;
    const int nq = li < nin ? in_cnt[li] : 0;
    const int e = (q & 1) ? Q - 1 - pos : pos;
    A[r] = 0;
    B[r] = 0;
    B1[r] = 0;
    P[r] = 0xffffffffu;
    if (e < nq) {
      const PickEnt en = in[(size_t)li * K + e];
      A[r] = en.key, B[r] = en.det, B1[r] = en.det1, P[r] = en.slot;
    }
  }
  block_sort_desc3<E, W == 2>(A, B, B1, P, SortScratch{ka, kb, pl, kc}, T * E / (F / 2));
  __syncthreads();
#pragma unroll
  for (int r = 0; r < E; ++r) {
    kb[threadIdx.x * E + r] = B[r];
    if constexpr (W == 2) kc[threadIdx.x * E + r] = B1[r];
  }
  __syncthreads();
  bool first[E];
  int cnt = 0;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = threadIdx.x * E + r;
    // Fillers are (0, 0, 0); a real determinant is never zero.
    bool dup = i > 0 && kb[i - 1] == B[r];
    if constexpr (W == 2) dup = dup && kc[i - 1] == B1[r];
    first[r] = (B[r] | B1[r]) != 0 && !dup;
    cnt += first[r] ? 1 : 0;
  }
  int distinct;
  int pos = block_excl_scan(cnt, wtot, distinct);
  if (emit) {
    if (distinct < K) {
      if (threadIdx.x == 0) st->error = ERR_CANDIDATES;
      return;
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if (!first[r]) continue;
      if (pos < K) {
        pdet[pos] = D2{B[r], B1[r]};
        pslot[pos] = P[r];
        const SlotW<W>& sl = tab[P[r]];
        px[pos] = sl.x;
        pz[pos] = fix_to_d(tz_join(sl.za, sl.zb));
      }
      ++pos;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < E; ++r) {
    if (!first[r]) continue;
    if (pos < K) out[(size_t)j * K + pos] = PickEnt{A[r], B[r], B1[r], P[r], 0};
    ++pos;
  }
  if (threadIdx.x == 0) out_cnt[j] = distinct < K ? distinct : K;
}

template <int W>
__global__ void k_pick_emit(DevState* st, int k, const PickEnt* in, const int* in_cnt, const SlotW<W>* tab,
                            D2* pdet, uint32_t* pslot, double* px, double* pz) {
  if (st->error) return;
  if (in_cnt[0] < k) {
    if (threadIdx.x == 0) st->error = ERR_CANDIDATES;
    return;
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const PickEnt e = in[i];
    pdet[i] = D2{e.det, e.det1};
    pslot[i] = e.slot;
    const SlotW<W>& sl = tab[e.slot];
    px[i] = sl.x;
    pz[i] = fix_to_d(tz_join(sl.za, sl.zb));
  }
}

// ------------------------------------------------------------------ init
// update_x of the reference determinant with dx = 1 (src/solver.h:864-872).
template <int W>
__global__ void k_init_ref(DevState* st, D2 det, SlotW<W>* tab, uint64_t mask, D2* pdet, uint32_t* pslot,
                           double* px, double* pz, double* dx) {
  if (threadIdx.x != 0) return;
  uint64_t b = dhash<W>(det) & mask & ~3ULL;
  long long si = -1;
  bool fresh = false;
  for (uint64_t n = 0; n <= mask && si < 0; n += 4) {
    for (int j = 0; j < 4; ++j) {
      const SlotW<W>& s = tab[b + j];
      bool empty, match;
      if constexpr (W == 1) {
        empty = s.key == 0;
        match = s.key == det.lo;
      } else {
        empty = s.k1 == 0;
        match = !empty && (s.k1 & kHiMask) == det.hi && s.k0 == det.lo;
      }
      if (match || empty) {
        si = (long long)(b + j);
        fresh = empty;
        break;
      }
    }
    b = (b + 4) & mask;
  }
  if (si < 0) {
    st->error = ERR_PROBE;
    return;
  }
  SlotW<W>* s = &tab[si];
  if constexpr (W == 1) {
    s->key = det.lo;
  } else {
    s->k0 = det.lo;
    s->k1 = det.hi | kOcc;
  }
  if (fresh) {
    s->x = 1.0;
    s->za = 0;
    s->zb = 0;
  } else {
    s->x = s->x + 1.0;
  }
  st->x_size = 1;
  st->z_size = 0;
  st->updates = 0;
  st->xx_hi = 1.0;
  st->xx_lo = 0.0;
  st->sc_hi = 1.0;
  st->sc_lo = 0.0;
  st->xz_hi = 0;
  st->xz_lo = 0;
  st->k_cur = 1;
  st->k_prev = 1;
  st->iterations = 0;
  pdet[0] = det;
  pslot[0] = (uint32_t)si;
  px[0] = 0.0;
  pz[0] = 0.0;
  dx[0] = 1.0;
}

// Empty table.  W = 1: all zero.  W = 2: word 0 = kSent (word 1 = 0 marks
// the slot empty).  With 2^32 slots the last slot index is the "absent"
// marker of the staging lists, so that slot is filled with a key no
// determinant of the sector has: all 63 bits set for one word (at most 62
// spin-orbitals), the empty determinant for two (at least one electron).
template <int W>
__global__ void k_table_init(SlotW<W>* tab, uint64_t nslots) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    SlotW<W> s;
    const bool reserved = i == 0xffffffffULL;
    if constexpr (W == 1) {
      s.key = reserved ? kKeyMask : 0;
    } else {
      s.k0 = reserved ? 0 : kSent;
      s.k1 = reserved ? kOcc : 0;
    }
    s.x = 0.0;
    s.za = 0;
    s.zb = 0;
    tab[i] = s;
  }
}

// ------------------------------------------------------------------ checkpoint load
// Inserts n distinct determinants (two words each in det2) with their x and
// z (zA, zB pairs) into an empty table.  Keys are unique, so a slot claimed
// by another thread is never this thread's: no waiting, the probe moves on.
template <int W>
__global__ void k_table_insert(SlotW<W>* tab, uint64_t mask, long long n, const uint64_t* det2, const double* x,
                               const int64_t* z2, int* fail) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const D2 D{det2[2 * i], det2[2 * i + 1]};
    uint64_t b = dhash<W>(D) & mask & ~3ULL;
    SlotW<W>* hit = nullptr;
    for (uint64_t m = 0; m <= mask && !hit; m += 4) {
      for (int jj = 0; jj < 4 && !hit; ++jj) {
        SlotW<W>* sl = &tab[b + jj];
        uint64_t expect = 0;
        if constexpr (W == 1) {
          if (__hip_atomic_compare_exchange_strong(&sl->key, &expect, D.lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
            hit = sl;
        } else {
          if (__hip_atomic_compare_exchange_strong(&sl->k1, &expect, D.hi | kOcc, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            sl->k0 = D.lo;
            hit = sl;
          }
        }
      }
      b = (b + 4) & mask;
    }
    if (!hit) {
      *fail = 1;
      continue;
    }
    hit->x = x[i];
    hit->za = z2[2 * i];
    hit->zb = z2[2 * i + 1];
  }
}

// Slot of every candidate (cand[t * M + i], i < cnt[t]) after a reload.
template <int W>
__global__ void k_cand_resolve(const SlotW<W>* tab, uint64_t mask, CandEnt* cand, const int* cnt, int k, int M,
                               int* fail) {
  const int t = blockIdx.x;
  if (t >= k) return;
  for (int i = threadIdx.x; i < cnt[t]; i += blockDim.x) {
    CandEnt& e = cand[(size_t)t * M + i];
    const long long si = find_slot<W>(tab, mask, D2{e.det, e.det1});
    if (si < 0) *fail = 1;
    else e.slot = (uint32_t)si;
  }
}

// ------------------------------------------------------------------ access ceiling
// Random-slot access rates on the table's own footprint, the ceiling the
// scatter and pass 2 are measured against.  n operations, operation i on
// slot mix64(i + seed) & mask.  mode 0 (insert): relaxed load of the key
// word, CAS of an empty key, one 64-bit atomic add; mode 1 (update): load
// of the key word and one atomic add (the scatter's z update for a
// contribution below 2^-11); mode 2 (read): load of the key, x and both z
// words; mode 3: load of the key word and one 64-bit floating-point atomic
// add (a z kept as a plain double, whose sums would depend on the order of
// the adds); mode 4: load of the key word and two int64 atomic adds (a
// large contribution, both z words).  A per-thread sum of what was read
// goes to *sink so that no load is dead.
template <int W>
__global__ __launch_bounds__(kWG) void k_access_probe(SlotW<W>* tab, uint64_t mask, long long n, int mode,
                                                      uint64_t seed, unsigned long long* sink) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  uint64_t acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    SlotW<W>* sl = &tab[mix64((uint64_t)i + seed) & mask];
    uint64_t* kw;
    if constexpr (W == 1) kw = &sl->key;
    else kw = &sl->k1;
    if (mode == 2) {
      acc += rload(kw) + (uint64_t)sl->za + (uint64_t)sl->zb + bitsd(sl->x);
      continue;
    }
    const uint64_t k = rload(kw);
    acc += k;
    if (mode == 3) {
      // one floating-point atomic add (a z kept as a plain double)
      unsafeAtomicAdd((double*)&sl->za, 1.0);
      continue;
    }
    if (mode == 0 && k == 0) {
      uint64_t expect = 0;
      __hip_atomic_compare_exchange_strong(kw, &expect, ((uint64_t)i & kHiMask) | kOcc, __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the scatter's z update: one int64 add to B, both words for a large
    // contribution (mode 4)
    if (mode == 4) __hip_atomic_fetch_add(&sl->za, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&sl->zb, (int64_t)3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (acc == 0x5a5a5a5a5a5a5a5aULL) atomicAdd(sink, 1ULL);
}

// ------------------------------------------------------------------ test hooks
template <bool MIXED, int W>
__global__ __launch_bounds__(kWG) void k_column_dump(DevHam H, D2 det, double diag, uint64_t* odet,
                                                     double* oval, unsigned long long* count, long long cap) {
  __shared__ ColCtxW<W> C;
  setup_ctx<W>(C, H, det);
  for (int r = 0; r < kEPT; ++r) {
    int e = blockIdx.x * kChunk + r * kWG + threadIdx.x;
    if (e >= H.n_enum) continue;
    D2 D;
    double V;
    bool valid;
    if (e == 0) {
      D = det;
      V = diag;
      valid = true;
    } else {
      valid = gen_elem<MIXED, W>(H, C, e, D, V);
    }
    if (!valid) continue;
    unsigned long long pos = atomicAdd(count, 1ULL);
    if ((long long)pos < cap) {
      odet[2 * pos] = D.lo;
      odet[2 * pos + 1] = D.hi;
      oval[pos] = V;
    }
  }
}

template <int W>
__global__ void k_table_compact(const SlotW<W>* tab, uint64_t nslots, uint64_t* odet, double* ox, int64_t* oz,
                                unsigned long long* count, long long cap) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    if (i == 0xffffffffULL) continue;
    const SlotW<W> s = tab[i];
    D2 d;
    if constexpr (W == 1) {
      if (s.key == 0) continue;
      d = D2{s.key & kKeyMask, 0};
    } else {
      if (s.k1 == 0) continue;
      d = D2{s.k0, s.k1 & kHiMask};
    }
    unsigned long long pos = atomicAdd(count, 1ULL);
    if ((long long)pos < cap) {
      odet[2 * pos] = d.lo;
      odet[2 * pos + 1] = d.hi;
      ox[pos] = s.x;
      oz[2 * pos] = s.za;
      oz[2 * pos + 1] = s.zb;
    }
  }
}

// FP64 Rayleigh quotient numerator of the table's x: sum over the
// determinants D with x != 0 (dets, xs; c = x * scale) and every element
// D' of D's column of c_D * H_{D'D} * c_D', with H_{D'D} from the fp64
// integrals (gen_elem<false>, get_diagonal's sequential order for D' = D)
// whatever precision the context runs in.  Each product is rounded once to
// the 2^-64 fixed point and summed exactly, so the result does not depend
// on the visiting order.  One workgroup per (column chunk, determinant).
template <int W>
__global__ __launch_bounds__(kWG) void k_rayleigh(DevHam H, const SlotW<W>* tab, uint64_t mask, const D2* dets,
                                                  const double* xs, long long n, double scale, int64_t* acc) {
  __shared__ ColCtxW<W> C;
  __shared__ i128 red[kWG / 64];
  const int chunk = blockIdx.x;
  i128 local = 0;
  for (long long di = blockIdx.y; di < n; di += gridDim.y) {
    setup_ctx<W>(C, H, dets[di]);
    const double cD = xs[di] * scale;
    for (int r = 0; r < kEPT; ++r) {
      const int e = chunk * kChunk + r * kWG + threadIdx.x;
      if (e >= H.n_enum) continue;
      D2 D;
      double V;
      if (e == 0) {
        double v = 0.0;
        for (int a = 0; a < C.ne; ++a) {
          v += H.d1[C.occ[a]];
          for (int b = a + 1; b < C.ne; ++b) v += H.d2[C.occ[a] * H.ns + C.occ[b]];
        }
        V = v + H.core;
        D = C.det;
      } else if (!gen_elem<false, W>(H, C, e, D, V)) {
        continue;
      }
      const long long si = find_slot<W>(tab, mask, D);
      if (si < 0) continue;
      const double xo = tab[si].x;
      if (!(__builtin_fabs(xo) != 0.0)) continue;
      local += to_fix(cD * V * (xo * scale));
    }
    __syncthreads();  // C is rebuilt for the next determinant
  }
  local = wave_sum_i128(local);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    i128 t = 0;
    for (int w = 0; w < kWG / 64; ++w) t += red[w];
    if (t != 0) {
      int64_t a, b;
      fix_split(t, a, b);
      __hip_atomic_fetch_add(&acc[0], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&acc[1], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Order-independent digest of the table: {occupied slots, sum of key
// hashes, sum of (key, x, z) hashes}, sums modulo 2^64.  The entry hash
// chains mix64 over word 0 (offset by the golden-ratio constant), word 1,
// the bits of x (0 for either zero) and the two halves of the 128-bit
// fixed-point z; the CPU oracle computes the same digest of its table.
__device__ __forceinline__ void digest_entry(const D2& d, double x, i128 z, uint64_t& kh, uint64_t& eh) {
  uint64_t e = mix64(d.lo + 0x9e3779b97f4a7c15ULL);
  e = mix64(e ^ d.hi);
  kh = e;
  e = mix64(e ^ (x == 0.0 ? 0ULL : bitsd(x)));
  e = mix64(e ^ (uint64_t)(z >> 64));
  eh = mix64(e ^ (uint64_t)z);
}

template <int W>
__global__ __launch_bounds__(256) void k_table_digest(const SlotW<W>* tab, uint64_t nslots,
                                                      unsigned long long* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long cnt = 0, ks = 0, es = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    if (i == 0xffffffffULL) continue;
    const SlotW<W> s = tab[i];
    D2 d;
    if constexpr (W == 1) {
      if (s.key == 0) continue;
      d = D2{s.key & kKeyMask, 0};
    } else {
      if (s.k1 == 0) continue;
      d = D2{s.k0, s.k1 & kHiMask};
    }
    uint64_t kh, eh;
    digest_entry(d, s.x, tz_join(s.za, s.zb), kh, eh);
    cnt += 1;
    ks += kh;
    es += eh;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    cnt += __shfl_xor(cnt, off, 64);
    ks += __shfl_xor(ks, off, 64);
    es += __shfl_xor(es, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && cnt) {
    atomicAdd(&out[0], cnt);
    atomicAdd(&out[1], ks);
    atomicAdd(&out[2], es);
  }
}

// ------------------------------------------------------------------ launch wrappers
#define MC_W(W_, ...)      \
  do {                     \
    if ((W_) == 1) {       \
      constexpr int W = 1; \
      __VA_ARGS__;         \
    } else {               \
      constexpr int W = 2; \
      __VA_ARGS__;         \
    }                      \
  } while (0)

void launch_submatrix(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* dets, int k,
                      double* Hs, int ldh) {
  MC_W(nw, hipLaunchKernelGGL(k_submatrix<W>, dim3(k), dim3(kWG), 0, s, H, st, dets, k, Hs, ldh));
}

template <int W, int EPT>
void launch_scatter_e(hipStream_t s, const DevHam& H, DevState* st, const D2* pdet, const double* dx,
                      const double* Hs, int ldh, SlotW<W>* tab, uint64_t mask, uint64_t* vis, int vis_words, int k,
                      long long* ctr, uint32_t* eslot) {
  const int nch = (H.n_enum + kWG * EPT - 1) / (kWG * EPT);
  const int estride = vis_words * 64;  // column elements rounded up to whole chunks
  if (H.mixed)
    hipLaunchKernelGGL((k_scatter<true, W, EPT>), dim3(nch, k), dim3(kWG), 0, s, H, st, pdet, dx, Hs, ldh, tab,
                       mask, vis, vis_words, ctr, eslot, estride);
  else
    hipLaunchKernelGGL((k_scatter<false, W, EPT>), dim3(nch, k), dim3(kWG), 0, s, H, st, pdet, dx, Hs, ldh, tab,
                       mask, vis, vis_words, ctr, eslot, estride);
}

void launch_scatter(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* pdet, const double* dx,
                    const double* Hs, int ldh, void* tab, uint64_t mask, uint64_t* vis, int vis_words, int k,
                    int ept, long long* ctr, uint32_t* eslot) {
  MC_W(nw, {
    SlotW<W>* tb = (SlotW<W>*)tab;
    switch (ept) {
      case 1: launch_scatter_e<W, 1>(s, H, st, pdet, dx, Hs, ldh, tb, mask, vis, vis_words, k, ctr, eslot); break;
      case 2: launch_scatter_e<W, 2>(s, H, st, pdet, dx, Hs, ldh, tb, mask, vis, vis_words, k, ctr, eslot); break;
      case 8: launch_scatter_e<W, 8>(s, H, st, pdet, dx, Hs, ldh, tb, mask, vis, vis_words, k, ctr, eslot); break;
      case 16: launch_scatter_e<W, 16>(s, H, st, pdet, dx, Hs, ldh, tb, mask, vis, vis_words, k, ctr, eslot); break;
      default: launch_scatter_e<W, 4>(s, H, st, pdet, dx, Hs, ldh, tb, mask, vis, vis_words, k, ctr, eslot); break;
    }
  });
}

void launch_pass2(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* pdet, const double* dx,
                  const double* Hs, int ldh, void* tab, uint64_t mask, const uint64_t* vis, int vis_words, int M,
                  int k, int nchunks, StageEnt* stage, int* stage_cnt, int64_t* part, const uint32_t* eslot,
                  uint32_t* znorm_list) {
  // Kept entries per chunk: up to 64 the top-L tournament, up to 256 the
  // top-256 merge, above a full sort.
  const int mw = M < kChunk ? M : kChunk;
  const int sel = mw <= 64 ? 1 : (mw <= 256 ? 2 : 0);
#define MC_P2(MX_, SEL_)                                                                                 \
  hipLaunchKernelGGL((k_pass2<MX_, W, SEL_>), dim3(nchunks, k), dim3(kWG), 0, s, H, st, pdet, dx, Hs, ldh, \
                     (SlotW<W>*)tab, mask, vis, vis_words, M, nchunks, stage, stage_cnt, part, eslot, znorm_list)
  MC_W(nw, {
    if (H.mixed) {
      if (sel == 1) MC_P2(true, 1);
      else if (sel == 2) MC_P2(true, 2);
      else MC_P2(true, 0);
    } else {
      if (sel == 1) MC_P2(false, 1);
      else if (sel == 2) MC_P2(false, 2);
      else MC_P2(false, 0);
    }
  });
#undef MC_P2
}

void launch_merge(hipStream_t s, int nw, DevState* st, const uint32_t* pslot, void* tab, int M, int k,
                  int nchunks, StageEnt* stage, const int* stage_cnt, const int64_t* part, int64_t* col_xz,
                  CandEnt* cand, int* cand_cnt, const double* dx, int count_iteration, long long* ctr,
                  uint64_t nslots, const uint32_t* znorm_list, int merge_tails) {
  // Up to 4096 entries: one sort of all chunk lists (256 threads up to 1024
  // entries, then four entries per thread).  Beyond: the streaming merge
  // with a window of 2 * Mp entries.
  const int mw = M < kChunk ? M : kChunk;
#define MC_MERGE(E_, T_, S_)                                                                                 \
  hipLaunchKernelGGL((k_merge<E_, T_, S_, W>), dim3(k), dim3(T_), 0, s, st, pslot, (SlotW<W>*)tab, M, nchunks, \
                     stage, stage_cnt, part, col_xz, cand, cand_cnt, dx, count_iteration, ctr, nslots, znorm_list, \
                     merge_tails)
  MC_W(nw, {
    if ((long long)nchunks * mw <= kSortCap) {
      int w = kWG;
      while (w < nchunks * mw) w <<= 1;
      switch (w) {
        case 256: MC_MERGE(1, 256, false); break;
        case 512: MC_MERGE(2, 256, false); break;
        case 1024: MC_MERGE(4, 256, false); break;
        case 2048: MC_MERGE(4, 512, false); break;
        default: MC_MERGE(4, 1024, false); break;
      }
    } else if (mw <= 256 && nchunks > 8) {
      // Sixteen waves, one per pair of lists: the first tournament level of
      // up to 32 lists in one round; at one word the heads of up to 32 lists
      // fit the 2048-entry LDS window (two words keep E = 1 within the
      // register budget of 1024 threads).
      if constexpr (W == 1) MC_MERGE(2, 1024, true);
      else MC_MERGE(1, 1024, true);
    } else {
      int w = kWG;
      while (w < 2 * mw) w <<= 1;
      switch (w) {
        case 256: MC_MERGE(1, 256, true); break;
        case 512: MC_MERGE(2, 256, true); break;
        case 1024: MC_MERGE(4, 256, true); break;
        default: MC_MERGE(8, 256, true); break;
      }
    }
  });
#undef MC_MERGE
}

template <int W>
void launch_pick_w(hipStream_t s, DevState* st, int k, int M, const CandEnt* cand, const int* cand_cnt,
                   const SlotW<W>* tb, D2* pdet, uint32_t* pslot, double* px, double* pz, PickEnt* pk0,
                   PickEnt* pk1, int* pc0, int* pc1, bool tree) {
  {
    if (!tree && (long long)k * M <= kSortCap) {
      // One workgroup sorts all candidates: the smallest power of two of at
      // least k * M entries (at least one wave); four entries per thread from
      // 256 entries up.
      int w = 64;
      while (w < k * M) w <<= 1;
      const int E = w >= 256 ? 4 : (w >= 128 ? 2 : 1);
      const int threads = w / E;
      switch (E) {
        case 1:
          hipLaunchKernelGGL((k_pick<1, W>), dim3(1), dim3(threads), 0, s, st, k, M, cand, cand_cnt, tb, pdet,
                             pslot, px, pz);
          break;
        case 2:
          hipLaunchKernelGGL((k_pick<2, W>), dim3(1), dim3(threads), 0, s, st, k, M, cand, cand_cnt, tb, pdet,
                             pslot, px, pz);
          break;
        default:
          hipLaunchKernelGGL((k_pick<4, W>), dim3(1), dim3(threads), 0, s, st, k, M, cand, cand_cnt, tb, pdet,
                             pslot, px, pz);
          break;
      }
    } else {
      const int L = M < k ? M : k;
      // Merge levels (lists of up to k): four lists per node while the window
      // (the next power of two >= fan-in x k) stays within 512 entries (half
      // as many levels; measured faster at k = 64), else two (a 1024-entry
      // sort per node costs more than the levels it saves: k = 256).
      const int fan = 4 * k <= 512 ? 4 : 2;
      int wm = 64;
      while (wm < fan * k) wm <<= 1;
      // Column sorts: width next power of two >= M (<= 1024); with four-list
      // nodes fused with the first merge level (four columns per workgroup)
      // when the window of 4 x that width holds at most 1024 entries and it
      // is not the root (k = 64: pick 27.9 -> 26.4 us; with two-list nodes
      // the 512-entry sorts in half as many workgroups were slower).
      int wc = 64;
      while (wc < M) wc <<= 1;
      const int nfused = (k + fan - 1) / fan;
      int nin = k;
      if (fan == 4 && fan * wc <= 1024 && nfused >= 2) {
        const int win = fan * wc;
#define MC_CM(E_, T_)                                                                                          \
  hipLaunchKernelGGL((k_pick_cols_merge<E_, T_, W, 4>), dim3(nfused), dim3(T_), 0, s, st, M, k, cand, cand_cnt, tb, \
                     pk0, pc0)
        switch (win) {
          case 256: MC_CM(1, 256); break;
          case 512: MC_CM(2, 256); break;
          default: MC_CM(4, 256); break;
        }
#undef MC_CM
        nin = nfused;
      } else {
#define MC_COLS(E_, T_)                                                                                       \
  hipLaunchKernelGGL((k_pick_cols<E_, T_, W>), dim3(k), dim3(T_), 0, s, st, M, L, k, cand, cand_cnt, tb, pk0, pc0)
        switch (wc) {
          case 64: MC_COLS(1, 64); break;
          case 128: MC_COLS(1, 128); break;
          case 256: MC_COLS(1, 256); break;
          case 512: MC_COLS(2, 256); break;
          default: MC_COLS(4, 256); break;
        }
#undef MC_COLS
      }
      PickEnt *in = pk0, *out = pk1;
      int *ic = pc0, *oc = pc1;
      while (nin > 1) {
        const int nout = (nin + fan - 1) / fan;
        const int root = nout == 1 ? 1 : 0;  // the root emits the picked coordinates
#define MC_PM(E_, T_)                                                                                      \
  if (fan == 4)                                                                                            \
    hipLaunchKernelGGL((k_pick_merge<E_, T_, W, 4>), dim3(nout), dim3(T_), 0, s, st, k, in, ic, nin, out, oc, \
                       root, tb, pdet, pslot, px, pz);                                                     \
  else                                                                                                     \
    hipLaunchKernelGGL((k_pick_merge<E_, T_, W, 2>), dim3(nout), dim3(T_), 0, s, st, k, in, ic, nin, out, oc, \
                       root, tb, pdet, pslot, px, pz)
        switch (wm) {
          case 64: MC_PM(1, 64); break;
          case 128: MC_PM(1, 128); break;
          case 256: MC_PM(1, 256); break;
          case 512: MC_PM(2, 256); break;
          default: MC_PM(4, 256); break;
        }
#undef MC_PM
        std::swap(in, out);
        std::swap(ic, oc);
        nin = nout;
      }
      if (k == 1)  // a single list: no merge level, nothing emitted yet
        hipLaunchKernelGGL(k_pick_emit<W>, dim3(1), dim3(256), 0, s, st, k, in, ic, tb, pdet, pslot, px, pz);
    }
  }
}

void launch_pick(hipStream_t s, int nw, DevState* st, int k, int M, const CandEnt* cand, const int* cand_cnt,
                 const void* tab, D2* pdet, uint32_t* pslot, double* px, double* pz, PickEnt* pk0, PickEnt* pk1,
                 int* pc0, int* pc1, bool tree) {
  MC_W(nw, launch_pick_w<W>(s, st, k, M, cand, cand_cnt, (const SlotW<W>*)tab, pdet, pslot, px, pz, pk0, pk1, pc0,
                            pc1, tree));
}

void launch_init_ref(hipStream_t s, int nw, DevState* st, D2 det, void* tab, uint64_t mask, D2* pdet,
                     uint32_t* pslot, double* px, double* pz, double* dx) {
  MC_W(nw, hipLaunchKernelGGL(k_init_ref<W>, dim3(1), dim3(64), 0, s, st, det, (SlotW<W>*)tab, mask, pdet, pslot,
                              px, pz, dx));
}

void launch_access_probe(hipStream_t s, int nw, void* tab, uint64_t mask, long long n, int mode, uint64_t seed,
                         unsigned long long* sink) {
  MC_W(nw, hipLaunchKernelGGL(k_access_probe<W>, dim3(4096), dim3(kWG), 0, s, (SlotW<W>*)tab, mask, n, mode, seed,
                              sink));
}

void launch_table_insert(hipStream_t s, int nw, void* tab, uint64_t mask, long long n, const uint64_t* det2,
                         const double* x, const int64_t* z2, int* fail) {
  const long long blocks = std::min<long long>((n + 255) / 256, 1 << 16);
  if (blocks < 1) return;
  MC_W(nw, hipLaunchKernelGGL(k_table_insert<W>, dim3((unsigned)blocks), dim3(256), 0, s, (SlotW<W>*)tab, mask, n,
                              det2, x, z2, fail));
}

void launch_cand_resolve(hipStream_t s, int nw, const void* tab, uint64_t mask, CandEnt* cand, const int* cnt,
                         int k, int M, int* fail) {
  MC_W(nw, hipLaunchKernelGGL(k_cand_resolve<W>, dim3(k), dim3(256), 0, s, (const SlotW<W>*)tab, mask, cand, cnt,
                              k, M, fail));
}

void launch_table_init(hipStream_t s, int nw, void* tab, uint64_t nslots) {
  const uint64_t blocks = std::min<uint64_t>((nslots + 255) / 256, kSlotBlocks);
  MC_W(nw, hipLaunchKernelGGL(k_table_init<W>, dim3((unsigned)blocks), dim3(256), 0, s, (SlotW<W>*)tab, nslots));
}

void launch_column_dump(hipStream_t s, int nw, const DevHam& H, D2 det, double diag, uint64_t* odet, double* oval,
                        unsigned long long* count, long long cap) {
  const int nchunks = (H.n_enum + kChunk - 1) / kChunk;
  MC_W(nw, {
    if (H.mixed)
      hipLaunchKernelGGL((k_column_dump<true, W>), dim3(nchunks), dim3(kWG), 0, s, H, det, diag, odet, oval, count,
                         cap);
    else
      hipLaunchKernelGGL((k_column_dump<false, W>), dim3(nchunks), dim3(kWG), 0, s, H, det, diag, odet, oval, count,
                         cap);
  });
}

void launch_rayleigh(hipStream_t s, int nw, const DevHam& H, const void* tab, uint64_t mask, const D2* dets,
                     const double* xs, long long n, double scale, int64_t* acc) {
  if (n <= 0) return;
  const int nchunks = (H.n_enum + kChunk - 1) / kChunk;
  const long long gy = n < 65535 ? n : 65535;
  MC_W(nw, hipLaunchKernelGGL(k_rayleigh<W>, dim3(nchunks, (unsigned)gy), dim3(kWG), 0, s, H, (const SlotW<W>*)tab,
                              mask, dets, xs, n, scale, acc));
}

void launch_table_digest(hipStream_t s, int nw, const void* tab, uint64_t nslots, unsigned long long* out) {
  const uint64_t blocks = std::min<uint64_t>((nslots + 255) / 256, kSlotBlocks / 16);
  MC_W(nw, hipLaunchKernelGGL(k_table_digest<W>, dim3((unsigned)blocks), dim3(256), 0, s, (const SlotW<W>*)tab,
                              nslots, out));
}

void launch_table_compact(hipStream_t s, int nw, const void* tab, uint64_t nslots, uint64_t* odet, double* ox,
                          int64_t* oz, unsigned long long* count, long long cap) {
  const uint64_t blocks = std::min<uint64_t>((nslots + 255) / 256, kSlotBlocks);
  MC_W(nw, hipLaunchKernelGGL(k_table_compact<W>, dim3((unsigned)blocks), dim3(256), 0, s, (const SlotW<W>*)tab,
                              nslots, odet, ox, oz, count, cap));
}

}  // namespace mc
