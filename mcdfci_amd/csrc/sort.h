// Workgroup bitonic sort with the entries held in registers.
//
// A workgroup of T threads (power of two, at least one wave) sorts T*E
// entries, E per thread in blocked order (entry i = threadIdx.x * E + r),
// descending by the 128-bit key (a, b), or the 192-bit key (a, b, c) when
// C3 is set (two-word determinants); the 32-bit payload p rides along.
// Stages whose partner lies in the same thread are register compare-swaps,
// stages whose partner lies in the same wave exchange through lane
// shuffles, and only strides of a wave's span or more go through LDS (two
// barriers each).  A sorting network over a total order produces one
// sequence of keys, and equal keys carry equal payloads wherever this sort is
// used, so the output does not depend on the network.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mc {

struct SortScratch {
  uint64_t* a;
  uint64_t* b;
  uint32_t* p;
  uint64_t* c = nullptr;  // third key word (C3 only)
};

__device__ __forceinline__ bool key_gt(uint64_t a0, uint64_t b0, uint64_t a1, uint64_t b1) {
  return a0 > a1 || (a0 == a1 && b0 > b1);
}

template <bool C3>
__device__ __forceinline__ bool key_gt3(uint64_t a0, uint64_t b0, uint64_t c0, uint64_t a1, uint64_t b1,
                                        uint64_t c1) {
  if constexpr (C3) return a0 > a1 || (a0 == a1 && (b0 > b1 || (b0 == b1 && c0 > c1)));
  return key_gt(a0, b0, a1, b1);
}

// Compare-swap of entries R and R | S (R & S == 0) of this thread, unrolled
// by template recursion so that every register index is a constant (a
// runtime-indexed register array would be placed in scratch memory).
template <int E, int S, int R, bool C3>
__device__ __forceinline__ void sort_reg_stage(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                               uint32_t (&p)[E], int size, int base) {
  if constexpr (R < E) {
    if constexpr ((R & S) == 0) {
      constexpr int Q = R | S;
      const bool desc = ((base + R) & size) == 0;
      const uint64_t ar = a[R], br = b[R], aq = a[Q], bq = b[Q];
      const uint64_t cr = c[R], cq = c[Q];
      const uint32_t pr = p[R], pq = p[Q];
      const bool sw = desc ? key_gt3<C3>(aq, bq, cq, ar, br, cr) : key_gt3<C3>(ar, br, cr, aq, bq, cq);
      a[R] = sw ? aq : ar;
      b[R] = sw ? bq : br;
      p[R] = sw ? pq : pr;
      a[Q] = sw ? ar : aq;
      b[Q] = sw ? br : bq;
      p[Q] = sw ? pr : pq;
      if constexpr (C3) {
        c[R] = sw ? cq : cr;
        c[Q] = sw ? cr : cq;
      }
    }
    sort_reg_stage<E, S, R + 1, C3>(a, b, c, p, size, base);
  }
}

// The in-thread tail of one bitonic level: strides min(size, E)/2 ... 1 in
// order.  The stages run in sequence (no alternative branches whose common
// code the compiler could merge into address-selected accesses).  base is
// the sort index of register 0.
template <int E, bool C3>
__device__ __forceinline__ void sort_reg_tail(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                              uint32_t (&p)[E], int size, int base) {
  if constexpr (E >= 16) {
    if (size >= 16) sort_reg_stage<E, 8, 0, C3>(a, b, c, p, size, base);
  }
  if constexpr (E >= 8) {
    if (size >= 8) sort_reg_stage<E, 4, 0, C3>(a, b, c, p, size, base);
  }
  if constexpr (E >= 4) {
    if (size >= 4) sort_reg_stage<E, 2, 0, C3>(a, b, c, p, size, base);
  }
  if constexpr (E >= 2) sort_reg_stage<E, 1, 0, C3>(a, b, c, p, size, base);
}

// Exchange with the entry `stride` away; the lower index keeps the larger
// key inside descending blocks and the smaller inside ascending ones.
template <bool C3>
__device__ __forceinline__ void sort_keep(uint64_t& a, uint64_t& b, uint64_t& c, uint32_t& p, uint64_t pa,
                                          uint64_t pb, uint64_t pc, uint32_t pp, bool lower, bool desc) {
  const bool take_max = (lower == desc);
  const bool take = take_max ? key_gt3<C3>(pa, pb, pc, a, b, c) : key_gt3<C3>(a, b, c, pa, pb, pc);
  a = take ? pa : a;
  b = take ? pb : b;
  p = take ? pp : p;
  if constexpr (C3) c = take ? pc : c;
}

// One cross-lane stage over blocked register entries: entry r exchanges
// with entry r of lane (lane ^ m) and keeps the larger (lower lane, desc) key.
template <int E, bool C3>
__device__ __forceinline__ void sort_xlane_stage(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                                 uint32_t (&p)[E], int m, bool desc) {
  const int lane = threadIdx.x & 63;
  const bool lower = (lane & m) == 0;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const uint64_t pa = __shfl_xor(a[r], m, 64);
    const uint64_t pb = __shfl_xor(b[r], m, 64);
    const uint64_t pc = C3 ? __shfl_xor(c[r], m, 64) : 0;
    const uint32_t pp = __shfl_xor(p[r], m, 64);
    sort_keep<C3>(a[r], b[r], c[r], p[r], pa, pb, pc, pp, lower, desc);
  }
}

// Bitonic merge (descending) of the L = 64*E entries a wave holds in
// blocked order (entry lane*E + r), given a bitonic input.
template <int E, bool C3>
__device__ __forceinline__ void wave_bitonic_merge(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                                   uint32_t (&p)[E], int L) {
  for (int stride = L >> 1; stride >= E; stride >>= 1) sort_xlane_stage<E, C3>(a, b, c, p, stride / E, true);
  sort_reg_tail<E, C3>(a, b, c, p, L, (int)(threadIdx.x & 63) * E);
}

// first_size = 2 sorts any input; first_size = n runs only the last level,
// which sorts a bitonic input (the first half descending, the second
// ascending) -- the merge of two sorted lists.
template <int E, bool C3>
__device__ __forceinline__ void block_sort_desc3(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                                 uint32_t (&p)[E], SortScratch x, int first_size = 2) {
  const int n = blockDim.x * E;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int base = tid * E;
  for (int size = first_size; size <= n; size <<= 1) {
    for (int stride = size >> 1; stride >= E; stride >>= 1) {
      const bool desc = (base & size) == 0;
      if (stride < 64 * E) {
        const int m = stride / E;
        const bool lower = (lane & m) == 0;
        // size > stride >= E, so every entry of this thread shares (i & size).
#pragma unroll
        for (int r = 0; r < E; ++r) {
          const uint64_t pa = __shfl_xor(a[r], m, 64);
          const uint64_t pb = __shfl_xor(b[r], m, 64);
          const uint64_t pc = C3 ? __shfl_xor(c[r], m, 64) : 0;
          const uint32_t pp = __shfl_xor(p[r], m, 64);
          sort_keep<C3>(a[r], b[r], c[r], p[r], pa, pb, pc, pp, lower, desc);
        }
      } else {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) {
          x.a[base + r] = a[r];
          x.b[base + r] = b[r];
          x.p[base + r] = p[r];
          if constexpr (C3) x.c[base + r] = c[r];
        }
        __syncthreads();
        const bool lower = (base & stride) == 0;
#pragma unroll
        for (int r = 0; r < E; ++r) {
          const int j = (base + r) ^ stride;
          sort_keep<C3>(a[r], b[r], c[r], p[r], x.a[j], x.b[j], C3 ? x.c[j] : 0, x.p[j], lower, desc);
        }
      }
    }
    sort_reg_tail<E, C3>(a, b, c, p, size, base);
  }
}

// One wave sorts its own 64*E entries (entry lane*E + r) descending, every
// stage in registers or lane shuffles: no LDS and no barrier.
template <int E, bool C3>
__device__ __forceinline__ void wave_sort_desc3(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                                uint32_t (&p)[E]) {
  const int lane = threadIdx.x & 63;
  const int base = lane * E;
  for (int size = 2; size <= 64 * E; size <<= 1) {
    for (int stride = size >> 1; stride >= E; stride >>= 1) {
      const bool desc = (base & size) == 0;
      const int m = stride / E;
      const bool lower = (lane & m) == 0;
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const uint64_t pa = __shfl_xor(a[r], m, 64);
        const uint64_t pb = __shfl_xor(b[r], m, 64);
        const uint64_t pc = C3 ? __shfl_xor(c[r], m, 64) : 0;
        const uint32_t pp = __shfl_xor(p[r], m, 64);
        sort_keep<C3>(a[r], b[r], c[r], p[r], pa, pb, pc, pp, lower, desc);
      }
    }
    sort_reg_tail<E, C3>(a, b, c, p, size, base);
  }
}

// Top L (a power of two, at most 64) of two lists X and Y of L entries,
// each sorted descending, into lane i (i < L) of the calling wave, sorted
// descending: lane i keeps the larger of X[i] and Y[L-1-i] (together the
// top L, a bitonic sequence), then a bitonic merge over lanes.  Lanes from L
// up get the filler key (0, 0, 0).
template <bool C3>
__device__ __forceinline__ void wave_merge_top(const SortScratch& x, int xo, int yo, int L, uint64_t& a,
                                               uint64_t& b, uint64_t& c, uint32_t& p) {
  const int lane = threadIdx.x & 63;
  a = 0, b = 0, c = 0, p = 0xffffffffu;
  if (lane < L) {
    const int i = xo + lane, j = yo + L - 1 - lane;
    const uint64_t xc = C3 ? x.c[i] : 0, yc = C3 ? x.c[j] : 0;
    const bool ty = key_gt3<C3>(x.a[j], x.b[j], yc, x.a[i], x.b[i], xc);
    a = ty ? x.a[j] : x.a[i];
    b = ty ? x.b[j] : x.b[i];
    c = ty ? yc : xc;
    p = ty ? x.p[j] : x.p[i];
  }
  for (int m = L >> 1; m >= 1; m >>= 1) {
    const uint64_t pa = __shfl_xor(a, m, 64);
    const uint64_t pb = __shfl_xor(b, m, 64);
    const uint64_t pc = C3 ? __shfl_xor(c, m, 64) : 0;
    const uint32_t pp = __shfl_xor(p, m, 64);
    sort_keep<C3>(a, b, c, p, pa, pb, pc, pp, (lane & m) == 0, true);
  }
}

// Top L (power of two, at most 64) of the workgroup's T*E entries (T = 256,
// four waves, E = 4), sorted descending, into lane i (i < L) of wave 0.  Each
// wave first finds its top L: the L-th largest key K of its 64*E entries,
// built bit by bit from ballots (the largest P with at least L keys >= P),
// stopping as soon as at most S keys are >= P (S = 64 for L <= 32, else 128,
// so that a tie across rank L -- determinants related by spin flip share
// their key -- still stops); those entries (they include the top L under the
// full order, ties on the key included) are compacted one or two per lane
// and sorted across the wave.  A wave with more than 128 entries at or above
// K, or fewer than L nonzero keys, sorts all its entries instead.  Either way
// the wave's top L list is the same.  Then two rounds of pairwise top-L
// merges (waves 0 and 1, then wave 0) reduce the four lists.  The scratch
// needs max(512, 6L) entries.  Returns in every thread; only wave 0's outputs
// are set.
template <int E, bool C3>
__device__ __forceinline__ void block_top_desc3(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                                uint32_t (&p)[E], SortScratch x, int L, uint64_t& oa,
                                                uint64_t& ob, uint64_t& oc, uint32_t& op) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ULL << lane) - 1;
  int nzero = 0;
#pragma unroll
  for (int r = 0; r < E; ++r) nzero += __popcll(__ballot(a[r] != 0));
  uint64_t P = 0;
  int cP = 64 * E;
  const int S = L <= 32 ? 64 : 128;
  if (nzero >= L) {
    for (int bit = 62; bit >= 0 && cP > S; --bit) {
      const uint64_t cand = P | (1ULL << bit);
      int cnt = 0;
#pragma unroll
      for (int r = 0; r < E; ++r) cnt += __popcll(__ballot(a[r] >= cand));
      if (cnt >= L) {
        P = cand;
        cP = cnt;
      }
    }
  }
  // one survivor per lane (f1), two (f2), or the full sort
  const bool f1 = cP <= 64, f2 = !f1 && cP <= 128, filtered = f1 || f2;
  if (filtered) {
    int base = w * 128;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const bool keep = a[r] >= P;
      const uint64_t m = __ballot(keep);
      if (keep) {
        const int i = base + __popcll(m & lt);
        x.a[i] = a[r];
        x.b[i] = b[r];
        x.p[i] = p[r];
        if constexpr (C3) x.c[i] = c[r];
      }
      base += __popcll(m);
    }
  }
  __syncthreads();
  // survivors in blocked order (entry lane * R + r), fillers past cP
  uint64_t sa[2] = {0, 0}, sb[2] = {0, 0}, sc[2] = {0, 0};
  uint32_t sp[2] = {0xffffffffu, 0xffffffffu};
  if (filtered) {
    const int R = f1 ? 1 : 2;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int j = lane * R + r;
      if (r < R && j < cP) {
        const int i = w * 128 + j;
        sa[r] = x.a[i];
        sb[r] = x.b[i];
        sp[r] = x.p[i];
        if constexpr (C3) sc[r] = x.c[i];
      }
    }
    if (f1) {
      uint64_t ra[1] = {sa[0]}, rb[1] = {sb[0]}, rc[1] = {sc[0]};
      uint32_t rp[1] = {sp[0]};
      wave_sort_desc3<1, C3>(ra, rb, rc, rp);
      sa[0] = ra[0], sb[0] = rb[0], sc[0] = rc[0], sp[0] = rp[0];
    } else {
      wave_sort_desc3<2, C3>(sa, sb, sc, sp);
    }
  } else {
    wave_sort_desc3<E, C3>(a, b, c, p);
  }
  __syncthreads();
  if (filtered) {
    const int R = f1 ? 1 : 2;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int j = lane * R + r;
      if (r < R && j < L) {
        const int i = w * L + j;
        x.a[i] = sa[r];
        x.b[i] = sb[r];
        x.p[i] = sp[r];
        if constexpr (C3) x.c[i] = sc[r];
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const int i = lane * E + r;
      if (i < L) {
        x.a[w * L + i] = a[r];
        x.b[w * L + i] = b[r];
        x.p[w * L + i] = p[r];
        if constexpr (C3) x.c[w * L + i] = c[r];
      }
    }
  }
  __syncthreads();
  if (w < 2) {
    wave_merge_top<C3>(x, 2 * w * L, (2 * w + 1) * L, L, oa, ob, oc, op);
    if (lane < L) {
      const int i = 4 * L + w * L + lane;
      x.a[i] = oa;
      x.b[i] = ob;
      x.p[i] = op;
      if constexpr (C3) x.c[i] = oc;
    }
  }
  __syncthreads();
  if (w == 0) wave_merge_top<C3>(x, 4 * L, 5 * L, L, oa, ob, oc, op);
}

// Top 256 of the workgroup's 1024 entries (T = 256, E = 4), sorted
// descending, into wave 0's registers (blocked, entry lane*4 + r).
// First the 256th largest key K of the block, built bit by bit from ballots
// summed over the four waves (the largest P with at least 256 keys >= P),
// stopping once at most 512 keys are >= P.  Those entries (they include the
// top 256 under the full order, ties on the key included) are compacted in
// LDS; with at most 256 of them wave 0 sorts them alone, with at most 512
// waves 0 and 1 each sort 256 and wave 0 keeps the top 256 of the pair (the
// larger of X[i] and Y[255-i], then a bitonic merge).  Otherwise (fewer
// than 256 nonzero keys, or more than 512 entries tied at or above K) every
// wave sorts its 256 entries, waves 1 and 3 hand their lists to waves 0 and
// 2, each keeps the top 256 of its pair, and wave 2 hands its result to
// wave 0.  The scratch needs 512 entries.
template <bool C3>
__device__ __forceinline__ void block_top256_desc3(uint64_t (&a)[4], uint64_t (&b)[4], uint64_t (&c)[4],
                                                   uint32_t (&p)[4], SortScratch x) {
  __shared__ int wc[2][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ULL << lane) - 1;
  auto publish = [&](int region) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = region * 256 + lane * 4 + r;
      x.a[i] = a[r];
      x.b[i] = b[r];
      x.p[i] = p[r];
      if constexpr (C3) x.c[i] = c[r];
    }
  };
  auto absorb = [&](int region) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = region * 256 + 255 - (lane * 4 + r);
      const uint64_t ya = x.a[j], yb = x.b[j], yc = C3 ? x.c[j] : 0;
      const uint32_t yp = x.p[j];
      if (key_gt3<C3>(ya, yb, yc, a[r], b[r], c[r])) {
        a[r] = ya;
        b[r] = yb;
        c[r] = yc;
        p[r] = yp;
      }
    }
    wave_bitonic_merge<4, C3>(a, b, c, p, 256);
  };
  // Block-wide counts: every wave writes its count, one barrier, every
  // thread sums the four (two buffers, so one barrier per count).
  int it = 0;
  auto block_count = [&](int wave_count) {
    const int buf = (it++) & 1;
    if (lane == 0) wc[buf][w] = wave_count;
    __syncthreads();
    return wc[buf][0] + wc[buf][1] + wc[buf][2] + wc[buf][3];
  };
  int nz = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) nz += __popcll(__ballot(a[r] != 0));
  uint64_t P = 0;
  int cP = 1024;
  if (block_count(nz) >= 256) {
    for (int bit = 62; bit >= 0 && cP > 512; --bit) {
      const uint64_t cand = P | (1ULL << bit);
      int cnt = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) cnt += __popcll(__ballot(a[r] >= cand));
      const int tot = block_count(cnt);
      if (tot >= 256) {
        P = cand;
        cP = tot;
      }
    }
  }
  if (cP <= 512) {
    // compact the entries >= P in wave order, then waves 0 (and 1) load
    // 256 each, blocked, fillers past cP
    int sw = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) sw += __popcll(__ballot(a[r] >= P));
    const int buf = (it++) & 1;
    if (lane == 0) wc[buf][w] = sw;
    __syncthreads();
    int base = 0;
    for (int v = 0; v < w; ++v) base += wc[buf][v];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const bool keep = a[r] >= P;
      const uint64_t m = __ballot(keep);
      if (keep) {
        const int i = base + __popcll(m & lt);
        x.a[i] = a[r];
        x.b[i] = b[r];
        x.p[i] = p[r];
        if constexpr (C3) x.c[i] = c[r];
      }
      base += __popcll(m);
    }
    __syncthreads();
    const bool two = cP > 256;
    if (w == 0 || (w == 1 && two)) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = w * 256 + lane * 4 + r;
        a[r] = 0, b[r] = 0, c[r] = 0, p[r] = 0xffffffffu;
        if (j < cP) {
          a[r] = x.a[j];
          b[r] = x.b[j];
          p[r] = x.p[j];
          if constexpr (C3) c[r] = x.c[j];
        }
      }
      wave_sort_desc3<4, C3>(a, b, c, p);
    }
    __syncthreads();
    if (two) {
      if (w == 1) publish(0);
      __syncthreads();
      if (w == 0) absorb(0);
    }
    return;
  }
  wave_sort_desc3<4, C3>(a, b, c, p);
  if (w & 1) publish(w >> 1);
  __syncthreads();
  if (!(w & 1)) absorb(w >> 1);
  __syncthreads();
  if (w == 2) publish(0);
  __syncthreads();
  if (w == 0) absorb(0);
}

template <int E>
__device__ __forceinline__ void block_sort_desc(uint64_t (&a)[E], uint64_t (&b)[E], uint32_t (&p)[E], SortScratch x,
                                                int first_size = 2) {
  uint64_t c[E];
#pragma unroll
  for (int r = 0; r < E; ++r) c[r] = 0;
  block_sort_desc3<E, false>(a, b, c, p, x, first_size);
}

}  // namespace mc
