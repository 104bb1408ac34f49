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
                                               uint32_t (&p)[E], int size) {
  if constexpr (R < E) {
    if constexpr ((R & S) == 0) {
      constexpr int Q = R | S;
      const bool desc = ((int)(threadIdx.x * E + R) & size) == 0;
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
    sort_reg_stage<E, S, R + 1, C3>(a, b, c, p, size);
  }
}

// The in-thread tail of one bitonic level: strides min(size, E)/2 ... 1 in
// order.  The stages run in sequence (no alternative branches whose common
// code the compiler could merge into address-selected accesses).
template <int E, bool C3>
__device__ __forceinline__ void sort_reg_tail(uint64_t (&a)[E], uint64_t (&b)[E], uint64_t (&c)[E],
                                              uint32_t (&p)[E], int size) {
  if constexpr (E >= 16) {
    if (size >= 16) sort_reg_stage<E, 8, 0, C3>(a, b, c, p, size);
  }
  if constexpr (E >= 8) {
    if (size >= 8) sort_reg_stage<E, 4, 0, C3>(a, b, c, p, size);
  }
  if constexpr (E >= 4) {
    if (size >= 4) sort_reg_stage<E, 2, 0, C3>(a, b, c, p, size);
  }
  if constexpr (E >= 2) sort_reg_stage<E, 1, 0, C3>(a, b, c, p, size);
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
    sort_reg_tail<E, C3>(a, b, c, p, size);
  }
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
