// Column context shared by the CDFCI kernels and the energy-gradient
// kernel: determinant decode into occupied / virtual orbital lists and pair
// tables, and the table lookup for one- and two-word determinants.
#pragma once

#include <hip/hip_runtime.h>

#include "det.h"
#include "device.h"

namespace mc {

// Lookup without concurrent inserts (pass 2, the gradient, the host hooks):
// bucketed linear probe from bucket b.  Returns the slot index holding det,
// or -1.
template <int W>
__device__ __forceinline__ long long find_slot(const SlotW<W>* tab, uint64_t mask, const D2& det, uint64_t b) {
  for (uint64_t n = 0; n <= mask; n += 4) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if constexpr (W == 1) {
        const uint64_t k = tab[b + j].key & kKeyMask;
        if (k == det.lo) return (long long)(b + j);
        if (k == 0) return -1;
      } else {
        const uint64_t k1 = tab[b + j].k1;
        if (k1 == 0) return -1;
        if ((k1 & kHiMask) == det.hi && tab[b + j].k0 == det.lo) return (long long)(b + j);
      }
    }
    b = (b + 4) & mask;
  }
  return -1;
}

template <int W>
__device__ __forceinline__ long long find_slot(const SlotW<W>* tab, uint64_t mask, const D2& det) {
  return find_slot<W>(tab, mask, det, dhash<W>(det) & mask & ~3ULL);
}

// The same lookup resumed at slot r0, a slot of det's probe sequence up to
// which every slot held another determinant (so det, if present, sits at r0
// or later: slots are never freed).  The probe sequence is linear in slot
// order, bucket after bucket.  r0 = 0xffffffff: from the start.
template <int W>
__device__ __forceinline__ long long find_slot_at(const SlotW<W>* tab, uint64_t mask, const D2& det, uint32_t r0) {
  if (r0 == 0xffffffffu) return find_slot<W>(tab, mask, det);
  uint64_t s = r0;
  for (uint64_t n = 0; n <= mask; ++n) {
    if constexpr (W == 1) {
      const uint64_t k = tab[s].key & kKeyMask;
      if (k == det.lo) return (long long)s;
      if (k == 0) return -1;
    } else {
      const uint64_t k1 = tab[s].k1;
      if (k1 == 0) return -1;
      if ((k1 & kHiMask) == det.hi && tab[s].k0 == det.lo) return (long long)s;
    }
    s = (s + 1) & mask;
  }
  return -1;
}

// ------------------------------------------------------------------ column context
template <int W>
struct ColCtxW {
  static constexpr int kOrb = W == 1 ? 32 : 64;                   // spatial orbitals per spin list
  static constexpr int kPairs = W == 1 ? 496 : 1953;             // C(kMaxOrb+1, 2) / C(kMaxOrb2, 2)
  D2 det;
  int ne;
  uint8_t occ[2 * kOrb];  // interleaved occupied spin-orbitals, ascending
  uint8_t oa[kOrb], ob[kOrb], va[kOrb], vb[kOrb];
  uint16_t poa[kPairs], pob[kPairs], pva[kPairs], pvb[kPairs];  // pairs x | y << 8, x < y
};
using ColCtx = ColCtxW<1>;

// Pairs x | y << 8 (x < y < n) of list positions, in the order x-major,
// y ascending; entry of pair (x, y) at x * n - x * (x + 1) / 2 + y - x - 1.
__device__ __forceinline__ void fill_pairs_row(uint16_t* out, int n, int x) {
  if (x >= n) return;
  const int base = x * n - x * (x + 1) / 2;
  for (int y = x + 1; y < n; ++y) out[base + (y - x - 1)] = (uint16_t)(x | (y << 8));
}

// Decodes det into the context.  Every spin-orbital s is placed by one
// thread at its rank: occ[#occupied below s], and in its spin's occupied or
// virtual list at #(same-spin occupied below s) or #(same-spin virtual
// below s); then the four pair lists are filled side by side, one row x
// per thread (a 256-thread workgroup covers all four lists at once).
template <int W>
__device__ inline void setup_ctx(ColCtxW<W>& C, const DevHam& H, const D2& det) {
  const int tid = threadIdx.x;
  const D2 alpha = dand(det, D2{0x5555555555555555ULL, 0x5555555555555555ULL});
  const D2 beta = dand(det, D2{0xAAAAAAAAAAAAAAAAULL, 0xAAAAAAAAAAAAAAAAULL});
  for (int s = tid; s < H.ns; s += blockDim.x) {
    const bool o = dtest<W>(det, s);
    const int p = s >> 1;
    if (o) C.occ[dbelow<W>(det, s)] = (uint8_t)s;
    const int same = dbelow<W>((s & 1) ? beta : alpha, s);  // same-spin occupied below s
    if (s & 1) {
      if (o) C.ob[same] = (uint8_t)p;
      else C.vb[p - same] = (uint8_t)p;
    } else {
      if (o) C.oa[same] = (uint8_t)p;
      else C.va[p - same] = (uint8_t)p;
    }
  }
  if (tid == 0) {
    C.det = det;
    C.ne = dpopc<W>(det);
  }
  __syncthreads();
  for (int r = tid; r < 4 * 64; r += blockDim.x) {
    const int list = r >> 6, x = r & 63;
    if (list == 0) fill_pairs_row(C.poa, H.na, x);
    else if (list == 1) fill_pairs_row(C.pob, H.nb, x);
    else if (list == 2) fill_pairs_row(C.pva, H.nva, x);
    else fill_pairs_row(C.pvb, H.nvb, x);
  }
  __syncthreads();
}

}  // namespace mc
