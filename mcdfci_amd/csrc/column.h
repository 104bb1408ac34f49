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

__device__ inline void fill_pairs(uint16_t* out, int n, int tid, int nthreads) {
  for (int x = tid; x < n; x += nthreads) {
    int base = x * n - x * (x + 1) / 2;
    for (int y = x + 1; y < n; ++y) out[base + (y - x - 1)] = (uint16_t)(x | (y << 8));
  }
}

template <int W>
__device__ inline void setup_ctx(ColCtxW<W>& C, const DevHam& H, const D2& det) {
  if (threadIdx.x == 0) {
    C.det = det;
    int ne = 0, na = 0, nb = 0, nva = 0, nvb = 0;
    for (int s = 0; s < H.ns; ++s) {
      bool o = dtest<W>(det, s);
      int p = s >> 1;
      if (o) C.occ[ne++] = (uint8_t)s;
      if (s & 1) {
        if (o) C.ob[nb++] = (uint8_t)p;
        else C.vb[nvb++] = (uint8_t)p;
      } else {
        if (o) C.oa[na++] = (uint8_t)p;
        else C.va[nva++] = (uint8_t)p;
      }
    }
    C.ne = ne;
  }
  __syncthreads();
  fill_pairs(C.poa, H.na, threadIdx.x, blockDim.x);
  fill_pairs(C.pob, H.nb, threadIdx.x, blockDim.x);
  fill_pairs(C.pva, H.nva, threadIdx.x, blockDim.x);
  fill_pairs(C.pvb, H.nvb, threadIdx.x, blockDim.x);
  __syncthreads();
}

}  // namespace mc
