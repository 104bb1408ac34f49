// Column context shared by the CDFCI kernels and the energy-gradient
// kernel: determinant decode into occupied / virtual orbital lists and pair
// tables, the fermionic sign of one excitation, and the table probe.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace mc {

__device__ __forceinline__ int parity1(uint64_t d, int i, int a) {
  uint64_t m = ((1ULL << i) - 1) ^ ((1ULL << a) - 1);
  return __popcll(d & m) & 1;
}

// Bucketed linear probe. Returns the slot index holding det, or -1.
__device__ __forceinline__ long long find_slot(const Slot* tab, uint64_t mask, uint64_t det) {
  uint64_t b = mix64(det) & mask & ~3ULL;
  for (uint64_t n = 0; n <= mask; n += 4) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint64_t k = tab[b + j].key & kKeyMask;
      if (k == det) return (long long)(b + j);
      if (k == 0) return -1;
    }
    b = (b + 4) & mask;
  }
  return -1;
}

// ------------------------------------------------------------------ column context
struct ColCtx {
  uint64_t det;
  int ne;
  uint8_t occ[64];  // interleaved occupied spin-orbitals, ascending
  uint8_t oa[32], ob[32], va[32], vb[32];
  uint16_t poa[496], pob[496], pva[496], pvb[496];  // pairs x | y << 8, x < y
};

__device__ inline void fill_pairs(uint16_t* out, int n, int tid, int nthreads) {
  for (int x = tid; x < n; x += nthreads) {
    int base = x * n - x * (x + 1) / 2;
    for (int y = x + 1; y < n; ++y) out[base + (y - x - 1)] = (uint16_t)(x | (y << 8));
  }
}

__device__ inline void setup_ctx(ColCtx& C, const DevHam& H, uint64_t det) {
  if (threadIdx.x == 0) {
    C.det = det;
    int ne = 0, na = 0, nb = 0, nva = 0, nvb = 0;
    for (int s = 0; s < H.ns; ++s) {
      bool o = (det >> s) & 1;
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
