// Determinant words on the device: one 64-bit word (up to 31 spatial
// orbitals) or two (up to 63), in the reference's interleaved layout --
// spin-orbital 2p is alpha of spatial orbital p, 2p+1 beta; bit s%64 of word
// s/64 (src/determinant.h:24-100, src/input.h:242-243).
//
// Table keys.  W = 1: the key word is the determinant with bit 63 marking an
// insert of the running iteration (a determinant uses at most 62 bits).
// W = 2: word 1 carries the claim -- bits 0..61 the determinant's word 1,
// bit 62 "occupied" (so that a determinant whose word 1 is 0 still claims a
// slot) and bit 63 the insert mark -- and word 0 is published after the
// claiming CAS; an empty slot holds kSent in word 0, a value no determinant
// of the sector takes (it needs alpha and beta counts of at least 32).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "num.h"

namespace mc {

constexpr uint64_t kMark = 0x8000000000000000ULL;
constexpr uint64_t kKeyMask = ~kMark;
constexpr uint64_t kOcc = 0x4000000000000000ULL;
constexpr uint64_t kHiMask = ~(kMark | kOcc);
constexpr uint64_t kSent = ~0ULL;

struct D2 {
  uint64_t lo, hi;
};

__device__ __forceinline__ bool deq(const D2& a, const D2& b) { return a.lo == b.lo && a.hi == b.hi; }

// Table hash (the slot layout is not observable; any fixed mix will do).
template <int W>
__device__ __forceinline__ uint64_t dhash(const D2& d) {
  if constexpr (W == 1) return mix64(d.lo);
  return mix64(d.lo ^ mix64(d.hi + 0x9e3779b97f4a7c15ULL));
}

template <int W>
__device__ __forceinline__ void dflip(D2& d, int s) {
  if (W == 1 || s < 64) d.lo ^= 1ULL << (s & 63);
  else d.hi ^= 1ULL << (s & 63);
}

template <int W>
__device__ __forceinline__ bool dtest(const D2& d, int s) {
  if (W == 1 || s < 64) return (d.lo >> (s & 63)) & 1;
  return (d.hi >> (s & 63)) & 1;
}

// Number of set bits below position s (0 <= s <= 64 * W).
template <int W>
__device__ __forceinline__ int dbelow(const D2& d, int s) {
  if (W == 1 || s <= 64) return s >= 64 ? __popcll(d.lo) : __popcll(d.lo & ((1ULL << s) - 1));
  return __popcll(d.lo) + __popcll(d.hi & ((1ULL << (s - 64)) - 1));
}

// Parity of the occupied spin-orbitals in [min(i,a), max(i,a))
// (src/determinant.h:102-111, called with i already cleared).
template <int W>
__device__ __forceinline__ int dparity(const D2& d, int i, int a) {
  if constexpr (W == 1) {
    const uint64_t m = ((1ULL << i) - 1) ^ ((1ULL << a) - 1);
    return __popcll(d.lo & m) & 1;
  }
  const int lo = i < a ? i : a, hi = i < a ? a : i;
  return (dbelow<W>(d, hi) - dbelow<W>(d, lo)) & 1;
}

template <int W>
__device__ __forceinline__ int dpopc(const D2& d) {
  return W == 1 ? __popcll(d.lo) : __popcll(d.lo) + __popcll(d.hi);
}

// Lowest set bit (-1 when none) and its removal.
template <int W>
__device__ __forceinline__ int dffs(const D2& d) {
  if (d.lo) return __ffsll((long long)d.lo) - 1;
  if (W == 2 && d.hi) return 64 + __ffsll((long long)d.hi) - 1;
  return -1;
}

template <int W>
__device__ __forceinline__ void dpop_low(D2& d) {
  if (d.lo) d.lo &= d.lo - 1;
  else if (W == 2) d.hi &= d.hi - 1;
}

__device__ __forceinline__ D2 dand(const D2& a, const D2& b) { return D2{a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ D2 dxor(const D2& a, const D2& b) { return D2{a.lo ^ b.lo, a.hi ^ b.hi}; }

}  // namespace mc
