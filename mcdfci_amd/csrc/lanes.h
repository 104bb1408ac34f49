// Exchange with lane (lane ^ M) of the same wave without the LDS crossbar:
// lane-swap instructions across the 32- and 16-lane halves, a DPP row
// rotation for 8, bank-masked DPP row shifts for 4 and DPP quad
// permutations for 2 and 1.  Every lane of the wave must be active.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mc {

template <int M>
__device__ __forceinline__ int lane_xor_i32(int x) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "lane mask");
  if constexpr (M == 32) {
    const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (threadIdx.x & 32) ? p[0] : p[1];
  } else if constexpr (M == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (threadIdx.x & 16) ? p[0] : p[1];
  } else if constexpr (M == 8) {
    return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (M == 4) {
    const int t = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0x5, false);  // row_shl:4, banks 0, 2
    return __builtin_amdgcn_update_dpp(t, x, 0x114, 0xF, 0xA, false);         // row_shr:4, banks 1, 3
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  }
}

template <int M>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  return (uint32_t)lane_xor_i32<M>((int)v);
}

template <int M>
__device__ __forceinline__ uint64_t lane_xor(uint64_t v) {
  const uint32_t lo = lane_xor<M>((uint32_t)v), hi = lane_xor<M>((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

template <int M>
__device__ __forceinline__ int64_t lane_xor(int64_t v) {
  return (int64_t)lane_xor<M>((uint64_t)v);
}

template <int M>
__device__ __forceinline__ double lane_xor(double v) {
  return __longlong_as_double((long long)lane_xor<M>((uint64_t)__double_as_longlong(v)));
}

template <int M>
struct LaneMask {
  static constexpr int value = M;
};

// Calls f(LaneMask<m>{}) for a wave-uniform m in {1, 2, 4, 8, 16, 32}.
template <class F>
__device__ __forceinline__ void with_lane_mask(int m, F&& f) {
  switch (m) {
    case 1: f(LaneMask<1>{}); break;
    case 2: f(LaneMask<2>{}); break;
    case 4: f(LaneMask<4>{}); break;
    case 8: f(LaneMask<8>{}); break;
    case 16: f(LaneMask<16>{}); break;
    default: f(LaneMask<32>{}); break;
  }
}

}  // namespace mc
