// Device helpers shared by the window trackers (Volume and Surface).
#pragma once

#include "window.hpp"

namespace thesis {
namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Record the first point met outside the boxes (later ones are dropped).
__device__ __forceinline__ void note_miss(int* s_miss, int* s_mp, int i, int j, int k) {
  if (atomicExch(s_miss, 1) == 0) {
    s_mp[0] = i;
    s_mp[1] = j;
    s_mp[2] = k;
  }
}

}  // namespace
}  // namespace thesis
