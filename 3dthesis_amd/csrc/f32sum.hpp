// FP32-compensated node sums over a tile of points (device code only).
//
// A node's exponent is c + sum over axes of n_a (p_a - o_a)^2 with n_a < 0
// (log2 units, NodeF32c). With s_a = sqrt(-n_a) and q_a = -s_a o_a, each axis
// term is -(s_a p_a + q_a)^2: one FMA forms e_a = s_a p_a + q_a and a second
// subtracts e_a^2. Offsets p and o are taken from a tile origin in FP64 and
// q_a is formed in FP64 and rounded once, so the cancellation near a node's
// centre happens before the rounding. The terms of 8 consecutive nodes are
// summed plainly and each 8-node sum enters a Kahan sum.
#pragma once

#include <cmath>

#include "device.hpp"

namespace thesis {

struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return double(s) - double(c); }
};

// Node m of a staged block in the scaled form: scl = (s_x, s_y, s_z, c),
// ofs = (q_x, q_y, q_z, 0) for the origin (x0, y0, z0); m >= cnt pads the
// block with zero terms (s = q = 0, c = -inf).
__device__ __forceinline__ void stage_scaled(const NodeF32c* nodes, int m, int cnt, double x0, double y0, double z0,
                                             float4* s_scl, float4* s_ofs) {
  float4 scl = make_float4(0.f, 0.f, 0.f, -INFINITY), ofs = make_float4(0.f, 0.f, 0.f, 0.f);
  if (m < cnt) {
    const NodeF32c nd = nodes[m];
    const float sx = sqrtf(-nd.nx), sy = sqrtf(-nd.ny), sz = sqrtf(-nd.nz);
    scl = make_float4(sx, sy, sz, nd.c);
    ofs = make_float4(float(-double(sx) * (nd.xb - x0)), float(-double(sy) * (nd.yb - y0)),
                      float(-double(sz) * (nd.zb - z0)), 0.f);
  }
  s_scl[m] = scl;
  s_ofs[m] = ofs;
}

// Nodes padded to a multiple of 8 within a block of kNodeBlock.
__device__ __forceinline__ int pad8(int cnt) { return (cnt + 7) & ~7; }

// Add the staged nodes [0, cnt8) to PPT points' sums.
template <int PPT>
__device__ __forceinline__ void sum_scaled(const float4* s_scl, const float4* s_ofs, int cnt8, const float* px,
                                           const float* py, const float* pz, Kahan* acc) {
  for (int m = 0; m < cnt8; m += 8) {
    float blk[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) blk[u] = 0.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const float4 sc = s_scl[m + v];
      const float4 of = s_ofs[m + v];
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        const float ex = fmaf(sc.x, px[u], of.x), ey = fmaf(sc.y, py[u], of.y), ez = fmaf(sc.z, pz[u], of.z);
        float a = fmaf(-ex, ex, sc.w);
        a = fmaf(-ey, ey, a);
        a = fmaf(-ez, ez, a);
        blk[u] += __builtin_amdgcn_exp2f(a);
      }
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) acc[u].add(blk[u]);
  }
}

}  // namespace thesis
