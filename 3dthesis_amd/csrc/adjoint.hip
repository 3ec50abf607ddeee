// Adjoint of the snapshot temperature field with respect to the node set.
//
// For a loss L(T) with w_p = dL/dT_p given on the device, this computes for
// every node n the seven sensitivities
//   S_x = sum_p w_p E_pn 6 dx/phix             (dL/dxb, likewise y, z)
//   S_fx = sum_p w_p E_pn (3 dx^2/phix^2 - 1/(2 phix))   (dL/dphix, likewise y, z)
//   S_a = sum_p w_p E_pn                       (dL/d ln(dtau qmod beta))
// where E_pn = dtau exp(-3 sum dx^2/phi + expmod) is the node's term in T_p.
// The host chains them through the quadrature schedule to the beam and path
// parameters (design.cpp).
//
// Work split: one lane per node (64 nodes per wavefront, 256 per workgroup),
// one workgroup per (node block, point chunk). Point offsets from the chunk
// origin and w_p are staged in LDS 256 at a time and read as broadcasts; the
// node centres are taken relative to the same origin in FP64 and rounded once,
// so the per-(point, node) work is FP32 (exp via v_exp_f32). Each 256-point
// slice is summed in FP32 and carried into FP64 accumulators; chunk partials
// go to HBM and a second kernel adds them up per node.
#include "adjoint.hpp"

namespace thesis {
namespace {

constexpr int kBlock = 256;
constexpr int kSlice = 256;

__global__ __launch_bounds__(kBlock) void k_adj_partial(const NodeF64* __restrict__ nodes, int n, GridGeom g,
                                                        int pnum, int chunk, const double* __restrict__ w,
                                                        double* __restrict__ partial) {
  __shared__ float4 s_pt[kSlice];  // point offset from the chunk origin (x, y, z) and w_p
  const int m = blockIdx.x * kBlock + threadIdx.x;
  const int c = blockIdx.y;
  const int p0 = c * chunk;
  const int p1 = min(pnum, p0 + chunk);
  int i, j, k;
  grid_ijk(g, p0, i, j, k);
  const double x0 = grid_x(g, i), y0 = grid_y(g, j), z0 = grid_z(g, k);
  const double log2e = 1.4426950408889634074;
  float bx = 0.f, by = 0.f, bz = 0.f, ix = 0.f, iy = 0.f, iz = 0.f, cc = -INFINITY;
  if (m < n) {
    const NodeF64 nd = nodes[m];
    bx = float(nd.xb - x0);
    by = float(nd.yb - y0);
    bz = float(nd.zb - z0);
    ix = float(nd.ix);
    iy = float(nd.iy);
    iz = float(nd.iz);
    cc = nd.dtau > 0.0 ? float(log2e * (nd.expmod + log(nd.dtau))) : -INFINITY;
  }
  const float kx = float(-3.0 * log2e) * ix, ky = float(-3.0 * log2e) * iy, kz = float(-3.0 * log2e) * iz;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int s0 = p0; s0 < p1; s0 += kSlice) {
    const int cnt = min(kSlice, p1 - s0);
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const int p = s0 + threadIdx.x;
      grid_ijk(g, p, i, j, k);
      s_pt[threadIdx.x] = make_float4(float(grid_x(g, i) - x0), float(grid_y(g, j) - y0), float(grid_z(g, k) - z0),
                                      float(w[p]));
    }
    __syncthreads();
    float a[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < cnt; ++q) {
      const float4 pt = s_pt[q];
      const float dx = pt.x - bx, dy = pt.y - by, dz = pt.z - bz;
      float e = fmaf(kx * dx, dx, cc);
      e = fmaf(ky * dy, dy, e);
      e = fmaf(kz * dz, dz, e);
      const float wE = pt.w * __builtin_amdgcn_exp2f(e);
      const float tx = ix * dx, ty = iy * dy, tz = iz * dz;
      a[0] = fmaf(wE, tx, a[0]);
      a[1] = fmaf(wE, ty, a[1]);
      a[2] = fmaf(wE, tz, a[2]);
      a[3] = fmaf(wE, fmaf(3.f * tx, tx, -0.5f * ix), a[3]);
      a[4] = fmaf(wE, fmaf(3.f * ty, ty, -0.5f * iy), a[4]);
      a[5] = fmaf(wE, fmaf(3.f * tz, tz, -0.5f * iz), a[5]);
      a[6] += wE;
    }
#pragma unroll
    for (int r = 0; r < 7; ++r) acc[r] += double(a[r]);
  }
  if (m < n) {
    double* out = partial + ((size_t)c * n + m) * 7;
    out[0] = 6.0 * acc[0];
    out[1] = 6.0 * acc[1];
    out[2] = 6.0 * acc[2];
#pragma unroll
    for (int r = 3; r < 7; ++r) out[r] = acc[r];
  }
}

__global__ void k_adj_reduce(const double* __restrict__ partial, int n, int nchunks, double* __restrict__ S) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 7) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += partial[(size_t)c * n * 7 + idx];
  S[idx] = s;
}

}  // namespace

void launch_adjoint(const NodeF64* nodes, int n, const GridGeom& g, int pnum, const double* w, int chunk,
                    double* partial, double* S, hipStream_t s) {
  if (n <= 0) return;
  const int nchunks = (pnum + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_adj_partial, dim3((n + kBlock - 1) / kBlock, nchunks), dim3(kBlock), 0, s, nodes, n, g, pnum,
                     chunk, w, partial);
  THESIS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_adj_reduce, dim3((n * 7 + 255) / 256), dim3(256), 0, s, partial, n, nchunks, S);
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
