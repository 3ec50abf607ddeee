// Per-point Gaussian-kernel sums (the reference's Grid::Calc_T, Grid.cpp:275-314).
//
//   T(p) = T_init + sum_n dtau_n exp(-3 (dx^2/phix + dy^2/phiy + dz^2/phiz) + expmod_n)
//
// The sum is exp-bound VALU work with O(1) HBM traffic per point, so the
// kernels are built around issue rate, not bandwidth:
//   * FP64 plain: the reference's arithmetic in its order (this file is compiled
//     with -ffp-contract=off so no FMA contraction changes the rounding); node
//     records are wavefront-uniform and come through the scalar cache.
//   * FP32 compensated: each workgroup owns a compact tile of points; node
//     blocks are staged in LDS in the scaled form relative to the tile's
//     middle point (f32sum.hpp), so the per-(point, node) work is 6 FMAs, one
//     v_exp_f32 and an add, plus a Kahan update per 8 nodes. The list form
//     (scattered points) forms each offset in FP64 instead.
#include <mutex>
#include <string>

#include "device.hpp"
#include "f32sum.hpp"

namespace thesis {
namespace {

constexpr int kBlock = 256;     // threads per workgroup
constexpr int kNodeBlock = 256; // nodes staged in LDS per pass

// ------------------------------------------------------------------ FP64 plain

__device__ __forceinline__ double node_term_f64(const NodeF64& nd, double xp, double yp, double zp) {
  const double dx = xp - nd.xb;
  const double dy = yp - nd.yb;
  const double dz = zp - nd.zb;
  const double phi = exp(-3.0 * ((dx * dx * nd.ix) + (dy * dy * nd.iy) + (dz * dz * nd.iz)) + nd.expmod);
  return nd.dtau * phi;
}

__global__ __launch_bounds__(kBlock) void k_tsum_grid_f64(const NodeF64* __restrict__ nodes, int n, GridGeom g,
                                                          int p_begin, int count, double T0,
                                                          double* __restrict__ T) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= count) return;
  double xp, yp, zp;
  point_coords(g, p_begin + q, xp, yp, zp);
  double dT = 0.0;
  for (int m = 0; m < n; ++m) dT += node_term_f64(nodes[m], xp, yp, zp);
  T[q] = T0 + dT;
}

__global__ __launch_bounds__(kBlock) void k_tsum_list_f64(const NodeF64* __restrict__ nodes, int n, GridGeom g,
                                                          const int* __restrict__ list, int count, double T0,
                                                          double* __restrict__ T) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= count) return;
  double xp, yp, zp;
  point_coords(g, list[q], xp, yp, zp);
  double dT = 0.0;
  for (int m = 0; m < n; ++m) dT += node_term_f64(nodes[m], xp, yp, zp);
  T[q] = T0 + dT;
}

// ---------------------------------------------------------- FP32 compensated


__device__ __forceinline__ float exponent_f32(float dx, float dy, float dz, const float4& cf, float c2) {
  float a = fmaf(cf.x * dx, dx, c2);
  a = fmaf(cf.y * dy, dy, a);
  a = fmaf(cf.z * dz, dz, a);
  return a;
}

// n_dev != nullptr: the node count is read from device memory (nodes packed
// on the device), otherwise it is n.
template <int PPT>
__global__ __launch_bounds__(kBlock) void k_tsum_grid_f32c(const NodeF32c* __restrict__ nodes, int n,
                                                           const int* __restrict__ n_dev, GridGeom g,
                                                           int p_begin, int count, double T0,
                                                           double* __restrict__ T) {
  if (n_dev) n = *n_dev;
  __shared__ float4 s_scl[kNodeBlock];
  __shared__ float4 s_ofs[kNodeBlock];
  const int tile0 = blockIdx.x * (kBlock * PPT);
  int i, j, k;
  grid_ijk(g, p_begin + min(tile0 + kBlock * PPT / 2, count - 1), i, j, k);  // origin: the tile's middle point
  const double x0 = grid_x(g, i), y0 = grid_y(g, j), z0 = grid_z(g, k);

  float px[PPT], py[PPT], pz[PPT];
  Kahan acc[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int q = min(tile0 + u * kBlock + (int)threadIdx.x, count - 1);
    grid_ijk(g, p_begin + q, i, j, k);
    px[u] = float(grid_x(g, i) - x0);
    py[u] = float(grid_y(g, j) - y0);
    pz[u] = float(grid_z(g, k) - z0);
  }
  for (int base = 0; base < n; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, n - base);
    __syncthreads();
    for (int m = threadIdx.x; m < pad8(cnt); m += kBlock) stage_scaled(nodes + base, m, cnt, x0, y0, z0, s_scl, s_ofs);
    __syncthreads();
    sum_scaled<PPT>(s_scl, s_ofs, pad8(cnt), px, py, pz, acc);
  }
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int q = tile0 + u * kBlock + (int)threadIdx.x;
    if (q < count) T[q] = T0 + acc[u].value();
  }
}

// list == nullptr: points p_begin + q (custom points, whose scattered
// coordinates get FP64 offsets here rather than the tile-relative form).
__global__ __launch_bounds__(kBlock) void k_tsum_list_f32c(const NodeF32c* __restrict__ nodes, int n, GridGeom g,
                                                           const int* __restrict__ list, int p_begin, int count,
                                                           double T0, double* __restrict__ T) {
  __shared__ double s_ctr[kNodeBlock * 3];
  __shared__ float4 s_coef[kNodeBlock];
  const int q = blockIdx.x * kBlock + threadIdx.x;
  const int qc = min(q, count - 1);
  double xp, yp, zp;
  point_coords(g, list ? list[qc] : p_begin + qc, xp, yp, zp);
  Kahan acc;
  for (int base = 0; base < n; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, n - base);
    __syncthreads();
    for (int m = threadIdx.x; m < cnt; m += kBlock) {
      const NodeF32c nd = nodes[base + m];
      s_ctr[3 * m + 0] = nd.xb;
      s_ctr[3 * m + 1] = nd.yb;
      s_ctr[3 * m + 2] = nd.zb;
      s_coef[m] = make_float4(nd.nx, nd.ny, nd.nz, nd.c);
    }
    __syncthreads();
    for (int m = 0; m < cnt; ++m) {
      const float4 cf = s_coef[m];
      const float dx = float(xp - s_ctr[3 * m + 0]);
      const float dy = float(yp - s_ctr[3 * m + 1]);
      const float dz = float(zp - s_ctr[3 * m + 2]);
      acc.add(__builtin_amdgcn_exp2f(exponent_f32(dx, dy, dz, cf, cf.w)));
    }
  }
  if (q < count) T[q] = T0 + acc.value();
}

inline int blocks_for(long long work, int per_block) { return (int)((work + per_block - 1) / per_block); }

// ------------------------------------------------ allocation-free evaluation
//
// The FP32c sums read node records in the NodeF32c form. The allocation-free
// entry point packs them on the device into a file-scope scratch array: a
// 256-byte header whose first int is the kept-node count, then up to
// kStaticNodes records. Every call on a device shares that device's copy, so
// calls that overlap on one device race.
constexpr int kStaticNodes = 1 << 16;
constexpr size_t kStaticHeader = 256;
constexpr size_t kStaticBytes = kStaticHeader + sizeof(NodeF32c) * kStaticNodes;  // 3 MiB + 256 B
static_assert(kStaticHeader % alignof(NodeF32c) == 0, "node records follow the header");

__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kPackBlock = 1024;

// One workgroup packs nodes [0, n) of the node-major FP64 layout into the
// NodeF32c form, dropping weight-0 nodes and keeping the order of the rest
// (the same records, in the same order, as the host's pack_f32c).
__global__ __launch_bounds__(kPackBlock) void k_pack_f32c(const NodeF64* __restrict__ in, int n,
                                                          unsigned char* __restrict__ buf) {
  __shared__ int s_wave[kPackBlock / 64];
  __shared__ int s_base;
  int* count = reinterpret_cast<int*>(buf);
  NodeF32c* out = reinterpret_cast<NodeF32c*>(buf + kStaticHeader);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double log2e = 1.4426950408889634074;
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += kPackBlock) {
    const int m = c0 + tid;
    NodeF64 a{};
    bool keep = false;
    if (m < n) {
      a = in[m];
      keep = a.dtau > 0.0;  // weight-0 nodes add nothing to T
    }
    const unsigned long long bal = __ballot(keep);
    if (lane == 0) s_wave[wave] = __popcll(bal);
    __syncthreads();
    int before = s_base;
    for (int v = 0; v < wave; ++v) before += s_wave[v];
    before += __popcll(bal & ((1ull << lane) - 1ull));
    if (keep) {
      NodeF32c b;
      b.xb = a.xb;
      b.yb = a.yb;
      b.zb = a.zb;
      b.nx = float(-3.0 * log2e * a.ix);
      b.ny = float(-3.0 * log2e * a.iy);
      b.nz = float(-3.0 * log2e * a.iz);
      b.c = float(log2e * (a.expmod + log(a.dtau)));
      out[before] = b;
    }
    __syncthreads();  // every wave has read s_base and s_wave
    if (tid == 0) {
      int tot = 0;
      for (int v = 0; v < kPackBlock / 64; ++v) tot += s_wave[v];
      s_base += tot;
    }
    __syncthreads();
  }
  if (tid == 0) *count = s_base;
}

}  // namespace

// Device address of `scratch` on the current device, looked up once per device.
static unsigned char* static_scratch() {
  static std::mutex mu;
  static unsigned char* ptr[64] = {};
  int dev = 0;
  THESIS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error(kDevice, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)",
                                        __FILE__, __LINE__);
    ptr[dev] = static_cast<unsigned char*>(p);
  }
  return ptr[dev];
}

int static_node_capacity() { return kStaticNodes; }

void launch_tsum_static(const double* nodes_dev, int n, int precision, const GridGeom& g, int p_begin, int count,
                        double T0, double* T, int ppt, hipStream_t s) {
  if (n < 0 || n > kStaticNodes)
    throw Error(kCapacity, "allocation-free evaluation takes at most " + std::to_string(kStaticNodes) + " nodes");
  if (g.pts) throw Error(kBadArgument, "allocation-free evaluation needs a structured grid (no custom point file)");
  if (count <= 0) return;
  const NodeF64* nd = reinterpret_cast<const NodeF64*>(nodes_dev);
  if (precision == kFP64) {
    hipLaunchKernelGGL(k_tsum_grid_f64, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nd, n, g, p_begin,
                       count, T0, T);
    THESIS_HIP(hipGetLastError());
    return;
  }
  unsigned char* buf = static_scratch();
  hipLaunchKernelGGL(k_pack_f32c, dim3(1), dim3(kPackBlock), 0, s, nd, n, buf);
  THESIS_HIP(hipGetLastError());
  const NodeF32c* packed = reinterpret_cast<const NodeF32c*>(buf + kStaticHeader);
  const int* n_dev = reinterpret_cast<const int*>(buf);
  switch (ppt) {
    case 1:
      hipLaunchKernelGGL(k_tsum_grid_f32c<1>, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, packed, 0, n_dev,
                         g, p_begin, count, T0, T);
      break;
    case 2:
      hipLaunchKernelGGL(k_tsum_grid_f32c<2>, dim3(blocks_for(count, 2 * kBlock)), dim3(kBlock), 0, s, packed, 0,
                         n_dev, g, p_begin, count, T0, T);
      break;
    case 4:
      hipLaunchKernelGGL(k_tsum_grid_f32c<4>, dim3(blocks_for(count, 4 * kBlock)), dim3(kBlock), 0, s, packed, 0,
                         n_dev, g, p_begin, count, T0, T);
      break;
    default:
      throw Error(kBadArgument, "points per thread must be 1, 2 or 4");
  }
  THESIS_HIP(hipGetLastError());
}

void launch_tsum_grid_f64(const NodeF64* nodes, int n, const GridGeom& g, int p_begin, int count, double T0,
                          double* T, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_tsum_grid_f64, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nodes, n, g, p_begin,
                     count, T0, T);
  THESIS_HIP(hipGetLastError());
}

void launch_tsum_list_f64(const NodeF64* nodes, int n, const GridGeom& g, const int* list, int count, double T0,
                          double* T, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_tsum_list_f64, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nodes, n, g, list,
                     count, T0, T);
  THESIS_HIP(hipGetLastError());
}

void launch_tsum_grid_f32c(const NodeF32c* nodes, int n, const GridGeom& g, int p_begin, int count, double T0,
                           double* T, int ppt, hipStream_t s) {
  if (count <= 0) return;
  if (g.pts) {
    hipLaunchKernelGGL(k_tsum_list_f32c, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nodes, n, g, nullptr,
                       p_begin, count, T0, T);
    THESIS_HIP(hipGetLastError());
    return;
  }
  switch (ppt) {
    case 1:
      hipLaunchKernelGGL(k_tsum_grid_f32c<1>, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nodes, n,
                         nullptr, g, p_begin, count, T0, T);
      break;
    case 2:
      hipLaunchKernelGGL(k_tsum_grid_f32c<2>, dim3(blocks_for(count, 2 * kBlock)), dim3(kBlock), 0, s, nodes, n,
                         nullptr, g, p_begin, count, T0, T);
      break;
    case 4:
      hipLaunchKernelGGL(k_tsum_grid_f32c<4>, dim3(blocks_for(count, 4 * kBlock)), dim3(kBlock), 0, s, nodes, n,
                         nullptr, g, p_begin, count, T0, T);
      break;
    default:
      throw Error(kBadArgument, "points per thread must be 1, 2 or 4");
  }
  THESIS_HIP(hipGetLastError());
}

void launch_tsum_list_f32c(const NodeF32c* nodes, int n, const GridGeom& g, const int* list, int count, double T0,
                           double* T, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_tsum_list_f32c, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, nodes, n, g, list, 0,
                     count, T0, T);
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
