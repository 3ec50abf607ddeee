// Batched snapshots: many single-beam parameter sets and scan paths of uneven
// length, sharing one grid and material, solved in one persistent launch.
//
// 1. k_count_nodes: one thread per set walks the interval schedule
//    (quadrature.hpp) and counts the nodes the reference would keep.
// 2. k_emit_nodes: one wavefront per set walks the schedule again, parks one
//    interval per lane, and every 64 intervals the lanes write their nodes at
//    offsets from a wavefront exclusive scan -- same order as the host
//    builder (instantaneous node, then intervals newest first, abscissae high
//    to low). Nodes are written in both layouts.
// 3. k_tsum_batched: a persistent grid pulls (set, tile) work items from an
//    atomic counter, most expensive sets first; every wave exits when the
//    queue is drained. Tiles are staged like k_tsum_grid_f32c (FP64
//    tile-relative node centres rounded once, FP32 exponent, v_exp_f32, Kahan).
#include <cmath>

#include "batch.hpp"
#include "f32sum.hpp"

namespace thesis {
namespace {

// The quadrature table in constant memory: a per-thread copy indexed at run
// time would live in scratch memory.
__constant__ const GLTable kGLDev = gl_table();


constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kNodeBlock = 256;

// Counting walker: mirrors EmitAll but only counts kept nodes.
struct CountKept {
  const PathSeg* P;
  const BeamQ& b;
  const RBox& rb;
  double alpha, t;
  const GLTable& gl;
  int n = 0;
  __device__ void instant(int seg) {
    RawNode r;
    if (instant_node(P, seg, b, rb, t, r)) ++n;
  }
  __device__ void interval(double t1, double t2, int seg, int order) {
    for (int a = order - 1; a >= 0; --a) {
      RawNode r;
      if (interval_node(P, seg, b, rb, alpha, t, t1, t2, order, a, gl, r)) ++n;
    }
  }
};

__global__ void k_count_nodes(BatchDev bd, int* counts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bd.nsets) return;
  const PathSeg* P = bd.segs + bd.seg_off[s];
  const int nseg = bd.seg_off[s + 1] - bd.seg_off[s];
  const double t = bd.times[s];
  const GLTable& gl = kGLDev;
  const int seg = locate_segment(P, nseg, t, 1);
  CountKept c{P, bd.beams[s], bd.rbox[s], bd.alpha, t, gl};
  walk_intervals(P, nseg, bd.beams[s], t, false, seg, c);
  counts[s] = c.n;
}

__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int u = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += u;
  }
  return incl - v;
}

__device__ __forceinline__ void write_node(const RawNode& r, NodeF64* o64, NodeF32c* o32) {
  NodeF64 n;
  n.xb = r.xb;
  n.yb = r.yb;
  n.zb = r.zb;
  n.ix = 1.0 / r.px;
  n.iy = 1.0 / r.py;
  n.iz = 1.0 / r.pz;
  n.dtau = r.dtau;
  n.expmod = node_expmod(r);
  *o64 = n;
  const double log2e = 1.4426950408889634074;
  NodeF32c c;
  c.xb = n.xb;
  c.yb = n.yb;
  c.zb = n.zb;
  c.nx = float(-3.0 * log2e * n.ix);
  c.ny = float(-3.0 * log2e * n.iy);
  c.nz = float(-3.0 * log2e * n.iz);
  // weight-0 nodes contribute exactly 0 through exp2(-inf)
  c.c = n.dtau > 0.0 ? float(log2e * (n.expmod + log(n.dtau))) : -INFINITY;
  *o32 = c;
}

struct WaveEmit {
  const PathSeg* P;
  const BeamQ& b;
  const RBox& rb;
  double alpha, t;
  const GLTable& gl;
  NodeF64* o64;
  NodeF32c* o32;
  int lane;
  int base = 0;  // next output index
  int cnt = 0;
  double b_t1 = 0, b_t2 = 0;
  int b_seg = 0, b_order = 0;

  __device__ void instant(int seg) {
    RawNode r;
    const bool keep = instant_node(P, seg, b, rb, t, r);  // uniform
    if (keep) {
      if (lane == 0) write_node(r, o64 + base, o32 + base);
      ++base;
    }
  }
  __device__ void interval(double t1, double t2, int seg, int order) {
    if (lane == cnt) {
      b_t1 = t1;
      b_t2 = t2;
      b_seg = seg;
      b_order = order;
    }
    if (++cnt == kWave) flush();
  }
  __device__ void flush() {
    int mine = 0;
    if (lane < cnt)
      for (int a = b_order - 1; a >= 0; --a) {
        RawNode r;
        if (interval_node(P, b_seg, b, rb, alpha, t, b_t1, b_t2, b_order, a, gl, r)) ++mine;
      }
    const int off = wave_excl_scan(mine, lane);
    const int total = __shfl(off + mine, kWave - 1, kWave);
    if (lane < cnt) {
      int k = base + off;
      for (int a = b_order - 1; a >= 0; --a) {
        RawNode r;
        if (interval_node(P, b_seg, b, rb, alpha, t, b_t1, b_t2, b_order, a, gl, r)) {
          write_node(r, o64 + k, o32 + k);
          ++k;
        }
      }
    }
    base += total;
    cnt = 0;
  }
};

__global__ __launch_bounds__(kBlock) void k_emit_nodes(BatchDev bd, const int* __restrict__ node_off,
                                                       NodeF64* __restrict__ o64, NodeF32c* __restrict__ o32) {
  const int s = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
  if (s >= bd.nsets) return;
  const PathSeg* P = bd.segs + bd.seg_off[s];
  const int nseg = bd.seg_off[s + 1] - bd.seg_off[s];
  const double t = bd.times[s];
  const GLTable& gl = kGLDev;
  const int seg = locate_segment(P, nseg, t, 1);
  WaveEmit e{P, bd.beams[s], bd.rbox[s], bd.alpha, t, gl, o64 + node_off[s], o32 + node_off[s],
             (int)(threadIdx.x & (kWave - 1))};
  walk_intervals(P, nseg, bd.beams[s], t, false, seg, e);
  e.flush();
}


template <int PPT>
__global__ __launch_bounds__(kBlock) void k_tsum_batched_f32c(BatchSumArgs a) {
  __shared__ float4 s_scl[kNodeBlock];
  __shared__ float4 s_ofs[kNodeBlock];
  __shared__ int s_item;
  const int pnum = a.g.nx * a.g.ny * a.g.nz;
  const int tile_pts = kBlock * PPT;
  const int tiles = (pnum + tile_pts - 1) / tile_pts;
  const long long total = (long long)a.nsets * tiles;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(a.queue, 1);
    __syncthreads();
    const long long item = s_item;
    if (item >= total) break;  // every wave of the block sees the same item
    const int set = a.order[item / tiles];
    const int tile0 = (int)(item % tiles) * tile_pts;
    const int count = min(tile_pts, pnum - tile0);
    const NodeF32c* nodes = a.n32 + a.node_off[set];
    const int n = a.node_off[set + 1] - a.node_off[set];
    int i, j, k;
    grid_ijk(a.g, tile0 + min(tile_pts / 2, count - 1), i, j, k);  // origin: the tile's middle point
    const double x0 = grid_x(a.g, i), y0 = grid_y(a.g, j), z0 = grid_z(a.g, k);
    float px[PPT], py[PPT], pz[PPT];
    Kahan acc[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int q = tile0 + min(u * kBlock + (int)threadIdx.x, count - 1);
      grid_ijk(a.g, q, i, j, k);
      px[u] = float(grid_x(a.g, i) - x0);
      py[u] = float(grid_y(a.g, j) - y0);
      pz[u] = float(grid_z(a.g, k) - z0);
    }
    for (int base = 0; base < n; base += kNodeBlock) {
      const int cnt = min(kNodeBlock, n - base);
      __syncthreads();
      for (int m = threadIdx.x; m < pad8(cnt); m += kBlock) stage_scaled(nodes + base, m, cnt, x0, y0, z0, s_scl, s_ofs);
      __syncthreads();
      sum_scaled<PPT>(s_scl, s_ofs, pad8(cnt), px, py, pz, acc);
    }
    double* T = a.T + (long long)set * pnum + tile0;
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int q = u * kBlock + (int)threadIdx.x;
      if (q < count) T[q] = a.T0 + acc[u].value();
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_tsum_batched_f64(BatchSumArgs a) {
  __shared__ int s_item;
  const int pnum = a.g.nx * a.g.ny * a.g.nz;
  const int tiles = (pnum + kBlock - 1) / kBlock;
  const long long total = (long long)a.nsets * tiles;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(a.queue, 1);
    __syncthreads();
    const long long item = s_item;
    if (item >= total) break;
    const int set = a.order[item / tiles];
    const int q = (int)(item % tiles) * kBlock + (int)threadIdx.x;
    if (q >= pnum) continue;
    const NodeF64* nodes = a.n64 + a.node_off[set];
    const int n = a.node_off[set + 1] - a.node_off[set];
    int i, j, k;
    grid_ijk(a.g, q, i, j, k);
    const double xp = grid_x(a.g, i), yp = grid_y(a.g, j), zp = grid_z(a.g, k);
    double dT = 0.0;
    for (int m = 0; m < n; ++m) {
      const NodeF64 nd = nodes[m];
      const double dx = xp - nd.xb, dy = yp - nd.yb, dz = zp - nd.zb;
      dT += nd.dtau * exp(-3.0 * ((dx * dx * nd.ix) + (dy * dy * nd.iy) + (dz * dz * nd.iz)) + nd.expmod);
    }
    a.T[(long long)set * pnum + q] = a.T0 + dT;
  }
}

}  // namespace

void launch_count_nodes(const BatchDev& bd, int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_count_nodes, dim3((bd.nsets + 63) / 64), dim3(64), 0, s, bd, counts);
  THESIS_HIP(hipGetLastError());
}

void launch_emit_nodes(const BatchDev& bd, const int* node_off, NodeF64* o64, NodeF32c* o32, hipStream_t s) {
  const int waves_per_block = kBlock / kWave;
  hipLaunchKernelGGL(k_emit_nodes, dim3((bd.nsets + waves_per_block - 1) / waves_per_block), dim3(kBlock), 0, s, bd,
                     node_off, o64, o32);
  THESIS_HIP(hipGetLastError());
}

void launch_tsum_batched(const BatchSumArgs& a, int precision, int ppt, int blocks, hipStream_t s) {
  if (precision == kFP64) {
    hipLaunchKernelGGL(k_tsum_batched_f64, dim3(blocks), dim3(kBlock), 0, s, a);
  } else if (ppt == 1) {
    hipLaunchKernelGGL(k_tsum_batched_f32c<1>, dim3(blocks), dim3(kBlock), 0, s, a);
  } else if (ppt == 2) {
    hipLaunchKernelGGL(k_tsum_batched_f32c<2>, dim3(blocks), dim3(kBlock), 0, s, a);
  } else if (ppt == 4) {
    hipLaunchKernelGGL(k_tsum_batched_f32c<4>, dim3(blocks), dim3(kBlock), 0, s, a);
  } else {
    throw Error(kBadArgument, "points per thread must be 1, 2 or 4");
  }
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
