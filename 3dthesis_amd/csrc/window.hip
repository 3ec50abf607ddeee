// Windowed melt-pool tracking kernels: the box evaluation, the one-workgroup
// tracker and the per-point replay. The contract of each is in window.hpp.
#include <cfloat>

#include "f32sum.hpp"
#include "window.hpp"
#include "window_dev.hpp"

namespace thesis {
namespace {

constexpr int kBlock = 256;
constexpr int kTrackBlock = 1024;
constexpr int kNodeBlock = 256;
constexpr int kSurfRows = 4;  // surface rows per wavefront in pass A (FP32c)

// ------------------------------------------------------------ box evaluation

// FP64: one thread per slot, nodes in the reference's order (Grid.cpp:286-296),
// so every cached T is bit-identical to the reference's Calc_T.
__global__ __launch_bounds__(kBlock) void k_window_eval_f64(const NodeF64* __restrict__ nodes,
                                                            const int* __restrict__ node_off, int row0, WinGeom w,
                                                            GridGeom g, double T0, double T_liq,
                                                            double* __restrict__ Tc,
                                                            unsigned long long* __restrict__ lq) {
  const int slot0 = blockIdx.x * kBlock + threadIdx.x;
  const int r = row0 + blockIdx.y;
  const bool live = slot0 < w.nslots;
  const int slot = live ? slot0 : w.nslots - 1;
  int i, j, k, col;
  win_ijk(w, slot, i, j, k, col);
  const double xp = grid_x(g, i), yp = grid_y(g, j), zp = grid_z(g, k);
  const int m0 = node_off[r], m1 = node_off[r + 1];
  double dT = 0.0;
  for (int m = m0; m < m1; ++m) {
    const NodeF64 nd = nodes[m];
    const double dx = xp - nd.xb;
    const double dy = yp - nd.yb;
    const double dz = zp - nd.zb;
    dT += nd.dtau * exp(-3.0 * ((dx * dx * nd.ix) + (dy * dy * nd.iy) + (dz * dz * nd.iz)) + nd.expmod);
  }
  const double T = T0 + dT;
  // liquid bits: the wavefront's 64 consecutive slots make one word
  const unsigned long long m = __ballot(live && T >= T_liq);
  if (live) Tc[(size_t)r * w.stride + slot] = T;
  if ((threadIdx.x & 63) == 0 && slot0 < w.stride) lq[(size_t)r * (w.stride / 64) + (slot0 >> 6)] = m;
}

// FP32 compensated: a tile of kBlock*PPT consecutive slots per workgroup,
// node blocks staged in LDS in the scaled form relative to the tile's middle
// point (f32sum.hpp).
template <int PPT>
__global__ __launch_bounds__(kBlock) void k_window_eval_f32c(const NodeF32c* __restrict__ nodes,
                                                             const int* __restrict__ node_off, int row0, WinGeom w,
                                                             GridGeom g, double T0, double T_liq,
                                                             double* __restrict__ Tc,
                                                             unsigned long long* __restrict__ lq) {
  __shared__ float4 s_scl[kNodeBlock];
  __shared__ float4 s_ofs[kNodeBlock];
  const int r = row0 + blockIdx.y;
  const int tile0 = blockIdx.x * (kBlock * PPT);
  int i, j, k, col;
  win_ijk(w, min(tile0 + kBlock * PPT / 2, w.nslots - 1), i, j, k, col);
  const double x0 = grid_x(g, i), y0 = grid_y(g, j), z0 = grid_z(g, k);
  float px[PPT], py[PPT], pz[PPT];
  Kahan acc[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    win_ijk(w, min(tile0 + u * kBlock + (int)threadIdx.x, w.nslots - 1), i, j, k, col);
    px[u] = float(grid_x(g, i) - x0);
    py[u] = float(grid_y(g, j) - y0);
    pz[u] = float(grid_z(g, k) - z0);
  }
  const int m0 = node_off[r], n = node_off[r + 1] - m0;
  for (int base = 0; base < n; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, n - base);
    __syncthreads();
    for (int m = threadIdx.x; m < pad8(cnt); m += kBlock) stage_scaled(nodes + m0 + base, m, cnt, x0, y0, z0, s_scl, s_ofs);
    __syncthreads();
    sum_scaled<PPT>(s_scl, s_ofs, pad8(cnt), px, py, pz, acc);
  }
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int slot = tile0 + u * kBlock + (int)threadIdx.x;
    const bool live = slot < w.nslots;
    const double T = T0 + acc[u].value();
    const unsigned long long m = __ballot(live && T >= T_liq);  // 64 consecutive slots per word
    if (live) Tc[(size_t)r * w.stride + slot] = T;
    if ((threadIdx.x & 63) == 0 && slot < w.stride) lq[(size_t)r * (w.stride / 64) + (slot >> 6)] = m;
  }
}

// One wavefront per (cache row, surface word): lane b reads the liquid bit of
// the surface slot of column b and the ballot is the word.
__global__ __launch_bounds__(kBlock) void k_surface_bits(WinGeom w, int row0, const unsigned long long* __restrict__ lq,
                                                         unsigned long long* __restrict__ sl, int sstride) {
  const int q = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int r = row0 + blockIdx.y;
  const int lane = threadIdx.x & 63;
  if (q >= sstride) return;  // whole wavefronts
  bool bit = false;
  if (q < w.nsurfw) {
    int b = 0;
    while (b + 1 < w.nbox && q >= w.b[b + 1].woff) ++b;
    const WinBox& x = w.b[b];
    const int rr = q - x.woff;
    const int li = rr / x.nw, ww = rr - li * x.nw;
    const int nbits = min(64, x.nj - ww * 64);
    if (lane < nbits) {
      const int slot = x.off + (li * x.nj + ww * 64 + lane) * x.nk + (x.nk - 1);
      bit = (lq[(size_t)r * (w.stride / 64) + (slot >> 6)] >> (slot & 63)) & 1ull;
    }
  }
  const unsigned long long m = __ballot(bit);
  if (lane == 0) sl[(size_t)r * sstride + q] = m;
}

// ------------------------------------------- two-pass box sums (Surface)
//
// Surface tracking reads a box's layers below the surface only in columns
// whose surface point is liquid at the step or at the step before: the
// recheck's solidification walk down a column that was liquid (Run.cpp:
// 628-643), calc_depth under a liquid surface point (Melt.cpp:271-317), and
// the T_last refresh of those levels from the step before. So Surface
// windows sum in two passes: pass A every surface slot at every step, which
// also gives the tracker's surface words; pass B, per step, the deeper
// layers of the columns listed from those words. Every temperature either
// pass writes is the one the full box would hold, and the tracker and the
// replay read no other.

__device__ __forceinline__ void surf_word_pos(const WinGeom& w, int q, int& b, int& li, int& ww) {
  b = 0;
  while (b + 1 < w.nbox && q >= w.b[b + 1].woff) ++b;
  const int rr = q - w.b[b].woff;
  li = rr / w.b[b].nw;
  ww = rr - li * w.b[b].nw;
}

// Pass A point of a lane: wavefront = surface word q, lane = bit (clamped to
// a real point of the window for the arithmetic; `live` says whether it is one).
__device__ __forceinline__ void surf_point(const WinGeom& w, int q, int lane, int& i, int& j, int& k, int& slot,
                                           bool& live) {
  const int qc = min(q, w.nsurfw - 1);
  int b, li, ww;
  surf_word_pos(w, qc, b, li, ww);
  const WinBox& x = w.b[b];
  const int nbits = min(64, x.nj - ww * 64);
  live = q < w.nsurfw && lane < nbits;
  const int lj = ww * 64 + min(lane, nbits - 1);
  i = x.i0 + li;
  j = x.j0 + lj;
  k = x.k0 + x.nk - 1;
  slot = x.off + (li * x.nj + lj) * x.nk + x.nk - 1;
}

// FP64 pass A: the reference's arithmetic and order per point (Grid.cpp:286-296).
__global__ __launch_bounds__(kBlock) void k_surf_eval_f64(const NodeF64* __restrict__ nodes,
                                                          const int* __restrict__ node_off, int row0, WinGeom w,
                                                          GridGeom g, double T0, double T_liq, double* __restrict__ Tc,
                                                          unsigned long long* __restrict__ sl, int sstride) {
  const int r = row0 + blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  int i, j, k, slot;
  bool live;
  surf_point(w, q, lane, i, j, k, slot, live);
  const double xp = grid_x(g, i), yp = grid_y(g, j), zp = grid_z(g, k);
  const int m0 = node_off[r], m1 = node_off[r + 1];
  double dT = 0.0;
  for (int m = m0; m < m1; ++m) {
    const NodeF64 nd = nodes[m];
    const double dx = xp - nd.xb;
    const double dy = yp - nd.yb;
    const double dz = zp - nd.zb;
    dT += nd.dtau * exp(-3.0 * ((dx * dx * nd.ix) + (dy * dy * nd.iy) + (dz * dz * nd.iz)) + nd.expmod);
  }
  const double T = T0 + dT;
  if (live) Tc[(size_t)r * w.stride + slot] = T;
  const unsigned long long bits = __ballot(live && T >= T_liq);
  if (lane == 0 && q < sstride) sl[(size_t)r * sstride + q] = bits;
}

// Pass A tasks of box b: its surface rows in groups of R, times its words per
// row.
__device__ __forceinline__ int surf_tasks(const WinBox& x, int R) { return x.nw * ((x.ni + R - 1) / R); }

// FP32c pass A: a wavefront per R surface rows of one box word (rows li ..
// li+R-1 of the box, the same 64 columns j, lane = column). A node's term
// 2^(c - e_x^2 - e_y^2 - e_z^2) factors into 2^(c - e_x^2 - e_z^2), the same
// for every lane of a row, and 2^(-e_y^2), the same for the R rows of a lane.
// The first factor is formed once per node and row while the node block is
// staged (offsets in FP64), the second per node and lane (e_y in the scaled
// form of f32sum.hpp relative to the wavefront's middle column), so a node
// costs a lane one exp and R FMAs. 8-node partial sums enter Kahan sums.
template <int R>
__global__ __launch_bounds__(kBlock) void k_surf_eval_f32c(const NodeF32c* __restrict__ nodes,
                                                           const int* __restrict__ node_off, int row0, WinGeom w,
                                                           GridGeom g, double T0, double T_liq,
                                                           double* __restrict__ Tc,
                                                           unsigned long long* __restrict__ sl, int sstride) {
  constexpr int kWaves = kBlock / 64;
  __shared__ float s_xz[kWaves][kNodeBlock][R];  // 2^(c - e_x^2 - e_z^2) per wavefront, node and row
  __shared__ float2 s_y[kWaves][kNodeBlock];     // (s_y, q_y) per wavefront and node
  __shared__ double s_x[kWaves][R], s_z[kWaves], s_y0[kWaves];
  const int r = row0 + blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // this wavefront's task: box b, rows li0 .. li0+R-1, word ww
  int b = -1, li0 = 0, ww = 0;
  {
    int t = blockIdx.x * kWaves + wv;
    for (int bb = 0; bb < w.nbox; ++bb) {
      const int nt = surf_tasks(w.b[bb], R);
      if (t < nt) {
        b = bb;
        li0 = (t / w.b[bb].nw) * R;
        ww = t % w.b[bb].nw;
        break;
      }
      t -= nt;
    }
  }
  const bool has = b >= 0;
  const WinBox& x = w.b[has ? b : 0];
  const int nbits = has ? min(64, x.nj - ww * 64) : 1;
  const int lj = ww * 64 + min(lane, nbits - 1);
  if (lane == 0) {
#pragma unroll
    for (int u = 0; u < R; ++u) s_x[wv][u] = grid_x(g, x.i0 + min(li0 + u, x.ni - 1));
    s_z[wv] = grid_z(g, x.k0 + x.nk - 1);
    s_y0[wv] = grid_y(g, x.j0 + ww * 64 + min(32, nbits - 1));
  }
  const float py = float(grid_y(g, x.j0 + lj) - grid_y(g, x.j0 + ww * 64 + min(32, nbits - 1)));
  Kahan acc[R];
  const int m0 = node_off[r], n = node_off[r + 1] - m0;
  for (int base = 0; base < n; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, n - base);
    __syncthreads();  // the last block's terms are summed (and, the first time, the rows are in)
    {
      // a thread per node: its factors for every wavefront of the block
      const int m = threadIdx.x;
      NodeF32c nd{};
      if (m < cnt) nd = nodes[m0 + base + m];
      const float sy = m < cnt ? sqrtf(-nd.ny) : 0.f;
#pragma unroll
      for (int v = 0; v < kWaves; ++v) {
        if (m < cnt) {
          const double dz = s_z[v] - nd.zb;
          const double cz = double(nd.c) + double(nd.nz) * dz * dz;
          s_y[v][m] = make_float2(sy, float(-double(sy) * (nd.yb - s_y0[v])));
#pragma unroll
          for (int u = 0; u < R; ++u) {
            const double dx = s_x[v][u] - nd.xb;
            s_xz[v][m][u] = __builtin_amdgcn_exp2f(float(cz + double(nd.nx) * dx * dx));
          }
        } else {
          s_y[v][m] = make_float2(0.f, 0.f);
#pragma unroll
          for (int u = 0; u < R; ++u) s_xz[v][m][u] = 0.f;
        }
      }
    }
    __syncthreads();
    if (has) {
      const int cnt8 = (cnt + 7) & ~7;
      for (int m = 0; m < cnt8; m += 8) {
        float blk[R];
#pragma unroll
        for (int u = 0; u < R; ++u) blk[u] = 0.f;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const float2 yv = s_y[wv][m + v];
          const float ey = fmaf(yv.x, py, yv.y);
          const float e = __builtin_amdgcn_exp2f(-ey * ey);
#pragma unroll
          for (int u = 0; u < R; ++u) blk[u] = fmaf(s_xz[wv][m + v][u], e, blk[u]);
        }
#pragma unroll
        for (int u = 0; u < R; ++u) acc[u].add(blk[u]);
      }
    }
  }
  if (has) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int li = li0 + u;
      if (li >= x.ni) break;  // wavefront-uniform
      const double T = T0 + acc[u].value();
      const bool live = lane < nbits;
      if (live) Tc[(size_t)r * w.stride + x.off + (li * x.nj + lj) * x.nk + x.nk - 1] = T;
      const unsigned long long bits = __ballot(live && T >= T_liq);
      if (lane == 0) sl[(size_t)r * sstride + x.woff + li * x.nw + ww] = bits;
    }
  }
  // the row's padding words past the last box's
  if (blockIdx.x == 0 && threadIdx.x < sstride - w.nsurfw) sl[(size_t)r * sstride + w.nsurfw + threadIdx.x] = 0ull;
}

// Columns to sum in depth, per cache row: surface point liquid at the row's
// step or at the step before (the row before row0 has none), in ascending
// column order. One workgroup per row; cols[r * ccap ..], count in ccount[r].
constexpr int kListBlock = 1024;
__global__ __launch_bounds__(kListBlock) void k_list_columns(WinGeom w, int row0,
                                                             const unsigned long long* __restrict__ sl, int sstride,
                                                             int* __restrict__ cols, int* __restrict__ ccount,
                                                             int ccap) {
  __shared__ int s_wsum[kListBlock / 64];
  __shared__ int s_base;
  const int r = row0 + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int q0 = 0; q0 < w.nsurfw; q0 += kListBlock) {
    const int q = q0 + tid;
    unsigned long long m = 0ull;
    if (q < w.nsurfw) {
      m = sl[(size_t)r * sstride + q];
      if (r > row0) m |= sl[(size_t)(r - 1) * sstride + q];
    }
    const int c = __popcll(m);
    int x = c;  // inclusive prefix within the wavefront
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    int e = s_base + x - c;
    for (int v = 0; v < wave; ++v) e += s_wsum[v];
    if (m) {
      int b, li, ww;
      surf_word_pos(w, q, b, li, ww);
      const WinBox& bx = w.b[b];
      const int col0 = bx.coff + li * bx.nj + ww * 64;
      for (unsigned long long t = m; t; t &= t - 1ull) {
        if (e < ccap) cols[(size_t)r * ccap + e] = col0 + __builtin_ctzll(t);
        ++e;
      }
    }
    __syncthreads();  // every wave has read s_base and s_wsum
    if (tid == 0) {
      int tot = 0;
      for (int v = 0; v < kListBlock / 64; ++v) tot += s_wsum[v];
      s_base += tot;
    }
    __syncthreads();
  }
  if (tid == 0) ccount[r] = min(s_base, ccap);
}

// A listed column's box and box-local position.
__device__ __forceinline__ void col_pos(const WinGeom& w, int col, int& b, int& cc, int& li, int& lj) {
  b = 0;
  while (b + 1 < w.nbox && col >= w.b[b + 1].coff) ++b;
  const WinBox& x = w.b[b];
  cc = col - x.coff;
  li = div_small(cc, x.nj, x.rnj);
  lj = cc - li * x.nj;
}

// FP64 pass B: thread per (listed column, layer below the surface), the
// reference's arithmetic per point; liquid bits OR-ed into zeroed rows.
__global__ __launch_bounds__(kBlock) void k_deep_eval_f64(const NodeF64* __restrict__ nodes,
                                                          const int* __restrict__ node_off, int row0, WinGeom w,
                                                          GridGeom g, double T0, double T_liq,
                                                          const int* __restrict__ cols,
                                                          const int* __restrict__ ccount, int ccap,
                                                          double* __restrict__ Tc,
                                                          unsigned long long* __restrict__ lq) {
  const int r = row0 + blockIdx.y;
  const int nd = w.b[0].nk - 1;  // every Surface box spans the same layers
  const int n = ccount[r] * nd;
  const int it = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= n || nd <= 0) return;
  const int itc = min(it, n - 1);
  const int li_col = itc / nd, lk = itc - li_col * nd;
  int b, cc, li, lj;
  col_pos(w, cols[(size_t)r * ccap + li_col], b, cc, li, lj);
  const WinBox& x = w.b[b];
  const double xp = grid_x(g, x.i0 + li), yp = grid_y(g, x.j0 + lj), zp = grid_z(g, x.k0 + lk);
  const int m0 = node_off[r], m1 = node_off[r + 1];
  double dT = 0.0;
  for (int m = m0; m < m1; ++m) {
    const NodeF64 nd8 = nodes[m];
    const double dx = xp - nd8.xb;
    const double dy = yp - nd8.yb;
    const double dz = zp - nd8.zb;
    dT += nd8.dtau * exp(-3.0 * ((dx * dx * nd8.ix) + (dy * dy * nd8.iy) + (dz * dz * nd8.iz)) + nd8.expmod);
  }
  if (it >= n) return;
  const double T = T0 + dT;
  const int slot = x.off + cc * x.nk + lk;
  Tc[(size_t)r * w.stride + slot] = T;
  if (T >= T_liq) atomicOr(&lq[(size_t)r * (w.stride / 64) + (slot >> 6)], 1ull << (slot & 63));
}

// FP32c pass B: G threads per listed column, kLpt layers each; a node's x and
// y terms are formed once per column and reused for its layers, in the same
// FMA order as sum_scaled. The column's liquid bits are gathered over its G
// lanes and OR-ed into the zeroed row by one of them.
constexpr int kLpt = 4;
template <int G>
__global__ __launch_bounds__(kBlock) void k_deep_eval_f32c(const NodeF32c* __restrict__ nodes,
                                                           const int* __restrict__ node_off, int row0, WinGeom w,
                                                           GridGeom g, double T0, double T_liq,
                                                           const int* __restrict__ cols,
                                                           const int* __restrict__ ccount, int ccap,
                                                           double* __restrict__ Tc,
                                                           unsigned long long* __restrict__ lq) {
  constexpr int kCols = kBlock / G;  // columns per workgroup
  __shared__ float4 s_scl[kNodeBlock];
  __shared__ float4 s_ofs[kNodeBlock];
  const int r = row0 + blockIdx.y;
  const int n = ccount[r];
  const int c0 = blockIdx.x * kCols;
  if (c0 >= n) return;
  const int nd = w.b[0].nk - 1;  // every Surface box spans the same layers
  const int* L = cols + (size_t)r * ccap;
  const int ci = threadIdx.x / G, gi = threadIdx.x % G;
  const bool live = c0 + ci < n;
  int b, cc, li, lj;
  col_pos(w, L[min(c0 + kCols / 2, n - 1)], b, cc, li, lj);  // origin: the middle column, middle layer
  const double x0 = grid_x(g, w.b[b].i0 + li), y0 = grid_y(g, w.b[b].j0 + lj), z0 = grid_z(g, w.b[0].k0 + nd / 2);
  col_pos(w, L[min(c0 + ci, n - 1)], b, cc, li, lj);
  const WinBox& x = w.b[b];
  const float px = float(grid_x(g, x.i0 + li) - x0), py = float(grid_y(g, x.j0 + lj) - y0);
  float pz[kLpt];
#pragma unroll
  for (int u = 0; u < kLpt; ++u) pz[u] = float(grid_z(g, x.k0 + min(gi * kLpt + u, nd - 1)) - z0);
  Kahan acc[kLpt];
  const int m0 = node_off[r], nn = node_off[r + 1] - m0;
  for (int base = 0; base < nn; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, nn - base);
    __syncthreads();
    for (int m = threadIdx.x; m < pad8(cnt); m += kBlock) stage_scaled(nodes + m0 + base, m, cnt, x0, y0, z0, s_scl, s_ofs);
    __syncthreads();
    for (int m = 0; m < pad8(cnt); m += 8) {
      float blk[kLpt];
#pragma unroll
      for (int u = 0; u < kLpt; ++u) blk[u] = 0.f;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const float4 sc = s_scl[m + v];
        const float4 of = s_ofs[m + v];
        const float ex = fmaf(sc.x, px, of.x), ey = fmaf(sc.y, py, of.y);
        const float axy = fmaf(-ey, ey, fmaf(-ex, ex, sc.w));
#pragma unroll
        for (int u = 0; u < kLpt; ++u) {
          const float ez = fmaf(sc.z, pz[u], of.z);
          blk[u] += __builtin_amdgcn_exp2f(fmaf(-ez, ez, axy));
        }
      }
#pragma unroll
      for (int u = 0; u < kLpt; ++u) acc[u].add(blk[u]);
    }
  }
  unsigned long long mask = 0ull;  // bit lk: liquid at layer lk of the column
  const int base_slot = x.off + cc * x.nk;
#pragma unroll
  for (int u = 0; u < kLpt; ++u) {
    const int lk = gi * kLpt + u;
    if (!live || lk >= nd) continue;
    const double T = T0 + acc[u].value();
    Tc[(size_t)r * w.stride + base_slot + lk] = T;
    if (T >= T_liq) mask |= 1ull << lk;
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1) mask |= __shfl_xor(mask, o);
  if (gi == 0 && live && mask) {
    unsigned long long* row = lq + (size_t)r * (w.stride / 64);
    const int w0 = base_slot >> 6, sh = base_slot & 63;
    atomicOr(&row[w0], mask << sh);
    if (sh + nd > 64) atomicOr(&row[w0 + 1], mask >> (64 - sh));
  }
}

// FP32c pass B for columns of at most 4G layers: G threads per listed column,
// kLpt layers each. A node's term factors into 2^(c - e_x^2 - e_y^2), formed
// once per node and column, and 2^(-e_z^2), the same for every column at a
// layer and formed once per node and layer while the node block is staged
// (offsets in FP64). A node then costs a thread one exp and kLpt FMAs.
template <int G>
__global__ __launch_bounds__(kBlock) void k_deep_eval_sep_f32c(const NodeF32c* __restrict__ nodes,
                                                               const int* __restrict__ node_off, int row0, WinGeom w,
                                                               GridGeom g, double T0, double T_liq,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ ccount, int ccap,
                                                               double* __restrict__ Tc,
                                                               unsigned long long* __restrict__ lq) {
  static_assert(kBlock == kNodeBlock, "a thread stages one node");
  constexpr int kCols = kBlock / G;  // columns per workgroup
  constexpr int kL = G * kLpt;       // layers held per node
  __shared__ float4 s_scl[kNodeBlock];
  __shared__ float4 s_ofs[kNodeBlock];
  __shared__ float4 s_ez[kNodeBlock][G];  // 2^(-e_z^2) per node and layer
  const int r = row0 + blockIdx.y;
  const int n = ccount[r];
  const int c0 = blockIdx.x * kCols;
  if (c0 >= n) return;
  const int nd = w.b[0].nk - 1;  // every Surface box spans the same layers
  const int* L = cols + (size_t)r * ccap;
  const int ci = threadIdx.x / G, gi = threadIdx.x % G;
  const bool live = c0 + ci < n;
  int b, cc, li, lj;
  col_pos(w, L[min(c0 + kCols / 2, n - 1)], b, cc, li, lj);  // origin: the middle column
  const double x0 = grid_x(g, w.b[b].i0 + li), y0 = grid_y(g, w.b[b].j0 + lj);
  const int k0 = w.b[0].k0;
  col_pos(w, L[min(c0 + ci, n - 1)], b, cc, li, lj);
  const WinBox& x = w.b[b];
  const float px = float(grid_x(g, x.i0 + li) - x0), py = float(grid_y(g, x.j0 + lj) - y0);
  Kahan acc[kLpt];
  const int m0 = node_off[r], nn = node_off[r + 1] - m0;
  for (int base = 0; base < nn; base += kNodeBlock) {
    const int cnt = min(kNodeBlock, nn - base);
    __syncthreads();
    {
      const int m = threadIdx.x;  // kBlock == kNodeBlock: a thread per node
      stage_scaled(nodes + m0 + base, m, cnt, x0, y0, 0.0, s_scl, s_ofs);
      float e[kL];
      if (m < cnt) {
        const NodeF32c nd8 = nodes[m0 + base + m];
#pragma unroll
        for (int u = 0; u < kL; ++u) {
          const double dz = grid_z(g, k0 + min(u, nd - 1)) - nd8.zb;
          e[u] = __builtin_amdgcn_exp2f(float(double(nd8.nz) * dz * dz));
        }
      } else {
#pragma unroll
        for (int u = 0; u < kL; ++u) e[u] = 0.f;
      }
#pragma unroll
      for (int q = 0; q < G; ++q) s_ez[m][q] = make_float4(e[4 * q], e[4 * q + 1], e[4 * q + 2], e[4 * q + 3]);
    }
    __syncthreads();
    for (int m = 0; m < pad8(cnt); m += 8) {
      float blk[kLpt];
#pragma unroll
      for (int u = 0; u < kLpt; ++u) blk[u] = 0.f;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const float4 sc = s_scl[m + v];
        const float4 of = s_ofs[m + v];
        const float ex = fmaf(sc.x, px, of.x), ey = fmaf(sc.y, py, of.y);
        const float exy = __builtin_amdgcn_exp2f(fmaf(-ey, ey, fmaf(-ex, ex, sc.w)));
        const float4 ez = s_ez[m + v][gi];
        blk[0] = fmaf(exy, ez.x, blk[0]);
        blk[1] = fmaf(exy, ez.y, blk[1]);
        blk[2] = fmaf(exy, ez.z, blk[2]);
        blk[3] = fmaf(exy, ez.w, blk[3]);
      }
#pragma unroll
      for (int u = 0; u < kLpt; ++u) acc[u].add(blk[u]);
    }
  }
  unsigned long long mask = 0ull;  // bit lk: liquid at layer lk of the column
  const int base_slot = x.off + cc * x.nk;
#pragma unroll
  for (int u = 0; u < kLpt; ++u) {
    const int lk = gi * kLpt + u;
    if (!live || lk >= nd) continue;
    const double T = T0 + acc[u].value();
    Tc[(size_t)r * w.stride + base_slot + lk] = T;
    if (T >= T_liq) mask |= 1ull << lk;
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1) mask |= __shfl_xor(mask, o);
  if (gi == 0 && live && mask) {
    unsigned long long* row = lq + (size_t)r * (w.stride / 64);
    const int w0 = base_slot >> 6, sh = base_slot & 63;
    atomicOr(&row[w0], mask << sh);
    if (sh + nd > 64) atomicOr(&row[w0 + 1], mask >> (64 - sh));
  }
}

// ------------------------------------------------------------------ tracker

// Flag bits per slot, live within the tracker. Frontier membership rotates
// over three bits (and three lists) so a wave can clear the list after next
// without a barrier of its own.
enum : uint32_t {
  kFLiqPrev = 0x01,  // liquid after the previous step
  kFLiqNew = 0x02,   // liquid after this step
  kFEval = kEvEval,  // evaluated / claimed this step (the reference's T_calc_flag)
  kFTlast = kEvTlast,
  kFSolid = kEvSolid,
};
__device__ __constant__ const uint32_t kFront[3] = {0x08u, 0x10u, 0x80u};

// Append from the lanes executing the call: one counter atomic per wave.
__device__ __forceinline__ void wave_push(int* cnt, int* list, int cap, int v) {
  const unsigned long long m = __ballot(1);
  const int lane = lane_id();
  const int leader = __ffsll(m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader);
  const int i = base + __popcll(m & ((1ull << lane) - 1ull));
  if (i < cap) list[i] = v;
}

// Lists: 0, 1 liquid (previous / this step, alternating), 2..4 frontiers,
// 5 the slots evaluated this step.
enum { kListTouched = 5 };

struct TrackState {
  uint8_t* F;                    // flag byte per slot (4-byte aligned, padded)
  const unsigned long long* LQ;  // liquid bit per slot of the current step
  short* Dc;          // committed column depth
  short* Dn;          // column depth being built this step
  int* lists;
  int* n;             // list counters
  int cap;
  __device__ __forceinline__ uint32_t set(int slot, uint32_t bits) const {
    uint32_t* wd = reinterpret_cast<uint32_t*>(F + (slot & ~3));
    const int sh = 8 * (slot & 3);
    return (atomicOr(wd, bits << sh) >> sh) & 0xffu;
  }
  __device__ __forceinline__ void clear(int slot, uint32_t bits) const {
    uint32_t* wd = reinterpret_cast<uint32_t*>(F + (slot & ~3));
    atomicAnd(wd, ~(bits << (8 * (slot & 3))));
  }
  __device__ __forceinline__ int* list(int l) const { return lists + l * cap; }
  __device__ __forceinline__ bool liq(int slot) const { return (LQ[slot >> 6] >> (slot & 63)) & 1ull; }
  // a list whose counter passes cap is abandoned for a flag scan
  __device__ __forceinline__ void push(int l, int v) const { wave_push(n + l, lists + l * cap, cap, v); }
  // set bits including kFEval; a slot evaluated for the first time this step joins the touched list
  __device__ __forceinline__ uint32_t eval(int slot, uint32_t bits) const {
    const uint32_t old = set(slot, bits | kFEval);
    if (!(old & kFEval)) push(kListTouched, slot);
    return old;
  }
};

constexpr int kPrefetch = 4;  // uint4 per thread: rows up to 64 KiB are prefetched into registers

template <bool LDS, bool SURF>
__device__ void track_body(const TrackArgs& a, uint8_t* lds) {
  if (a.prev_result && (a.prev_result[kTrDone] != a.prev_nsteps || a.prev_result[kTrStatus] != 0)) {
    if (threadIdx.x == 0) {  // the window before ended early: this one was planned for a start that never came
      for (int q = 0; q < kTrInts; ++q) a.result[q] = 0;
      a.result[kTrStatus] = 4;
    }
    return;
  }
  const WinGeom& w = a.w;
  const GridGeom& g = a.g;
  const int tid = threadIdx.x;
  const int nz = g.nz;
  constexpr int kNb = SURF ? 9 : 18;
  __shared__ int s_miss, s_mp[3], s_stop, s_cnt[2][4], s_done, s_nl, s_ext[6], s_n[kTrackLists];
  // phase clock: wall_clock64 ticks at 100 MHz
  long long tick = a.phase_times ? wall_clock64() : 0;
  long long ph0 = 0, ph1 = 0, ph2 = 0, ph4 = 0, ph5 = 0, ph6 = 0, ph7 = 0;
  int levels = 0, scans = 0, max_front = 0, max_liq = 0;
#define LAP(v)                               \
  do {                                       \
    if (a.phase_times) {                     \
      const long long now_ = wall_clock64(); \
      v += now_ - tick;                      \
      tick = now_;                           \
    }                                        \
  } while (0)

  TrackState S;
  const int fbytes = w.stride;        // flag bytes (stride is a multiple of 128)
  const int lbytes = w.stride / 8;    // liquid bits of one step
  const int cols8 = (w.ncols + 7) & ~7;
  S.cap = a.list_cap;
  S.n = s_n;
  if (LDS) {
    S.F = lds;
    S.LQ = reinterpret_cast<const unsigned long long*>(lds + fbytes);
    S.Dc = reinterpret_cast<short*>(lds + fbytes + lbytes);
    S.Dn = S.Dc + cols8;
    S.lists = reinterpret_cast<int*>(lds + fbytes + lbytes + 4 * cols8);
  } else {
    S.F = a.gflags;
    S.LQ = nullptr;
    S.Dc = a.gdepth;
    S.Dn = a.gdepth + cols8;
    S.lists = a.glists;
  }
  uint8_t* LQbuf = LDS ? lds + fbytes : nullptr;
  const size_t lwords = (size_t)w.stride / 64;
  const int n16 = lbytes / 16;
  const bool prefetch = LDS && n16 <= kPrefetch * kTrackBlock;
  // the next step's liquid row, held in registers while the current step runs
  uint4 pv0 = make_uint4(0, 0, 0, 0), pv1 = pv0, pv2 = pv0, pv3 = pv0;
#define ROW_LOAD(row)                                                   \
  do {                                                                  \
    const uint4* src_ = reinterpret_cast<const uint4*>(row);            \
    if (tid < n16) pv0 = src_[tid];                                     \
    if (tid + kTrackBlock < n16) pv1 = src_[tid + kTrackBlock];         \
    if (tid + 2 * kTrackBlock < n16) pv2 = src_[tid + 2 * kTrackBlock]; \
    if (tid + 3 * kTrackBlock < n16) pv3 = src_[tid + 3 * kTrackBlock]; \
  } while (0)
  int lp = 0;  // list index of the previous step's liquid points

  if (tid == 0) {
    s_miss = 0;
    s_stop = 0;
    s_done = 0;
    for (int l = 0; l < kTrackLists; ++l) s_n[l] = 0;
    for (int l = 0; l < 4; ++l) s_cnt[0][l] = s_cnt[1][l] = 0;
  }
  for (int q = tid; q < fbytes / 4; q += kTrackBlock) reinterpret_cast<uint32_t*>(S.F)[q] = 0u;
  for (int c = tid; c < w.ncols; c += kTrackBlock) {
    int b = 0;
    while (b + 1 < w.nbox && c >= w.b[b + 1].coff) ++b;
    const WinBox& x = w.b[b];
    const int cc = c - x.coff;
    const int li = div_small(cc, x.nj, x.rnj);
    const short d = (short)a.depths[(x.i0 + li) * g.ny + x.j0 + (cc - li * x.nj)];
    S.Dc[c] = d;
    S.Dn[c] = d;
  }
  if (prefetch && a.nsteps >= 1) ROW_LOAD(a.lq + lwords);
  __syncthreads();
  const int n_in = a.prev_result ? a.prev_result[kTrNliq] : a.n_liq_in;  // a queued window takes the last one's list
  for (int q = tid; q < n_in; q += kTrackBlock) {
    int i, j, k;
    grid_ijk(g, a.liq_in[q], i, j, k);
    const int sl = win_slot(w, i, j, k);
    if (sl < 0) {
      note_miss(&s_miss, s_mp, i, j, k);
      continue;
    }
    S.set(sl, kFLiqPrev);
    S.push(lp, sl);
  }
  __syncthreads();
  LAP(ph0);

  for (int r = 1; r <= a.nsteps && !s_miss; ++r) {
    const int ln = 1 - lp;
    if (LDS) {
      uint4* dst = reinterpret_cast<uint4*>(LQbuf);
      if (prefetch) {
        if (tid < n16) dst[tid] = pv0;
        if (tid + kTrackBlock < n16) dst[tid + kTrackBlock] = pv1;
        if (tid + 2 * kTrackBlock < n16) dst[tid + 2 * kTrackBlock] = pv2;
        if (tid + 3 * kTrackBlock < n16) dst[tid + 3 * kTrackBlock] = pv3;
      } else {
        const uint4* src = reinterpret_cast<const uint4*>(a.lq + (size_t)r * lwords);
        for (int q = tid; q < n16; q += kTrackBlock) dst[q] = src[q];
      }
      S.LQ = reinterpret_cast<const unsigned long long*>(LQbuf);
    } else {
      S.LQ = a.lq + (size_t)r * lwords;
    }
    if (tid == 0) {
      s_n[ln] = 0;
      s_n[2] = 0;
      s_n[3] = 0;
      s_n[kListTouched] = 0;
    }
    __syncthreads();
    if (prefetch && r < a.nsteps) ROW_LOAD(a.lq + (size_t)(r + 1) * lwords);  // lands while this step runs
    LAP(ph1);

    // Recheck the previous step's liquid points (Run.cpp:509-535 / :617-654).
    {
      auto recheck = [&](int sl) __attribute__((always_inline)) {
        const bool liq = S.liq(sl);
        const bool front = SURF || liq;  // Surface: the search restarts from every previous liquid point
        S.eval(sl, (liq ? kFLiqNew : kFSolid) | (front ? kFront[0] : 0u));
        if (SURF && !liq) {  // solidify the column beneath (Run.cpp:628-643)
          int b, li, lj, lk;
          win_local(w, sl, b, li, lj, lk);
          const WinBox& x = w.b[b];
          const int col = x.coff + li * x.nj + lj;
          const int d0 = S.Dc[col];
          for (int d = d0; d > 0; --d) {
            if (lk - d < 0) {  // the column's slots are sl - d down to the box floor
              note_miss(&s_miss, s_mp, x.i0 + li, x.j0 + lj, nz - 1 - d);
              break;
            }
            S.eval(sl - d, kFSolid | (d != d0 ? kFTlast : 0u));
          }
          S.Dn[col] = 0;
        }
        if (liq) S.push(ln, sl);
        if (front) S.push(2, sl);
      };
      const int nlp = s_n[lp];
      if (nlp <= S.cap) {
        const int* L = S.list(lp);
        for (int q = tid; q < nlp; q += kTrackBlock) recheck(L[q]);
      } else {
        for (int sl = tid; sl < w.nslots; sl += kTrackBlock)
          if (S.F[sl] & kFLiqPrev) recheck(sl);
      }
    }
    // Beam-trace seeds: unclaimed ones are claimed and evaluated (Melt.cpp:98-183).
    for (int q = a.seed_off[r - 1] + tid; q < a.seed_off[r]; q += kTrackBlock) {
      int i, j, k;
      grid_ijk(g, a.seeds[q], i, j, k);
      const int sl = win_slot(w, i, j, k);
      if (sl < 0) {
        note_miss(&s_miss, s_mp, i, j, k);
        continue;
      }
      const uint32_t old = S.eval(sl, 0u);
      if (!(old & kFEval) && S.liq(sl)) {
        S.set(sl, kFLiqNew | kFront[0]);
        S.push(ln, sl);
        S.push(2, sl);
      }
    }
    __syncthreads();
    LAP(ph2);

    // Neighbour waves (Melt::neighbor_check, Melt.cpp:185-269): wave L reads
    // frontier L % 3, fills (L + 1) % 3 and clears (L + 2) % 3, whose last
    // readers finished before wave L - 1's barrier. One thread per frontier
    // point visits its neighbours.
    for (int level = 0;; ++level) {
      const int c = level % 3, nx = (level + 1) % 3, zz = (level + 2) % 3;
      const uint32_t nbit = kFront[nx];
      if (tid == 0) s_n[2 + zz] = 0;
      auto expand = [&](int sl) __attribute__((always_inline)) {
        int b, li, lj, lk;
        win_local(w, sl, b, li, lj, lk);
        const WinBox& x = w.b[b];
#pragma unroll 3
        for (int e = 0; e < kNb; ++e) {
          const int di = (e % 3) - 1, dj = (e / 3) % 3 - 1, dk = SURF ? 0 : e / 9 - 1;
          const int ni = x.i0 + li + di, nj = x.j0 + lj + dj, nk = x.k0 + lk + dk;
          if (ni < 0 || ni > g.nx - 1 || nj < 0 || nj > g.ny - 1 || nk < 0) continue;
          int s2;
          if (li + di >= 0 && li + di < x.ni && lj + dj >= 0 && lj + dj < x.nj && lk + dk >= 0)
            s2 = sl + (di * x.nj + dj) * x.nk + dk;
          else
            s2 = win_slot(w, ni, nj, nk);
          if (s2 < 0) {
            note_miss(&s_miss, s_mp, ni, nj, nk);
            continue;
          }
          const uint32_t old = S.eval(s2, 0u);
          if (!(old & kFEval) && S.liq(s2)) {
            S.set(s2, kFLiqNew | nbit);
            S.push(ln, s2);
            S.push(2 + nx, s2);
          }
        }
      };
      const int ncur = s_n[2 + c];
      max_front = max(max_front, ncur);
      if (ncur <= S.cap) {
        const int* L = S.list(2 + c);
        for (int q = tid; q < ncur; q += kTrackBlock) expand(L[q]);
      } else {
        ++scans;
        const uint32_t cbit = kFront[c];
        for (int sl = tid; sl < w.nslots; sl += kTrackBlock) {
          if (!(S.F[sl] & cbit)) continue;
          S.clear(sl, cbit);
          expand(sl);
        }
      }
      __syncthreads();
      ++levels;
      if (s_n[2 + nx] == 0 || s_miss) break;
    }
    LAP(ph4);

    // Column depths under the liquid surface points (Melt::calc_depth,
    // Melt.cpp:271-317); a column is nk consecutive slots ending at the surface.
    if (SURF && nz > 1 && !s_miss) {
      auto column = [&](int sl) __attribute__((always_inline)) {
        int b, li, lj, lk;
        win_local(w, sl, b, li, lj, lk);
        const WinBox& x = w.b[b];
        const int col = x.coff + li * x.nj + lj;
        const int d0 = S.Dc[col];
        int depth = d0;
        bool deeper = false;
        while (true) {
          ++depth;
          if (depth == nz) break;
          if (lk - depth < 0) {
            note_miss(&s_miss, s_mp, x.i0 + li, x.j0 + lj, nz - 1 - depth);
            return;
          }
          S.eval(sl - depth, 0u);
          if (!S.liq(sl - depth)) break;
          deeper = true;
        }
        while (true) {
          --depth;
          if (depth == 0 || deeper) break;
          if (lk - depth < 0) {
            note_miss(&s_miss, s_mp, x.i0 + li, x.j0 + lj, nz - 1 - depth);
            return;
          }
          if (S.liq(sl - depth)) {
            S.eval(sl - depth, 0u);
            break;
          }
          S.eval(sl - depth, kFSolid | (depth != d0 ? kFTlast : 0u));
        }
        if (depth == nz) --depth;
        S.Dn[col] = (short)depth;
        if (a.depth_events) a.depth_ev[(size_t)r * w.ncols + col] = (short)depth;
      };
      const int nln = s_n[ln];
      max_liq = max(max_liq, nln);
      if (nln <= S.cap) {
        const int* L = S.list(ln);
        for (int q = tid; q < nln; q += kTrackBlock) column(L[q]);
      } else {
        for (int sl = tid; sl < w.nslots; sl += kTrackBlock)
          if (S.F[sl] & kFLiqNew) column(sl);
      }
    }
    __syncthreads();
    LAP(ph5);
    if (s_miss) break;

    // Commit the step: event bytes of the touched slots out (the rows start
    // zeroed), liquid-new becomes liquid-prev, every other flag clears.
    // Counts go to s_cnt[r & 1]; thread 0 clears the other set, which only it
    // has read. A step whose touched list overflowed commits every word.
    int c_ev = 0, c_tl = 0, c_so = 0, c_lq = 0;
    const int ntouch = s_n[kListTouched];
    if (ntouch <= S.cap) {
      const int* L = S.list(kListTouched);
      uint8_t* ev_row = a.ev + (size_t)r * w.stride;
      for (int q = tid; q < ntouch; q += kTrackBlock) {
        const int sl = L[q];
        const uint32_t f = S.F[sl];
        ev_row[sl] = (uint8_t)(f & (kEvEval | kEvTlast | kEvSolid));
        c_ev += 1;
        c_tl += (f & kFTlast) ? 1 : 0;
        c_so += (f & kFSolid) ? 1 : 0;
        c_lq += (f & kFLiqNew) ? 1 : 0;
        S.F[sl] = (f & kFLiqNew) ? (uint8_t)kFLiqPrev : (uint8_t)0;
      }
    } else {
      uint32_t* ev_row = reinterpret_cast<uint32_t*>(a.ev + (size_t)r * w.stride);
      uint32_t* F32 = reinterpret_cast<uint32_t*>(S.F);
      for (int q = tid; q < fbytes / 4; q += kTrackBlock) {
        const uint32_t f = F32[q];
        ev_row[q] = f & 0x64646464u;  // kEvEval | kEvTlast | kEvSolid in every byte
        c_ev += __popc(f & 0x04040404u);
        c_tl += __popc(f & 0x20202020u);
        c_so += __popc(f & 0x40404040u);
        c_lq += __popc(f & 0x02020202u);
        F32[q] = (f & 0x02020202u) >> 1;
      }
    }
    for (int c = tid; c < w.ncols; c += kTrackBlock) S.Dc[c] = S.Dn[c];
    c_ev = wave_sum_i(c_ev);
    c_tl = wave_sum_i(c_tl);
    c_so = wave_sum_i(c_so);
    c_lq = wave_sum_i(c_lq);
    int* cnt = s_cnt[r & 1];
    if (lane_id() == 0) {
      atomicAdd(&cnt[0], c_ev);
      atomicAdd(&cnt[1], c_tl);
      atomicAdd(&cnt[2], c_so);
      atomicAdd(&cnt[3], c_lq);
    }
    lp = ln;
    __syncthreads();
    const bool stop = a.finish[r - 1] && cnt[3] == 0;  // Util::sim_finish (Util.cpp:239-243)
    if (tid == 0) {
      int* sc = a.step_counts + 4 * (r - 1);
      sc[0] = cnt[0];
      sc[1] = cnt[1];
      sc[2] = cnt[2];
      sc[3] = cnt[3];
      s_done = r;
      s_stop = stop ? 1 : 0;
      int* other = s_cnt[(r + 1) & 1];
      other[0] = other[1] = other[2] = other[3] = 0;
    }
    LAP(ph6);
    if (stop) break;
  }
  __syncthreads();

  // Hand-off: liquid list, column depths, extents.
  if (tid == 0) {
    s_nl = 0;
    s_ext[0] = 0;     // max depth
    s_ext[1] = g.nz;  // min k of a liquid point
    s_ext[2] = g.nx;  // i_lo
    s_ext[3] = -1;    // i_hi
    s_ext[4] = g.ny;  // j_lo
    s_ext[5] = -1;    // j_hi
  }
  __syncthreads();
  auto handoff = [&](int sl) __attribute__((always_inline)) {
    int i, j, k, col;
    win_ijk(w, sl, i, j, k, col);
    const int q = atomicAdd(&s_nl, 1);
    if (q < a.liq_cap) a.liq_out[q] = (i * g.ny + j) * g.nz + k;
    atomicMin(&s_ext[1], k);
    atomicMin(&s_ext[2], i);
    atomicMax(&s_ext[3], i);
    atomicMin(&s_ext[4], j);
    atomicMax(&s_ext[5], j);
  };
  const int nlp = s_n[lp];
  if (nlp <= S.cap) {
    const int* L = S.list(lp);
    for (int q = tid; q < nlp; q += kTrackBlock) handoff(L[q]);
  } else {
    for (int sl = tid; sl < w.nslots; sl += kTrackBlock)
      if (S.F[sl] & kFLiqPrev) handoff(sl);
  }
  for (int c = tid; c < w.ncols; c += kTrackBlock) {
    int b = 0;
    while (b + 1 < w.nbox && c >= w.b[b + 1].coff) ++b;
    const WinBox& x = w.b[b];
    const int cc = c - x.coff;
    const int li = div_small(cc, x.nj, x.rnj);
    a.depths[(x.i0 + li) * g.ny + x.j0 + (cc - li * x.nj)] = S.Dc[c];
    if (S.Dc[c] > 0) atomicMax(&s_ext[0], (int)S.Dc[c]);
  }
  __syncthreads();
  if (tid == 0) {
    a.result[kTrDone] = s_done;
    a.result[kTrStatus] = s_miss ? 2 : (s_stop ? 1 : 0);
    a.result[kTrNliq] = s_nl;
    a.result[kTrMissI] = s_miss ? s_mp[0] : -1;
    a.result[kTrMissJ] = s_miss ? s_mp[1] : -1;
    a.result[kTrMissK] = s_miss ? s_mp[2] : -1;
    a.result[kTrMaxDepth] = s_ext[0];
    a.result[kTrMinK] = s_ext[1];
    a.result[kTrLiqBox + 0] = s_ext[2];
    a.result[kTrLiqBox + 1] = s_ext[3];
    a.result[kTrLiqBox + 2] = s_ext[4];
    a.result[kTrLiqBox + 3] = s_ext[5];
    a.result[kTrLevels] = levels;
    LAP(ph7);
    a.result[kTrTime + 0] = (int)ph0;
    a.result[kTrTime + 1] = (int)ph1;
    a.result[kTrTime + 2] = (int)ph2;
    a.result[kTrTime + 3] = 0;
    a.result[kTrTime + 4] = (int)ph4;
    a.result[kTrTime + 5] = (int)ph5;
    a.result[kTrTime + 6] = (int)ph6;
    a.result[kTrTime + 7] = (int)ph7;
    a.result[kTrScans] = scans;
    a.result[kTrMaxFront] = max_front;
    a.result[kTrMaxLiq] = max_liq;
  }
#undef LAP
#undef ROW_LOAD
}

__global__ __launch_bounds__(kTrackBlock) void k_track_window_lds(TrackArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  track_body<true, false>(a, lds);
}

__global__ __launch_bounds__(kTrackBlock) void k_track_window_global(TrackArgs a) { track_body<false, false>(a, nullptr); }

// ----------------------------------------------------------------- replay

__global__ __launch_bounds__(kBlock) void k_window_apply(WinGeom w, GridGeom g, int step0, const uint8_t* __restrict__ ev,
                                                         const double* __restrict__ Tc, double T_liq,
                                                         const int* __restrict__ result, PointState st, SolQueue q,
                                                         DevCounters dc, MeltTimes mt) {
  const int slot = blockIdx.x * kBlock + threadIdx.x;
  const int done = result[kTrDone];
  const size_t S = w.stride;
  uint8_t any = 0;
  int nsolid = 0;  // this slot's solidifications in the window
  if (slot < w.nslots)
    for (int r = 1; r <= done; ++r) {
      const uint8_t e = ev[r * S + slot];
      any |= e;
      nsolid += (e & kEvSolid) ? 1 : 0;
    }
  // the wavefront's queue entries in one reservation (a queue entry per
  // solidification, in step order per slot)
  int qn = nsolid;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(qn, o);
    if ((threadIdx.x & 63) >= o) qn += y;
  }
  const int qtot = __shfl(qn, 63);
  int qbase = 0;
  if ((threadIdx.x & 63) == 63 && qtot) qbase = atomicAdd(q.n, qtot);
  qbase = __shfl(qbase, 63) + qn - nsolid;
  if (slot >= w.nslots || !any) return;
  int i, j, k, col;
  win_ijk(w, slot, i, j, k, col);
  const int p = (i * g.ny + j) * g.nz + k;
  double T = st.T[p], Tl = st.T_last[p];
  int nm = st.numMelt ? st.numMelt[p] : 0;
  int sl_liq = -1, sl_sol = -1;
  bool liq_mark = false;
  // melt times (Run.cpp:607): at the reset of a step, a point of the previous
  // step that is liquid opens one unless one is open (Grid.h:400-404)
  int ntm = mt.step_time ? st.ntm[p] : 0, nsol = st.nsol ? st.nsol[p] : 0;
  auto reset = [&](int r) {
    if (mt.step_time && T >= mt.T_liq && ntm == nsol) {
      ++ntm;
      const int n = atomicAdd(st.tm_log.n, 1);
      if (n < st.tm_log.cap) {
        st.tm_log.p[n] = p;
        st.tm_log.a[n] = mt.step_time[r] - mt.dt;
      } else {
        atomicOr(dc.overflow, 1);
      }
    }
    Tl = T;
  };
  uint8_t prev = 0;
  for (int r = 1; r <= done; ++r) {
    const uint8_t e = ev[r * S + slot];
    const int step = step0 + r - 1;
    if (prev & kEvEval) reset(r - 1);  // the step's reset of the previous step's points
    if (e & kEvEval) {
      T = Tc[r * S + slot];
      if (st.hist_log.n) log_history(st, p, step, T);
      if (T >= T_liq) {  // the liquid bit the box sums derived from this T
        liq_mark = true;
        sl_liq = step;
      }
    }
    if (e & kEvTlast) Tl = Tc[(r - 1) * S + slot];
    if (e & kEvSolid) {
      ++nm;
      ++nsol;
      sl_liq = sl_sol = step;
      const int n = qbase++;
      if (n < q.cap) q.v[n] = SolEntry{p, step, T, Tl};
      else atomicOr(dc.overflow, 1);
    }
    prev = e;
  }
  if (prev & kEvEval) reset(done);  // reset at the start of the next step
  st.T[p] = T;
  st.T_last[p] = Tl;
  if (st.numMelt) st.numMelt[p] = (uint16_t)nm;
  if (mt.step_time) st.ntm[p] = ntm;
  if (st.nsol) st.nsol[p] = nsol;
  if (liq_mark) {
    if (st.G) st.G[p] = DBL_MIN;
    if (st.V) st.V[p] = DBL_MIN;
    if (st.eqFrac) st.eqFrac[p] = -1.0;
    if (st.oflag) st.oflag[p] = 1;
  }
  if (st.stamp_liq && sl_liq >= 0) st.stamp_liq[p] = sl_liq;
  if (st.stamp_sol && sl_sol >= 0) st.stamp_sol[p] = sl_sol;
}

__global__ __launch_bounds__(kBlock) void k_window_depth(WinGeom w, GridGeom g, const short* __restrict__ depth_ev,
                                                         const double* __restrict__ Tc, const int* __restrict__ result,
                                                         double T_liq, PointState st) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= w.ncols) return;
  int b = 0;
  while (b + 1 < w.nbox && c >= w.b[b + 1].coff) ++b;
  const WinBox& x = w.b[b];
  const int cc = c - x.coff;
  const int i = x.i0 + cc / x.nj, j = x.j0 + cc % x.nj;
  const int dnum = i * g.ny + j;
  const int done = result[kTrDone];
  for (int r = 1; r <= done; ++r) {
    const int dl = depth_ev[(size_t)r * w.ncols + c];
    if (dl < 0) continue;
    double depth;
    if (dl == g.nz - 1) {
      depth = dl;
    } else {
      const int sl = win_slot(w, i, j, g.nz - 1 - dl);
      const int ss = win_slot(w, i, j, g.nz - 1 - (dl + 1));
      if (sl < 0 || ss < 0) continue;  // both were evaluated this step, so inside the boxes
      const double Tl = Tc[(size_t)r * w.stride + sl], Ts = Tc[(size_t)r * w.stride + ss];
      depth = dl + (T_liq - Tl) / (Ts - Tl);
    }
    if (depth > st.depth_max[dnum]) {
      st.depth_max[dnum] = depth;
      if (st.depth)
        for (int d = dl; d >= 0; --d) st.depth[dnum * g.nz + (g.nz - 1 - d)] = depth;
    }
  }
}

// Staging of a window on the compute queue: each block reads its range from
// the block map and moves its part of it (a copy-engine transfer would make
// the box sums behind it on the stream wait for a cross-engine hand-off).
__global__ __launch_bounds__(kBlock) void k_window_upload(const UploadRange* __restrict__ r,
                                                          const unsigned short* __restrict__ map) {
  upload_block<kBlock>(r, map, blockIdx.x);
}

// The tracker's result block, the liquid list's first entries (at most
// liq_max, fewer when the result holds fewer) and the column counts, written
// to pinned host memory in one launch.
__global__ __launch_bounds__(kBlock) void k_window_export(const int* __restrict__ res, int nres,
                                                          const int* __restrict__ liq, int liq_max,
                                                          const int* __restrict__ ccnt, int nccnt,
                                                          int* __restrict__ res_h, int* __restrict__ liq_h,
                                                          int* __restrict__ ccnt_h) {
  const int nl = min(liq_max, max(res[kTrNliq], 0));
  const int step = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nres + nl + nccnt; i += step) {
    if (i < nres) res_h[i] = res[i];
    else if (i < nres + nl) liq_h[i - nres] = liq[i - nres];
    else ccnt_h[i - nres - nl] = ccnt[i - nres - nl];
  }
}

}  // namespace

void launch_window_eval_f64(const NodeF64* nodes, const int* node_off, int nrows, int row0, const WinGeom& w,
                            const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                            hipStream_t s) {
  if (nrows <= 0 || w.nslots <= 0) return;
  hipLaunchKernelGGL(k_window_eval_f64, dim3((w.nslots + kBlock - 1) / kBlock, nrows), dim3(kBlock), 0, s, nodes,
                     node_off, row0, w, g, T0, T_liq, Tc, lq);
  THESIS_HIP(hipGetLastError());
}

void launch_window_eval_f32c(const NodeF32c* nodes, const int* node_off, int nrows, int row0, const WinGeom& w,
                             const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                             hipStream_t s) {
  if (nrows <= 0 || w.nslots <= 0) return;
  constexpr int PPT = 2;
  hipLaunchKernelGGL(k_window_eval_f32c<PPT>, dim3((w.nslots + kBlock * PPT - 1) / (kBlock * PPT), nrows),
                     dim3(kBlock), 0, s, nodes, node_off, row0, w, g, T0, T_liq, Tc, lq);
  THESIS_HIP(hipGetLastError());
}

size_t track_lds_limit() {
  int dev = 0, v = 0;
  THESIS_HIP(hipGetDevice(&dev));
  THESIS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  return v > 8192 ? static_cast<size_t>(v) - 8192 : 0;  // room for the static counters and seed caches
}

size_t track_lds_bytes(int nslots, int ncols, int list_cap) {
  const size_t stride = (static_cast<size_t>(nslots) + 127) & ~static_cast<size_t>(127);
  const size_t cols = (static_cast<size_t>(ncols) + 7) & ~static_cast<size_t>(7);
  return stride + stride / 8 + 2 * cols * sizeof(short) + kTrackLists * static_cast<size_t>(list_cap) * sizeof(int);
}

void launch_track_window(const TrackArgs& a, hipStream_t s) {
  if (a.surface) {
    launch_track_surface(a, s);
    return;
  }
  if (a.lds_state) {
    const size_t bytes = track_lds_bytes(a.w.nslots, a.w.ncols, a.list_cap);
    THESIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_track_window_lds),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    hipLaunchKernelGGL(k_track_window_lds, dim3(1), dim3(kTrackBlock), bytes, s, a);
  } else {
    hipLaunchKernelGGL(k_track_window_global, dim3(1), dim3(kTrackBlock), 0, s, a);
  }
  THESIS_HIP(hipGetLastError());
}

void launch_window_apply(const WinGeom& w, const GridGeom& g, int nsteps, int step0, const uint8_t* ev,
                         const double* Tc, double T_liq, const int* result, const PointState& st,
                         const SolQueue& q, const DevCounters& dc, const MeltTimes& mt, hipStream_t s) {
  if (nsteps <= 0 || w.nslots <= 0) return;
  if (mt.step_time && (!st.ntm || !st.nsol || !st.tm_log.n))
    throw Error(kBadArgument, "melt times need the melt and solidification counts and their log");
  hipLaunchKernelGGL(k_window_apply, dim3((w.nslots + kBlock - 1) / kBlock), dim3(kBlock), 0, s, w, g, step0, ev, Tc,
                     T_liq, result, st, q, dc, mt);
  THESIS_HIP(hipGetLastError());
}

int surf_stride(const WinGeom& w) { return (w.nsurfw + 1) & ~1; }

bool two_pass_fits(const WinGeom& w) { return w.nbox > 0 && w.b[0].nk - 1 <= 16 * kLpt; }

void launch_surface_two_pass(const void* nodes, bool f32, const int* node_off, int nrows, int row0, const WinGeom& w,
                             const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                             unsigned long long* sl, int* cols, int* ccount, hipStream_t s) {
  if (nrows <= row0 || w.nslots <= 0) return;
  if (!two_pass_fits(w)) throw Error(kBadArgument, "two-pass box sums need at most 64 layers below the surface");
  const int ss = surf_stride(w);
  constexpr int kWords = kBlock / 64;
  const dim3 ga((ss + kWords - 1) / kWords, nrows - row0);
  if (f32) {
    constexpr int kR = kSurfRows;
    int tasks = 0;
    for (int b = 0; b < w.nbox; ++b) tasks += w.b[b].nw * ((w.b[b].ni + kR - 1) / kR);
    const dim3 ga2(std::max(1, (tasks + kWords - 1) / kWords), nrows - row0);
    hipLaunchKernelGGL(k_surf_eval_f32c<kR>, ga2, dim3(kBlock), 0, s, static_cast<const NodeF32c*>(nodes), node_off,
                       row0, w, g, T0, T_liq, Tc, sl, ss);
  } else {
    hipLaunchKernelGGL(k_surf_eval_f64, ga, dim3(kBlock), 0, s, static_cast<const NodeF64*>(nodes), node_off, row0, w,
                       g, T0, T_liq, Tc, sl, ss);
  }
  THESIS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_list_columns, dim3(nrows - row0), dim3(kListBlock), 0, s, w, row0, sl, ss, cols, ccount,
                     w.ncols);
  THESIS_HIP(hipGetLastError());
  const int nd = w.b[0].nk - 1;
  if (nd <= 0) return;
  if (!f32) {
    const dim3 gb((w.ncols * nd + kBlock - 1) / kBlock, nrows - row0);
    hipLaunchKernelGGL(k_deep_eval_f64, gb, dim3(kBlock), 0, s, static_cast<const NodeF64*>(nodes), node_off, row0, w,
                       g, T0, T_liq, cols, ccount, w.ncols, Tc, lq);
    THESIS_HIP(hipGetLastError());
    return;
  }
  int G = 1;
  while (G * kLpt < nd) G *= 2;
  const dim3 gb((w.ncols + kBlock / G - 1) / (kBlock / G), nrows - row0);
  const NodeF32c* n32 = static_cast<const NodeF32c*>(nodes);
  switch (G) {
    case 1: hipLaunchKernelGGL(k_deep_eval_sep_f32c<1>, gb, dim3(kBlock), 0, s, n32, node_off, row0, w, g, T0, T_liq, cols, ccount, w.ncols, Tc, lq); break;
    case 2: hipLaunchKernelGGL(k_deep_eval_sep_f32c<2>, gb, dim3(kBlock), 0, s, n32, node_off, row0, w, g, T0, T_liq, cols, ccount, w.ncols, Tc, lq); break;
    case 4: hipLaunchKernelGGL(k_deep_eval_sep_f32c<4>, gb, dim3(kBlock), 0, s, n32, node_off, row0, w, g, T0, T_liq, cols, ccount, w.ncols, Tc, lq); break;
    case 8: hipLaunchKernelGGL(k_deep_eval_f32c<8>, gb, dim3(kBlock), 0, s, n32, node_off, row0, w, g, T0, T_liq, cols, ccount, w.ncols, Tc, lq); break;
    default: hipLaunchKernelGGL(k_deep_eval_f32c<16>, gb, dim3(kBlock), 0, s, n32, node_off, row0, w, g, T0, T_liq, cols, ccount, w.ncols, Tc, lq); break;
  }
  THESIS_HIP(hipGetLastError());
}

void launch_surface_bits(const WinGeom& w, int nrows, int row0, const unsigned long long* lq, unsigned long long* sl,
                         hipStream_t s) {
  const int ss = surf_stride(w);
  if (nrows <= row0 || ss <= 0) return;
  constexpr int kWavesPerBlock = kBlock / 64;
  hipLaunchKernelGGL(k_surface_bits, dim3((ss + kWavesPerBlock - 1) / kWavesPerBlock, nrows - row0), dim3(kBlock), 0, s,
                     w, row0, lq, sl, ss);
  THESIS_HIP(hipGetLastError());
}

void launch_window_upload(const UploadRange* ranges, const unsigned short* map, int n, unsigned blocks,
                          hipStream_t s) {
  if (n <= 0 || blocks == 0) return;
  hipLaunchKernelGGL(k_window_upload, dim3(blocks), dim3(kBlock), 0, s, ranges, map);
  THESIS_HIP(hipGetLastError());
}

const unsigned short* UploadList::write_map(const UploadRange* r_dev) const {
  if (n > kUploadMaxRanges) throw Error(kBadArgument, "too many upload ranges");
  const size_t off = bytes_for(static_cast<size_t>(n), 0);
  unsigned short* map = reinterpret_cast<unsigned short*>(reinterpret_cast<uint8_t*>(r) + off);
  for (int q = 0; q < n; ++q) {
    const unsigned b1 = q + 1 < n ? r[q + 1].blk0 : blocks;
    for (unsigned b = r[q].blk0; b < b1; ++b) map[b] = static_cast<unsigned short>(q);
  }
  return reinterpret_cast<const unsigned short*>(reinterpret_cast<const uint8_t*>(r_dev) + off);
}

void UploadList::launch(const UploadRange* r_dev, hipStream_t s) const {
  launch_window_upload(r_dev, write_map(r_dev), n, blocks, s);
}

void launch_window_export(const int* result, int nres, const int* liq, int liq_max, const int* ccnt, int nccnt,
                          int* res_h, int* liq_h, int* ccnt_h, hipStream_t s) {
  const int most = nres + liq_max + nccnt;
  const int blocks = std::max(1, std::min((most + kBlock - 1) / kBlock, 64));
  hipLaunchKernelGGL(k_window_export, dim3(blocks), dim3(kBlock), 0, s, result, nres, liq, liq_max, ccnt, nccnt, res_h,
                     liq_h, ccnt_h);
  THESIS_HIP(hipGetLastError());
}

void launch_window_depth(const WinGeom& w, const GridGeom& g, int nsteps, const short* depth_ev, const double* Tc,
                         const int* result, double T_liq, const PointState& st, hipStream_t s) {
  if (nsteps <= 0 || w.ncols <= 0) return;
  hipLaunchKernelGGL(k_window_depth, dim3((w.ncols + kBlock - 1) / kBlock), dim3(kBlock), 0, s, w, g, depth_ev, Tc,
                     result, T_liq, st);
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
