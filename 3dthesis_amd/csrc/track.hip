// Solidification and melt-pool tracking kernels.
//
// The tracking lists of one timestep are small (10^2-10^4 points) and every
// point costs a full node sum, so evaluation here is one wavefront per point:
// the 64 lanes split the node list and reduce with an xor butterfly (which
// leaves the identical total in every lane, keeping control flow uniform).
//
// Per-point solidification (Grid::Solidify, Grid.cpp:316-338) needs node sets
// at the point's own liquidus-crossing time. Those are never stored: a
// wavefront walks the interval schedule of quadrature.hpp uniformly, parks one
// interval per lane, and every 64 intervals the lanes evaluate their own
// interval's Gauss nodes against the point. This file is compiled with
// -ffp-contract=off so the FP64 arithmetic rounds like the reference's.
#include <cfloat>

#include "track.hpp"
#include "window.hpp"

namespace thesis {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Solidification: a group of kSolGroup lanes per queue entry, four entries
// per wavefront. Neighbouring entries solidify at nearly the same time, so
// their interval schedules walk in step.
constexpr int kSolGroup = 16;
constexpr int kSolGroups = kWave / kSolGroup;
__device__ __forceinline__ int group_lane() { return threadIdx.x & (kSolGroup - 1); }
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = kSolGroup / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ int wave_index() { return blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); }

__device__ __forceinline__ int list_len(const DList& L) { return min(*L.n, L.cap); }

__device__ __forceinline__ void push(const DList& L, int v, const DevCounters& dc) {
  const int i = atomicAdd(L.n, 1);
  if (i < L.cap) L.v[i] = v;
  else atomicOr(dc.overflow, 1);
}

__device__ __forceinline__ double term_f64(const NodeF64& nd, double xp, double yp, double zp) {
  const double dx = xp - nd.xb;
  const double dy = yp - nd.yb;
  const double dz = zp - nd.zb;
  const double phi = exp(-3.0 * ((dx * dx * nd.ix) + (dy * dy * nd.iy) + (dz * dz * nd.iz)) + nd.expmod);
  return nd.dtau * phi;
}

// Node sum at one point, wavefront-cooperative; identical result in all lanes.
// The nodes may be in device memory or in LDS (same arithmetic and order).
__device__ __forceinline__ double wave_dT_at(const NodeF64* __restrict__ f64, const NodeF32c* __restrict__ f32,
                                             int n64, int n32, int prec, double xp, double yp, double zp) {
  const int lane = lane_id();
  if (prec == kFP64) {
    double acc = 0.0;
    for (int m = lane; m < n64; m += kWave) acc += term_f64(f64[m], xp, yp, zp);
    return wave_sum(acc);
  }
  float s = 0.f, c = 0.f;  // Kahan
  for (int m = lane; m < n32; m += kWave) {
    const NodeF32c nd = f32[m];
    const float dx = float(xp - nd.xb), dy = float(yp - nd.yb), dz = float(zp - nd.zb);
    float a = fmaf(nd.nx * dx, dx, nd.c);
    a = fmaf(nd.ny * dy, dy, a);
    a = fmaf(nd.nz * dz, dz, a);
    const float y = __builtin_amdgcn_exp2f(a) - c;
    const float tt = s + y;
    c = (tt - s) - y;
    s = tt;
  }
  return wave_sum(double(s) - double(c));
}

__device__ double wave_dT(const NodeSetDev& ns, double xp, double yp, double zp) {
  return wave_dT_at(ns.f64, ns.f32, ns.n64, ns.n32, ns.prec, xp, yp, zp);
}

// wave_dT_at for P points at once: each node is read once for all of them,
// and every point's sum has the same terms, order and reduction as
// wave_dT_at's (bit-identical results). Points past `np` are not summed.
template <int P>
__device__ __forceinline__ void wave_dT_multi(const NodeF64* __restrict__ f64, const NodeF32c* __restrict__ f32,
                                              int n64, int n32, int prec, const double* xp, const double* yp,
                                              const double* zp, int np, double* out) {
  const int lane = lane_id();
  if (prec == kFP64) {
    double acc[P];
#pragma unroll
    for (int u = 0; u < P; ++u) acc[u] = 0.0;
    for (int m = lane; m < n64; m += kWave) {
      const NodeF64 nd = f64[m];
#pragma unroll
      for (int u = 0; u < P; ++u)
        if (u < np) acc[u] += term_f64(nd, xp[u], yp[u], zp[u]);
    }
#pragma unroll
    for (int u = 0; u < P; ++u) out[u] = wave_sum(acc[u]);
    return;
  }
  float s[P], c[P];  // Kahan
#pragma unroll
  for (int u = 0; u < P; ++u) s[u] = c[u] = 0.f;
  for (int m = lane; m < n32; m += kWave) {
    const NodeF32c nd = f32[m];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      if (u >= np) continue;
      const float dx = float(xp[u] - nd.xb), dy = float(yp[u] - nd.yb), dz = float(zp[u] - nd.zb);
      float a = fmaf(nd.nx * dx, dx, nd.c);
      a = fmaf(nd.ny * dy, dy, a);
      a = fmaf(nd.nz * dz, dz, a);
      const float y = __builtin_amdgcn_exp2f(a) - c[u];
      const float tt = s[u] + y;
      c[u] = (tt - s[u]) - y;
      s[u] = tt;
    }
  }
#pragma unroll
  for (int u = 0; u < P; ++u) out[u] = wave_sum(double(s[u]) - double(c[u]));
}

__device__ __forceinline__ void point_xyz(const GridGeom& g, int p, double& x, double& y, double& z) {
  point_coords(g, p, x, y, z);
}

// Grid::Calc_T's "set" branch (Grid.cpp:300-311).
__device__ __forceinline__ void set_point(const PointState& st, int p, double T, double T_liq) {
  st.T[p] = T;
  if (st.hist_log.n) log_history(st, p, st.step, T);
  if (st.flag) st.flag[p] = 1u;
  if (T >= T_liq) {
    if (st.G) st.G[p] = DBL_MIN;
    if (st.V) st.V[p] = DBL_MIN;
    if (st.eqFrac) st.eqFrac[p] = -1.0;
    if (st.oflag) st.oflag[p] = 1;
    if (st.stamp_liq) st.stamp_liq[p] = st.step;
  }
}

// Grid::Solidify's immediate part (numMelt, Grid.cpp:317) and its queue entry;
// T[p] and T_last[p] are final for this step when this is called.
__device__ __forceinline__ void solidify_marks(const PointState& st, int p) {
  if (st.numMelt) st.numMelt[p] += 1;
  if (st.nsol) atomicAdd(st.nsol + p, 1);
  if (st.stamp_liq) st.stamp_liq[p] = st.step;
  if (st.stamp_sol) st.stamp_sol[p] = st.step;
}
__device__ __forceinline__ void put_solidify(const SolQueue& q, int i, const PointState& st, int p,
                                             const DevCounters& dc) {
  if (i < q.cap) q.v[i] = SolEntry{p, st.step, st.T[p], st.T_last[p]};
  else atomicOr(dc.overflow, 1);
}
__device__ __forceinline__ void queue_solidify(const SolQueue& q, const PointState& st, int p,
                                               const DevCounters& dc) {
  solidify_marks(st, p);
  put_solidify(q, atomicAdd(q.n, 1), st, p, dc);
}
__device__ __forceinline__ void put(const DList& L, int i, int v, const DevCounters& dc) {
  if (i < L.cap) L.v[i] = v;
  else atomicOr(dc.overflow, 1);
}

// The point-list kernels with many appends run 16 wavefronts per block and
// advance each list's counter once per block and round: appends reserve
// their places with LDS atomics, then one lane per list adds the block's
// count to the global counter. Same-address device atomics serialise across
// the chip, so one per appended point made these kernels atomic-bound.
constexpr int kListWaves = 16;
constexpr int kListBlock = kListWaves * kWave;
constexpr int kListPPW = 4;  // points per wavefront in the LDS-staged list kernels
// BFS levels: few claimed points per frontier point, and a round waits for
// its slowest wavefront, so smaller blocks spread a level over more CUs
constexpr int kBfsWaves = 4;
constexpr int kAggLists = 4;

// The block's reservations of one round -> global bases; clears the counts.
// Called by every thread of the block between two barriers.
__device__ __forceinline__ void block_bases(int* s_cnt, int* s_base, int* const* counters) {
  __syncthreads();
  if (threadIdx.x < kAggLists) {
    const int c = s_cnt[threadIdx.x];
    s_base[threadIdx.x] = c > 0 ? atomicAdd(counters[threadIdx.x], c) : 0;
    s_cnt[threadIdx.x] = 0;
  }
  __syncthreads();
}

// Evaluation counts summed over the block, one device atomic per block.
__device__ __forceinline__ void block_evals(unsigned long long* s_ev, unsigned long long mine,
                                            const DevCounters& dc) {
  if (mine) atomicAdd(s_ev, mine);
  __syncthreads();
  if (threadIdx.x == 0 && *s_ev) atomicAdd(dc.evals, *s_ev);
}

// --------------------------------------------------------------- list kernels

__global__ __launch_bounds__(kBlock) void k_reset(DList reset, PointState st, int reset_oflag) {
  const int n = list_len(reset);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
    const int p = reset.v[q];
    if (st.T_last) st.T_last[p] = st.T[p];
    st.flag[p] = 0u;
    if (reset_oflag) st.oflag[p] = 0;
  }
}

// The liquid marks of set_point on the listed points' G, V and eqFrac.
__global__ __launch_bounds__(kBlock) void k_remark_liquid(const int* pts, int n, PointState st) {
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
    const int p = pts[q];
    if (st.G) st.G[p] = DBL_MIN;
    if (st.V) st.V[p] = DBL_MIN;
    if (st.eqFrac) st.eqFrac[p] = -1.0;
  }
}

__global__ __launch_bounds__(kListBlock) void k_eval_list(NodeSetDev ns, GridGeom g, DList pts, double T_init,
                                                          double T_liq, PointState st, EvalAction act,
                                                          DevCounters dc) {
  __shared__ int s_cnt[kAggLists], s_base[kAggLists];
  int* const counters[kAggLists] = {act.liq.n, act.next.n, act.solid.n, act.solid2.n};
  const int len = list_len(pts);
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  if (threadIdx.x < kAggLists) s_cnt[threadIdx.x] = 0;
  // every listed point is evaluated once: one count for the launch
  if (blockIdx.x == 0 && threadIdx.x == 0 && len > 0)
    atomicAdd(dc.evals, static_cast<unsigned long long>(len) * ns.count());
  for (int w0 = blockIdx.x * kListWaves; w0 < len; w0 += gridDim.x * kListWaves) {  // block-uniform rounds
    const int w = w0 + wv;
    int p = -1, at[kAggLists] = {0, 0, 0, 0};
    bool liquid = false;
    __syncthreads();  // the counts were cleared
    if (w < len) {
      p = pts.v[w];
      double x, y, z;
      point_xyz(g, p, x, y, z);
      const double T = T_init + wave_dT(ns, x, y, z);
      liquid = T >= T_liq;
      if (lane == 0) {
        if (act.set) set_point(st, p, T, T_liq);
        if (liquid) {
          if (act.liq.v) at[0] = atomicAdd(&s_cnt[0], 1);
          if (act.next.v) at[1] = atomicAdd(&s_cnt[1], 1);
        } else {
          if (act.solid.v) {
            solidify_marks(st, p);
            at[2] = atomicAdd(&s_cnt[2], 1);
          }
          if (act.solid2.v) at[3] = atomicAdd(&s_cnt[3], 1);
          if (act.zero_depth) {
            int i, j, k;
            grid_ijk(g, p, i, j, k);
            st.depths[i * g.ny + j] = 0;
          }
        }
      }
    }
    block_bases(s_cnt, s_base, counters);
    if (w < len && lane == 0) {
      if (liquid) {
        if (act.liq.v) put(act.liq, s_base[0] + at[0], p, dc);
        if (act.next.v) put(act.next, s_base[1] + at[1], p, dc);
      } else {
        if (act.solid.v) put_solidify(act.solid, s_base[2] + at[2], st, p, dc);
        if (act.solid2.v) put(act.solid2, s_base[3] + at[3], p, dc);
      }
    }
  }
}

// Copy n8 8-byte units from device memory into LDS with all nt threads of the
// block, eight loads in flight per thread before their stores.
__device__ __forceinline__ void stage_units(unsigned long long* __restrict__ dst,
                                            const unsigned long long* __restrict__ src, int n8, int nt) {
  constexpr int kU = 8;
  for (int base = 0; base < n8; base += kU * nt) {
    unsigned long long v[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int u = base + k * nt + static_cast<int>(threadIdx.x);
      v[k] = u < n8 ? src[u] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int u = base + k * nt + static_cast<int>(threadIdx.x);
      if (u < n8) dst[u] = v[k];
    }
  }
}

// The node set of the precision in use into LDS with all nt threads of the
// block (8-byte units: both record sizes, 64 and 40 bytes, are whole units).
__device__ __forceinline__ void stage_node_set(uint8_t* s_nodes, const NodeSetDev& ns, int nt) {
  static_assert(sizeof(NodeF64) % 8 == 0 && sizeof(NodeF32c) % 8 == 0, "node records of whole 8-byte units");
  const bool f64 = ns.prec == kFP64;
  const int n8 = f64 ? ns.n64 * static_cast<int>(sizeof(NodeF64) / 8) : ns.n32 * static_cast<int>(sizeof(NodeF32c) / 8);
  const unsigned long long* src = f64 ? reinterpret_cast<const unsigned long long*>(ns.f64)
                                      : reinterpret_cast<const unsigned long long*>(ns.f32);
  stage_units(reinterpret_cast<unsigned long long*>(s_nodes), src, n8, nt);
}

// k_eval_list with the node set staged in LDS once per block: the
// wavefronts of a block then read it there instead of each reading the whole
// set from device memory (the same sums, in the same order).
__global__ __launch_bounds__(kListBlock) void k_eval_list_lds(NodeSetDev ns, GridGeom g, DList pts, double T_init,
                                                              double T_liq, PointState st, EvalAction act,
                                                              DevCounters dc) {
  extern __shared__ __align__(16) uint8_t s_nodes[];
  __shared__ int s_cnt[kAggLists], s_base[kAggLists];
  int* const counters[kAggLists] = {act.liq.n, act.next.n, act.solid.n, act.solid2.n};
  const int len = list_len(pts);
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  if (threadIdx.x < kAggLists) s_cnt[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && len > 0)
    atomicAdd(dc.evals, static_cast<unsigned long long>(len) * ns.count());
  stage_node_set(s_nodes, ns, kListBlock);
  const NodeF64* l64 = reinterpret_cast<const NodeF64*>(s_nodes);
  const NodeF32c* l32 = reinterpret_cast<const NodeF32c*>(s_nodes);
  constexpr int kPer = kListWaves * kListPPW;  // points per block and round
  for (int w0 = blockIdx.x * kPer; w0 < len; w0 += gridDim.x * kPer) {  // block-uniform rounds
    const int q0 = w0 + wv * kListPPW;  // this wavefront's first point
    const int np = max(0, min(kListPPW, len - q0));
    // lane u < np keeps point u's bookkeeping
    int p = -1, at[kAggLists] = {0, 0, 0, 0};
    bool liquid = false, mine = false;
    __syncthreads();  // the counts were cleared (and, the first time, the nodes are in)
    if (np > 0) {
      double x[kListPPW], y[kListPPW], z[kListPPW], dT[kListPPW];
#pragma unroll
      for (int u = 0; u < kListPPW; ++u) {
        const int pu = u < np ? pts.v[q0 + u] : pts.v[q0];
        point_xyz(g, pu, x[u], y[u], z[u]);
        if (lane == u) p = pu;
      }
      wave_dT_multi<kListPPW>(l64, l32, ns.n64, ns.n32, ns.prec, x, y, z, np, dT);
      double Tm = 0.0;
#pragma unroll
      for (int u = 0; u < kListPPW; ++u)
        if (lane == u) Tm = T_init + dT[u];
      mine = lane < np;
      liquid = mine && Tm >= T_liq;
      if (mine) {
        if (act.set) set_point(st, p, Tm, T_liq);
        if (liquid) {
          if (act.liq.v) at[0] = atomicAdd(&s_cnt[0], 1);
          if (act.next.v) at[1] = atomicAdd(&s_cnt[1], 1);
        } else {
          if (act.solid.v) {
            solidify_marks(st, p);
            at[2] = atomicAdd(&s_cnt[2], 1);
          }
          if (act.solid2.v) at[3] = atomicAdd(&s_cnt[3], 1);
          if (act.zero_depth) {
            int i, j, k;
            grid_ijk(g, p, i, j, k);
            st.depths[i * g.ny + j] = 0;
          }
        }
      }
    }
    block_bases(s_cnt, s_base, counters);
    if (mine) {
      if (liquid) {
        if (act.liq.v) put(act.liq, s_base[0] + at[0], p, dc);
        if (act.next.v) put(act.next, s_base[1] + at[1], p, dc);
      } else {
        if (act.solid.v) put_solidify(act.solid, s_base[2] + at[2], st, p, dc);
        if (act.solid2.v) put(act.solid2, s_base[3] + at[3], p, dc);
      }
    }
  }
}

// Run.cpp:908-957: probe down while liquid, else walk up until liquid; the
// column down to the depth is liquid. Evaluations do not set T.
__global__ __launch_bounds__(kBlock) void k_stork_depth(DList liq, NodeSetDev ns, GridGeom g, double T_init,
                                                        double T_liq, PointState st, uint8_t* is_liq,
                                                        DevCounters dc) {
  const int len = list_len(liq);
  const bool lead = lane_id() == 0;
  for (int w = wave_index(); w < len; w += gridDim.x * kWavesPerBlock) {
    const int p = liq.v[w];
    int i, j, k;
    grid_ijk(g, p, i, j, k);
    const int dnum = i * g.ny + j;
    const int col0 = (i * g.ny + j) * g.nz;
    int depth = st.depths[dnum];
    bool deeper = false;
    int evals = 0;
    while (true) {
      ++depth;
      if (depth == g.nz) break;  // the reference prints WARNING::MAXIMUM EXCEEDED
      double x, y, z;
      point_xyz(g, col0 + (g.nz - 1 - depth), x, y, z);
      const double T = T_init + wave_dT(ns, x, y, z);
      ++evals;
      if (!(T >= T_liq)) break;
      deeper = true;
    }
    while (true) {
      --depth;
      if (depth == 0 || deeper) break;
      double x, y, z;
      point_xyz(g, col0 + (g.nz - 1 - depth), x, y, z);
      const double T = T_init + wave_dT(ns, x, y, z);
      ++evals;
      if (T >= T_liq) break;
    }
    if (depth == g.nz - 1) --depth;
    if (lead) {
      st.depths[dnum] = depth;
      atomicAdd(dc.evals, (unsigned long long)evals * (unsigned long long)ns.count());
    }
    for (int d = lane_id(); d <= depth; d += kWave) is_liq[col0 + (g.nz - 1 - d)] = 1;
  }
}

// Run.cpp:959-991
__global__ __launch_bounds__(kBlock) void k_stork_cells(GridGeom g, const uint8_t* __restrict__ is_liq,
                                                        const uint8_t* __restrict__ c_prev, uint8_t* __restrict__ c_cur,
                                                        DList rel, DevCounters dc) {
  const int cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
  const long long n = (long long)cx * cy * cz;
  for (long long c = blockIdx.x * (long long)kBlock + threadIdx.x; c < n; c += (long long)gridDim.x * kBlock) {
    const int ci = (int)((c / cy) / cz), cj = (int)((c / cz) % cy), ck = (int)(c % cz);
    int cnt = 0;
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk) cnt += is_liq[((ci + di) * g.ny + (cj + dj)) * g.nz + (ck + dk)];
    c_cur[c] = (uint8_t)cnt;
    const int pv = c_prev[c];
    if (cnt != pv || pv % 8 || cnt % 8) push(rel, (int)c, dc);
  }
}

// Run.cpp:993-1061: every corner's T at t - dt (previous nodes) and t.
__global__ __launch_bounds__(kBlock) void k_stork_records(const int* __restrict__ rel, int n, GridGeom g,
                                                          NodeSetDev prev, NodeSetDev cur, double T_init,
                                                          double* __restrict__ out, DevCounters dc) {
  const int cy = g.ny - 1, cz = g.nz - 1;
  for (int w = wave_index(); w < 16 * n; w += gridDim.x * kWavesPerBlock) {
    const int r = w >> 4, m = w & 15;
    const int c = rel[r];
    const int ci = (c / cy) / cz, cj = (c / cz) % cy, ck = c % cz;
    const int v = m & 7;  // corner (di, dj, dk) = bits (2, 1, 0)
    const int p = ((ci + (v >> 2)) * g.ny + (cj + ((v >> 1) & 1))) * g.nz + (ck + (v & 1));
    double x, y, z;
    point_xyz(g, p, x, y, z);
    const NodeSetDev& ns = m < 8 ? prev : cur;
    const double T = T_init + wave_dT(ns, x, y, z);
    if (lane_id() == 0) {
      out[16 * r + m] = T;
      atomicAdd(dc.evals, (unsigned long long)ns.count());
    }
  }
}

// Melt::neighbor_check's claim step (Melt.cpp:195-257): 3x3 at the same layer
// (surface) or 3x3 x {k-1, k} (volume); first claimer of a point evaluates it.
__global__ __launch_bounds__(kBlock) void k_claim_neighbors(DList test, GridGeom g, int surface_only, PointState st,
                                                            DList cand, DList reset, DevCounters dc) {
  const int len = list_len(test);
  for (int q0 = blockIdx.x * kBlock + threadIdx.x; q0 < len; q0 += gridDim.x * kBlock) {
    const int p = test.v[q0];
    int i, j, k;
    grid_ijk(g, p, i, j, k);
    const int dk0 = surface_only ? 0 : -1;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = dk0; dk < 1; ++dk) {
          const int ni = i + di, nj = j + dj, nk = k + dk;
          if (ni < 0 || ni > g.nx - 1 || nj < 0 || nj > g.ny - 1 || nk < 0) continue;
          const int q = (ni * g.ny + nj) * g.nz + nk;
          if (atomicCAS(&st.flag[q], 0u, 1u) == 0u) {
            push(cand, q, dc);
            push(reset, q, dc);
          }
        }
  }
}

__global__ __launch_bounds__(kBlock) void k_claim_seeds(const int* __restrict__ seeds, int count, PointState st,
                                                        DList cand, DList reset, DevCounters dc) {
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < count; q += gridDim.x * kBlock) {
    const int p = seeds[q];
    if (atomicCAS(&st.flag[p], 0u, 1u) == 0u) {
      push(cand, p, dc);
      push(reset, p, dc);
    }
  }
}

// One BFS wave of Melt::neighbor_check fused with its evaluations: a wavefront
// per frontier point claims that point's neighbours, one lane per neighbour
// (atomicCAS on the claim flag, the reference's per-point lock,
// Melt.cpp:213-218), evaluates each point it wins with all 64 lanes, and
// appends the won and the liquid points with one counter atomic per list.
// Claims and appends are single round trips per wavefront; one after another
// they made the level latency-bound.
__global__ __launch_bounds__(kBfsWaves * kWave) void k_bfs_level(NodeSetDev ns, GridGeom g, DList test, int surface_only,
                                                          double T_init, double T_liq, PointState st, DList liq,
                                                          DList next, DList reset, int* zero, DevCounters dc) {
  __shared__ int s_cnt[kAggLists], s_base[kAggLists];
  __shared__ unsigned long long s_ev;
  int* const counters[kAggLists] = {reset.n, liq.n, next.n, nullptr};
  if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
  const int len = list_len(test);
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  const int nslots = surface_only ? 9 : 18;
  const unsigned long long below = (1ull << lane) - 1ull;
  if (threadIdx.x < kAggLists) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_ev = 0ull;
  unsigned long long evals = 0;
  for (int w0 = blockIdx.x * kBfsWaves; w0 < len; w0 += gridDim.x * kBfsWaves) {  // block-uniform rounds
    const int w = w0 + wv;
    int q = -1;
    bool won = false, hot = false;
    unsigned long long claimed = 0ull, melted = 0ull;
    int at_reset = 0, at_hot = 0;
    __syncthreads();  // the counts were cleared
    if (w < len) {
      const int p = test.v[w];
      int i, j, k;
      grid_ijk(g, p, i, j, k);
      // lane `slot` claims one neighbour: all claims of the wave in one round trip
      if (lane < nslots) {
        const int di = lane % 3 - 1, dj = (lane / 3) % 3 - 1, dk = surface_only ? 0 : lane / 9 - 1;
        const int ni = i + di, nj = j + dj, nk = k + dk;
        if (ni >= 0 && ni <= g.nx - 1 && nj >= 0 && nj <= g.ny - 1 && nk >= 0) {
          q = (ni * g.ny + nj) * g.nz + nk;
          // claims only go 0 -> 1 within a step: a point already claimed
          // (most neighbours of a pool's interior) costs a load, not an atomic
          won = st.flag[q] == 0u && atomicCAS(&st.flag[q], 0u, 1u) == 0u;
        }
      }
      claimed = __ballot(won);
      // the wave sums each claimed neighbour in turn; lane b keeps slot b's T
      double T = 0.0;
      for (unsigned long long rem = claimed; rem; rem &= rem - 1ull) {
        const int b = __ffsll(static_cast<long long>(rem)) - 1;
        double x, y, z;
        point_xyz(g, __shfl(q, b, kWave), x, y, z);
        const double Tb = T_init + wave_dT(ns, x, y, z);
        if (lane == b) T = Tb;
      }
      hot = won && T >= T_liq;
      melted = __ballot(hot);
      if (won) set_point(st, q, T, T_liq);
      const int n_claimed = __popcll(claimed), n_hot = __popcll(melted);
      if (lane == 0) {
        evals += static_cast<unsigned long long>(n_claimed) * ns.count();
        if (n_claimed) at_reset = atomicAdd(&s_cnt[0], n_claimed);
        if (n_hot) {
          at_hot = atomicAdd(&s_cnt[1], n_hot);
          atomicAdd(&s_cnt[2], n_hot);  // the same places in liq and next
        }
      }
      at_reset = __shfl(at_reset, 0, kWave);
      at_hot = __shfl(at_hot, 0, kWave);
    }
    block_bases(s_cnt, s_base, counters);
    if (won) put(reset, s_base[0] + at_reset + __popcll(claimed & below), q, dc);
    if (hot) {
      const int h = at_hot + __popcll(melted & below);
      put(liq, s_base[1] + h, q, dc);
      put(next, s_base[2] + h, q, dc);
    }
  }
  block_evals(&s_ev, evals, dc);
}

// The LDS-staged BFS kernels: a block of kBfsLdsBlock threads claims the
// neighbours of kBfsRound frontier points per round (kBfsSlotLanes threads
// per point), then evaluates the won points.
constexpr int kBfsRound = 32;     // frontier points per round and block
constexpr int kBfsSlotLanes = 32; // threads per frontier point (18 or 9 used)
constexpr int kBfsLdsBlock = kBfsRound * kBfsSlotLanes;  // 1024 threads
constexpr int kBfsLdsWaves = kBfsLdsBlock / kWave;

// Evaluate up to kBfsLdsBlock won points of a round, one set of kListPPW per
// wavefront in turn, and append them to reset (all), liq and next (the liquid
// ones); the body of k_bfs_level_lds's second phase. Block-uniform.
__device__ __forceinline__ void eval_won(const int* s_won, int nwon, const NodeF64* l64, const NodeF32c* l32,
                                         const NodeSetDev& ns, const GridGeom& g, double T_init, double T_liq,
                                         const PointState& st, const DList& reset, const DList& liq,
                                         const DList& next, int* s_cnt, int* s_base, int* const* counters,
                                         unsigned long long& evals, const DevCounters& dc) {
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  for (int e0 = 0; e0 < nwon; e0 += kBfsLdsWaves * kListPPW) {
    const int ea = e0 + wv * kListPPW;
    const int np = max(0, min(kListPPW, nwon - ea));
    int q = -1, at_reset = 0, at_hot = 0;
    bool hot = false, mine = false;
    if (np > 0) {
      double x[kListPPW], y[kListPPW], z[kListPPW], dT[kListPPW];
#pragma unroll
      for (int u = 0; u < kListPPW; ++u) {
        const int qu = s_won[ea + (u < np ? u : 0)];
        point_xyz(g, qu, x[u], y[u], z[u]);
        if (lane == u) q = qu;
      }
      wave_dT_multi<kListPPW>(l64, l32, ns.n64, ns.n32, ns.prec, x, y, z, np, dT);
      double T = 0.0;
#pragma unroll
      for (int u = 0; u < kListPPW; ++u)
        if (lane == u) T = T_init + dT[u];
      mine = lane < np;
      hot = mine && T >= T_liq;
      if (mine) {
        set_point(st, q, T, T_liq);
        evals += static_cast<unsigned long long>(ns.count());
        at_reset = atomicAdd(&s_cnt[0], 1);
        if (hot) {
          at_hot = atomicAdd(&s_cnt[1], 1);
          atomicAdd(&s_cnt[2], 1);  // the same places in liq and next
        }
      }
    }
    block_bases(s_cnt, s_base, counters);
    if (mine) {
      put(reset, s_base[0] + at_reset, q, dc);
      if (hot) {
        put(liq, s_base[1] + at_hot, q, dc);
        put(next, s_base[2] + at_hot, q, dc);
      }
    }
  }
}

// k_bfs_level with the node set staged in LDS once per block, in two phases
// per round: the block's threads claim the neighbours of kBfsRound frontier
// points (a thread per point and neighbour slot), collecting the won points
// in LDS; then its wavefronts evaluate the won points, one each in turn, so
// that a round's evaluations spread over all of them. The claims, sums and
// appends are those of k_bfs_level; only the order of list entries differs.
__global__ __launch_bounds__(kBfsLdsBlock) void k_bfs_level_lds(NodeSetDev ns, GridGeom g, DList test,
                                                                int surface_only, double T_init, double T_liq,
                                                                PointState st, DList liq, DList next, DList reset,
                                                                int* zero, const int* __restrict__ seeds, int nseeds,
                                                                DevCounters dc) {
  extern __shared__ __align__(16) uint8_t s_nodes[];
  __shared__ int s_won[kBfsLdsBlock];
  __shared__ int s_nwon;
  __shared__ int s_cnt[kAggLists], s_base[kAggLists];
  __shared__ unsigned long long s_ev;
  int* const counters[kAggLists] = {reset.n, liq.n, next.n, nullptr};
  if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
  const int len = list_len(test);
  // claim items: kBfsSlotLanes per frontier point, then one per seed
  const int nfront = len * kBfsSlotLanes, total = nfront + nseeds;
  if (total == 0) return;  // an empty frontier (levels queued past the end of a step's waves)
  // the grid is sized for the largest frontier the host expects: a block
  // without a round stages nothing
  if (blockIdx.x * kBfsLdsBlock >= total) return;
  const int nslots = surface_only ? 9 : 18;
  if (threadIdx.x < kAggLists) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_ev = 0ull;
  stage_node_set(s_nodes, ns, kBfsLdsBlock);
  const NodeF64* l64 = reinterpret_cast<const NodeF64*>(s_nodes);
  const NodeF32c* l32 = reinterpret_cast<const NodeF32c*>(s_nodes);
  unsigned long long evals = 0;
  for (int r0 = blockIdx.x * kBfsLdsBlock; r0 < total; r0 += gridDim.x * kBfsLdsBlock) {  // block-uniform rounds
    __syncthreads();  // the last round's won points are evaluated (and the nodes are in)
    if (threadIdx.x == 0) s_nwon = 0;
    __syncthreads();
    {
      const int item = r0 + static_cast<int>(threadIdx.x);
      int q = -1;
      if (item < nfront) {
        const int f = item / kBfsSlotLanes, slot = item % kBfsSlotLanes;
        if (slot < nslots) {
          const int p = test.v[f];
          int i, j, k;
          grid_ijk(g, p, i, j, k);
          const int di = slot % 3 - 1, dj = (slot / 3) % 3 - 1, dk = surface_only ? 0 : slot / 9 - 1;
          const int ni = i + di, nj = j + dj, nk = k + dk;
          if (ni >= 0 && ni <= g.nx - 1 && nj >= 0 && nj <= g.ny - 1 && nk >= 0) q = (ni * g.ny + nj) * g.nz + nk;
        }
      } else if (item < total) {
        q = seeds[item - nfront];  // a beam-trace seed (Melt.cpp:98-183) is claimed like a neighbour
      }
      // claims only go 0 -> 1 within a step: a point already claimed costs a load, not an atomic
      if (q >= 0 && st.flag[q] == 0u && atomicCAS(&st.flag[q], 0u, 1u) == 0u) s_won[atomicAdd(&s_nwon, 1)] = q;
    }
    __syncthreads();
    eval_won(s_won, s_nwon, l64, l32, ns, g, T_init, T_liq, st, reset, liq, next, s_cnt, s_base, counters, evals, dc);
  }
  block_evals(&s_ev, evals, dc);
}

// The recheck of a Volume step and its first BFS level in one launch
// (Run.cpp:511-531, the beam-trace seeds of :533-545, the first
// Melt::neighbor_check wave of :547-551). A round of a
// block rechecks kRecheckRound of the previous liquid points (set; liquid ones
// join liq, the others are queued for solidification), claims the neighbours
// of the ones still liquid, and evaluates the won points as k_bfs_level_lds
// does; rounds past the rechecks claim beam-trace seeds. The step prologue
// has already claimed every previous liquid point (launch_step_prologue with
// preclaim), so no neighbour claim takes one of them before its recheck: the
// claims, sums and appends are those of k_eval_list_lds followed by
// k_bfs_level_lds; only the order of list entries differs.
constexpr int kRecheckRound = 32;
constexpr int kRecheckPPW = kRecheckRound / kBfsLdsWaves;  // 2 points per wavefront
static_assert(kRecheckRound * kBfsSlotLanes <= kBfsLdsBlock, "a round's claims fit the won-point list");
__global__ __launch_bounds__(kBfsLdsBlock) void k_recheck_bfs_lds(NodeSetDev ns, GridGeom g, DList last, double T_init,
                                                                  double T_liq, PointState st, DList liq, DList next,
                                                                  DList reset, SolQueue solid, int* zero,
                                                                  const int* __restrict__ seeds, int nseeds,
                                                                  DevCounters dc) {
  extern __shared__ __align__(16) uint8_t s_nodes[];
  __shared__ int s_won[kBfsLdsBlock];
  __shared__ int s_front[kRecheckRound];
  __shared__ int s_nwon, s_nfront;
  __shared__ int s_cnt[kAggLists], s_base[kAggLists];
  __shared__ unsigned long long s_ev;
  int* const counters[kAggLists] = {reset.n, liq.n, next.n, solid.n};
  if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
  const int len = list_len(last);
  const int r_check = (len + kRecheckRound - 1) / kRecheckRound;
  const int rounds = r_check + (nseeds + kBfsLdsBlock - 1) / kBfsLdsBlock;
  if (static_cast<int>(blockIdx.x) >= rounds) return;  // no round for this block (none at all: nothing to do)
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  if (threadIdx.x < kAggLists) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_ev = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0 && len > 0)
    atomicAdd(dc.evals, static_cast<unsigned long long>(len) * ns.count());
  stage_node_set(s_nodes, ns, kBfsLdsBlock);
  const NodeF64* l64 = reinterpret_cast<const NodeF64*>(s_nodes);
  const NodeF32c* l32 = reinterpret_cast<const NodeF32c*>(s_nodes);
  unsigned long long evals = 0;
  for (int r = blockIdx.x; r < rounds; r += gridDim.x) {  // block-uniform rounds
    __syncthreads();  // the last round's won points are evaluated (and the nodes are in)
    if (threadIdx.x == 0) {
      s_nwon = 0;
      s_nfront = 0;
    }
    __syncthreads();
    int q = -1;
    if (r < r_check) {
      // recheck: kRecheckPPW points per wavefront
      const int ea = r * kRecheckRound + wv * kRecheckPPW;
      const int np = max(0, min(kRecheckPPW, len - ea));
      int p = -1, at = 0;
      bool liquid = false, mine = false;
      if (np > 0) {
        double x[kRecheckPPW], y[kRecheckPPW], z[kRecheckPPW], dT[kRecheckPPW];
#pragma unroll
        for (int u = 0; u < kRecheckPPW; ++u) {
          const int pu = last.v[ea + (u < np ? u : 0)];
          point_xyz(g, pu, x[u], y[u], z[u]);
          if (lane == u) p = pu;
        }
        wave_dT_multi<kRecheckPPW>(l64, l32, ns.n64, ns.n32, ns.prec, x, y, z, np, dT);
        double Tm = 0.0;
#pragma unroll
        for (int u = 0; u < kRecheckPPW; ++u)
          if (lane == u) Tm = T_init + dT[u];
        mine = lane < np;
        liquid = mine && Tm >= T_liq;
        if (mine) {
          set_point(st, p, Tm, T_liq);
          if (liquid) {
            at = atomicAdd(&s_cnt[1], 1);
            s_front[atomicAdd(&s_nfront, 1)] = p;
          } else {
            solidify_marks(st, p);
            at = atomicAdd(&s_cnt[3], 1);
          }
        }
      }
      block_bases(s_cnt, s_base, counters);  // also: every still-liquid point is in s_front
      if (mine) {
        if (liquid) put(liq, s_base[1] + at, p, dc);
        else put_solidify(solid, s_base[3] + at, st, p, dc);
      }
      // a thread per still-liquid point and neighbour slot
      const int f = threadIdx.x / kBfsSlotLanes, slot = threadIdx.x % kBfsSlotLanes;
      if (f < s_nfront && slot < 18) {
        int i, j, k;
        grid_ijk(g, s_front[f], i, j, k);
        const int ni = i + slot % 3 - 1, nj = j + (slot / 3) % 3 - 1, nk = k + slot / 9 - 1;
        if (ni >= 0 && ni <= g.nx - 1 && nj >= 0 && nj <= g.ny - 1 && nk >= 0) q = (ni * g.ny + nj) * g.nz + nk;
      }
    } else {
      const int item = (r - r_check) * kBfsLdsBlock + static_cast<int>(threadIdx.x);
      if (item < nseeds) q = seeds[item];  // a beam-trace seed (Melt.cpp:98-183) is claimed like a neighbour
    }
    if (q >= 0 && st.flag[q] == 0u && atomicCAS(&st.flag[q], 0u, 1u) == 0u) s_won[atomicAdd(&s_nwon, 1)] = q;
    __syncthreads();
    eval_won(s_won, s_nwon, l64, l32, ns, g, T_init, T_liq, st, reset, liq, next, s_cnt, s_base, counters, evals, dc);
  }
  block_evals(&s_ev, evals, dc);
}

// Beam-trace seeds (Melt.cpp:98-183) claimed and evaluated in one pass.
__global__ __launch_bounds__(kBlock) void k_seed_eval(const int* __restrict__ seeds, int count, NodeSetDev ns,
                                                      GridGeom g, double T_init, double T_liq, PointState st,
                                                      DList liq, DList next, DList reset, DevCounters dc) {
  const int lane = lane_id();
  for (int w = wave_index(); w < count; w += gridDim.x * kWavesPerBlock) {
    const int p = seeds[w];
    int won = 0;
    if (lane == 0) won = atomicCAS(&st.flag[p], 0u, 1u) == 0u;
    won = __shfl(won, 0, kWave);
    if (!won) continue;
    double x, y, z;
    point_xyz(g, p, x, y, z);
    const double T = T_init + wave_dT(ns, x, y, z);
    if (lane == 0) {
      atomicAdd(dc.evals, (unsigned long long)ns.count());
      set_point(st, p, T, T_liq);
      push(reset, p, dc);
      if (T >= T_liq) {
        push(liq, p, dc);
        push(next, p, dc);
      }
    }
  }
}

// Start of a tracked Solidification step (Run.cpp:497-509 / :603-616): the
// previous step's touched points get T_last = T and lose their claim flag;
// the previous liquid list becomes the new reset list (and, for Surface
// tracking, the first BFS frontier); the step's other lists start empty.
// Blocks bid of nb of the prologue.
__device__ __forceinline__ void step_prologue(const DList& old_reset, const DList& last, const DList& new_reset,
                                              const DList& frontier, int* zero0, int* zero1, int* zero2, int* zero3,
                                              const PointState& st, const DevCounters& dc, double tm_time, double T_liq,
                                              int preclaim, int bid, int nb) {
  const int n_old = list_len(old_reset);
  const int n_last = list_len(last);
  const int stride = nb * kBlock;
  const int tid = bid * kBlock + threadIdx.x;
  for (int q = tid; q < n_old; q += stride) {
    const int p = old_reset.v[q];
    // Run.cpp:607: a point liquid at the previous step opens a melt time
    // unless one is open (Grid.h:400-404); a point listed twice opens one
    if (st.ntm && st.T[p] >= T_liq) {
      const int open = st.nsol[p];
      if (atomicCAS(st.ntm + p, open, open + 1) == open) {
        const int i = atomicAdd(st.tm_log.n, 1);
        if (i < st.tm_log.cap) {
          st.tm_log.p[i] = p;
          st.tm_log.a[i] = tm_time;
        } else {
          atomicOr(dc.overflow, 1);
        }
      }
    }
    st.T_last[p] = st.T[p];
    // preclaim: the points liquid at the previous step (the last list: every
    // point of old_reset was evaluated once then, and the liquid ones are the
    // ones listed) stay claimed for their recheck
    st.flag[p] = preclaim && st.T[p] >= T_liq ? 1u : 0u;
  }
  for (int q = tid; q < n_last; q += stride) {
    if (q < new_reset.cap) new_reset.v[q] = last.v[q];
    if (frontier.v && q < frontier.cap) frontier.v[q] = last.v[q];
  }
  if (tid == 0) {
    *new_reset.n = n_last;
    if (frontier.v) *frontier.n = n_last;
    if (n_last > new_reset.cap || (frontier.v && n_last > frontier.cap)) atomicOr(dc.overflow, 1);
    if (zero0) *zero0 = 0;
    if (zero1) *zero1 = 0;
    if (zero2) *zero2 = 0;
    if (zero3) *zero3 = 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_step_prologue(DList old_reset, DList last, DList new_reset,
                                                          DList frontier, int* zero0, int* zero1, int* zero2,
                                                          int* zero3, PointState st, DevCounters dc, double tm_time,
                                                          double T_liq, int preclaim) {
  step_prologue(old_reset, last, new_reset, frontier, zero0, zero1, zero2, zero3, st, dc, tm_time, T_liq, preclaim,
                blockIdx.x, gridDim.x);
}

// The step's upload (blocks 0 .. up_blocks - 1) and its prologue (the rest)
// in one launch: they touch disjoint memory, and the kernels after them need
// both.
__global__ __launch_bounds__(kBlock) void k_step_start(const UploadRange* __restrict__ ranges,
                                                       const unsigned short* __restrict__ map, unsigned up_blocks,
                                                       DList old_reset, DList last, DList new_reset, DList frontier,
                                                       int* zero0, int* zero1, int* zero2, int* zero3, PointState st,
                                                       DevCounters dc, double tm_time, double T_liq, int preclaim) {
  if (blockIdx.x < up_blocks) {
    upload_block<kBlock>(ranges, map, blockIdx.x);
    return;
  }
  step_prologue(old_reset, last, new_reset, frontier, zero0, zero1, zero2, zero3, st, dc, tm_time, T_liq, preclaim,
                static_cast<int>(blockIdx.x - up_blocks), static_cast<int>(gridDim.x - up_blocks));
}

// Solidify the column under a surface point that just solidified
// (Run.cpp:628-643).
__device__ void column_one(int p, const NodeSetDev& cur, const NodeSetDev& last, const GridGeom& g, double T_init,
                           double T_liq, const PointState& st, const SolQueue& solid, const DList& reset,
                           const DevCounters& dc) {
  int i, j, k;
  grid_ijk(g, p, i, j, k);
  const int dnum = i * g.ny + j;
  const int depth = st.depths[dnum];
  const bool lead = lane_id() == 0;
  for (int d = depth; d > 0; --d) {
    const int q = (i * g.ny + j) * g.nz + (g.nz - 1 - d);
    double x, y, z;
    point_xyz(g, q, x, y, z);
    const double T = T_init + wave_dT(cur, x, y, z);
    if (lead) {
      set_point(st, q, T, T_liq);
      atomicAdd(dc.evals, (unsigned long long)cur.count());
    }
    if (d != depth) {
      const double Tl = T_init + wave_dT(last, x, y, z);
      if (lead) {
        st.T_last[q] = Tl;
        atomicAdd(dc.evals, (unsigned long long)last.count());
      }
    }
    if (lead) {
      queue_solidify(solid, st, q, dc);
      push(reset, q, dc);
    }
  }
  if (lead) st.depths[dnum] = 0;
}

__global__ __launch_bounds__(kBlock) void k_column_expand(DList cols, NodeSetDev cur, NodeSetDev last, GridGeom g,
                                                          double T_init, double T_liq, PointState st, SolQueue solid,
                                                          DList reset, DevCounters dc) {
  const int len = list_len(cols);
  for (int w = wave_index(); w < len; w += gridDim.x * kWavesPerBlock)
    column_one(cols.v[w], cur, last, g, T_init, T_liq, st, solid, reset, dc);
}

// Melt::calc_depth (Melt.cpp:271-317), one wavefront per liquid surface point.
__device__ void depth_one(int p, const NodeSetDev& cur, const NodeSetDev& last, const GridGeom& g, double T_init,
                          double T_liq, const PointState& st, const SolQueue& solid, const DList& reset,
                          const DevCounters& dc) {
  int i, j, k;
  grid_ijk(g, p, i, j, k);
  const int dnum = i * g.ny + j;
  const int d0 = st.depths[dnum];
  const bool lead = lane_id() == 0;
  int depth = d0;
  bool deeper = false;
  while (true) {
    ++depth;
    if (depth == g.nz) break;
    const int q = (i * g.ny + j) * g.nz + (g.nz - 1 - depth);
    double x, y, z;
    point_xyz(g, q, x, y, z);
    const double T = T_init + wave_dT(cur, x, y, z);
    if (lead) {
      push(reset, q, dc);
      set_point(st, q, T, T_liq);
      atomicAdd(dc.evals, (unsigned long long)cur.count());
    }
    if (!(T >= T_liq)) break;
    deeper = true;
  }
  while (true) {
    --depth;
    if (depth == 0 || deeper) break;
    const int q = (i * g.ny + j) * g.nz + (g.nz - 1 - depth);
    double x, y, z;
    point_xyz(g, q, x, y, z);
    const double T = T_init + wave_dT(cur, x, y, z);
    if (lead) {
      push(reset, q, dc);
      set_point(st, q, T, T_liq);
      atomicAdd(dc.evals, (unsigned long long)cur.count());
    }
    if (!(T >= T_liq)) {
      if (depth != d0) {
        const double Tl = T_init + wave_dT(last, x, y, z);
        if (lead) {
          st.T_last[q] = Tl;
          atomicAdd(dc.evals, (unsigned long long)last.count());
        }
      }
      if (lead) queue_solidify(solid, st, q, dc);
    } else {
      break;
    }
  }
  if (depth == g.nz) --depth;
  if (lead) st.depths[dnum] = depth;
}

__global__ __launch_bounds__(kBlock) void k_calc_depth(DList liq, NodeSetDev cur, NodeSetDev last, GridGeom g,
                                                       double T_init, double T_liq, PointState st, SolQueue solid,
                                                       DList reset, DevCounters dc) {
  const int len = list_len(liq);
  for (int w = wave_index(); w < len; w += gridDim.x * kWavesPerBlock)
    depth_one(liq.v[w], cur, last, g, T_init, T_liq, st, solid, reset, dc);
}

// Melt::calc_depth_max (Melt.cpp:319-357).
__global__ __launch_bounds__(kBlock) void k_depth_max(DList liq, GridGeom g, double T_liq, PointState st) {
  const int len = list_len(liq);
  for (int q0 = blockIdx.x * kBlock + threadIdx.x; q0 < len; q0 += gridDim.x * kBlock) {
    const int p = liq.v[q0];
    int i, j, k;
    grid_ijk(g, p, i, j, k);
    const int dnum = i * g.ny + j;
    const int dl = st.depths[dnum];
    double depth;
    if (dl == g.nz - 1) {
      depth = dl;
    } else {
      const int pl = (i * g.ny + j) * g.nz + (g.nz - 1 - dl);
      const int ps = (i * g.ny + j) * g.nz + (g.nz - 1 - (dl + 1));
      depth = dl + (T_liq - st.T[pl]) / (st.T[ps] - st.T[pl]);
    }
    if (depth > st.depth_max[dnum]) {
      st.depth_max[dnum] = depth;
      if (st.depth)
        for (int d = dl; d >= 0; --d) st.depth[(i * g.ny + j) * g.nz + (g.nz - 1 - d)] = depth;
    }
  }
}

// ---------------------------------------------------- per-point solidification

struct Accum {
  double dT = 0, gx = 0, gy = 0, gz = 0, lap = 0, s = 0, nodes = 0;
};

// What a wavefront's walk over the quadrature does with the nodes: sum T,
// sum T with gradient/Laplacian/source terms, or only count them (the
// reference's node-list size, before wall images).
enum BuildKind { kBuildT = 0, kBuildGrad = 1, kBuildCount = 2 };

template <int KIND>
struct WaveBuild {
  const SolParams& sp;
  const PathSeg* P;
  const BeamQ& B;
  double tq, xp, yp, zp;
  int lane;      // lane within the entry's group
  int used = 0;  // lanes of the group holding a node of the current batch (uniform in the group)
  // this lane's node of the batch: node a of interval [b_t1, b_t2] on segment
  // b_seg, or of a merged interval (b_seg < 0) with centre (b_cx, b_cy) and
  // qmod b_q * beta^(b_order - a)
  double b_t1 = 0, b_t2 = 0, b_cx = 0, b_cy = 0, b_q = 0;
  int b_seg = 0, b_order = 0, b_a = -1;
  Accum acc;

  __device__ WaveBuild(const SolParams& s, const PathSeg* p, const BeamQ& b, double t, double x, double y, double z)
      : sp(s), P(p), B(b), tq(t), xp(x), yp(y), zp(z), lane(group_lane()) {}

  // Grid::Calc_Solidficiaton_Primary's accumulation (Grid.cpp:422-448) for a
  // node centred at (xb, yb, zb).
  __device__ void add_at(const RawNode& r, double xb, double yb, double zb) {
    const double ix = 1.0 / r.px, iy = 1.0 / r.py, iz = 1.0 / r.pz;
    const double dx = xp - xb, dy = yp - yb, dz = zp - zb;
    const double phi = exp(-3.0 * ((dx * dx * ix) + (dy * dy * iy) + (dz * dz * iz)) + node_expmod(r));
    const double dT = r.dtau * phi;
    acc.dT += dT;
    acc.nodes += 1.0;
    if (KIND == kBuildGrad) {
      const double ddpx = -6.0 * ix, ddpy = -6.0 * iy, ddpz = -6.0 * iz;
      const double dpx = ddpx * dx, dpy = ddpy * dy, dpz = ddpz * dz;
      acc.gx += dT * dpx;
      acc.gy += dT * dpy;
      acc.gz += dT * dpz;
      acc.lap += dT * (dpx * dpx + dpy * dpy + dpz * dpz + ddpx + ddpy + ddpz);
      acc.s += (r.dtau == 0) ? phi : 0;
    }
  }
  __device__ void add(const RawNode& r) {
    if (KIND == kBuildCount) {
      acc.nodes += 1.0;
      return;
    }
    if (sp.nimages == 0) {
      add_at(r, r.xb, r.yb, r.zb);
      return;
    }
    for (int m = 0; m < sp.nimages; ++m) {
      const WallImage& w = sp.images[m];
      add_at(r, w.ox + w.sx * r.xb, w.oy + w.sy * r.yb, w.oz + w.sz * r.zb);
    }
  }
  __device__ void instant(int seg) {
    // weight 0: only the source term S and the node count see it (the
    // reference's list holds it, and Calc_T evaluates it, once per image)
    if (lane != 0) return;
    RawNode n;
    if (!instant_node(P, seg, B, sp.rb, tq, n)) return;
    if (KIND == kBuildT) acc.nodes += sp.nimages > 0 ? sp.nimages : 1;
    else add(n);
  }
  // Hand the interval's nodes a = order-1 .. 0 to consecutive lanes, building
  // a batch whenever all lanes of the group hold one.
  __device__ void place(double t1, double t2, int seg, int order, double cx, double cy, double q) {
    for (int a0 = 0; a0 < order;) {
      const int take = min(order - a0, kSolGroup - used);
      if (lane >= used && lane < used + take) {
        b_t1 = t1;
        b_t2 = t2;
        b_seg = seg;
        b_order = order;
        b_a = order - 1 - (a0 + lane - used);
        b_cx = cx;
        b_cy = cy;
        b_q = q;
      }
      used += take;
      a0 += take;
      if (used == kSolGroup) flush();
    }
  }
  __device__ void interval(double t1, double t2, int seg, int order) { place(t1, t2, seg, order, 0.0, 0.0, 0.0); }
  __device__ void merged(double t1, double t2, int order, double cx, double cy, double q) {
    place(t1, t2, -1, order, cx, cy, q);
  }
  __device__ void flush() {
    if (b_a >= 0) {
      RawNode n;
      bool keep;
      if (b_seg >= 0) {
        keep = interval_node(P, b_seg, B, sp.rb, sp.alpha, tq, b_t1, b_t2, b_order, b_a, sp.gl, n);
      } else {
        double q = b_q;
        for (int k = b_a; k < b_order; ++k) q *= B.beta;  // compounds per node, as Calc.cpp:463 does
        keep = merged_node(B, sp.alpha, tq, b_t1, b_t2, b_order, b_a, b_cx, b_cy, q, sp.gl, n);
      }
      if (keep) add(n);
    }
    used = 0;
    b_a = -1;
  }
};

template <int KIND>
__device__ Accum wave_build_sum(const SolParams& sp, const int* hints, double tq, bool is_sol, bool compressed,
                                double x, double y, double z) {
  Accum tot;
  for (int b = 0; b < sp.nbeams; ++b) {
    const PathSeg* P = sp.segs + sp.seg_off[b];
    const int nseg = sp.seg_off[b + 1] - sp.seg_off[b];
    const int seg = locate_segment(P, nseg, tq, hints[b]);
    WaveBuild<KIND> wb(sp, P, sp.beams[b], tq, x, y, z);
    if (compressed) walk_compressed(P, nseg, sp.beams[b], sp.rb, sp.alpha, tq, is_sol, seg, wb);
    else walk_intervals(P, nseg, sp.beams[b], tq, is_sol, seg, wb);
    wb.flush();
    tot.dT += wb.acc.dT;
    tot.nodes += wb.acc.nodes;
    if (KIND == kBuildGrad) {
      tot.gx += wb.acc.gx;
      tot.gy += wb.acc.gy;
      tot.gz += wb.acc.gz;
      tot.lap += wb.acc.lap;
      tot.s += wb.acc.s;
    }
  }
  tot.dT = group_sum(tot.dT);
  tot.nodes = group_sum(tot.nodes);
  if (KIND == kBuildGrad) {
    tot.gx = group_sum(tot.gx);
    tot.gy = group_sum(tot.gy);
    tot.gz = group_sum(tot.gz);
    tot.lap = group_sum(tot.lap);
    tot.s = group_sum(tot.s);
  }
  return tot;
}

// abs() at Grid.cpp:381/:535: int truncation (g++ build) or fabs (clang build).
__device__ __forceinline__ double ref_abs(double v, int abs_mode) {
  if (abs_mode == 1) return fabs(v);
  int iv;
  if (!(v == v) || v >= 2147483648.0 || v <= -2147483649.0) iv = INT_MIN;  // x86 cvttsd2si
  else iv = (int)v;
  if (iv == INT_MIN) return (double)INT_MIN;
  return (double)(iv < 0 ? -iv : iv);
}

// Returns the entry's counted evaluations on its group's lane 0 (0 elsewhere).
__device__ double solidify_one(const SolEntry& e, double t, const GridGeom& g, const SolParams& sp, const int* hints,
                               int do_sol, const PointState& st, const DevCounters& dc) {
  const bool lead = group_lane() == 0;
  if (!do_sol) return 0.0;
  const int p = e.p;
  double x, y, z;
  point_xyz(g, p, x, y, z);
  double T2 = e.T, T1 = e.T_last;
  const double T0 = sp.T_liq;
  double t2 = t, t1 = t - sp.dt, t0 = 0.0;
  int iter = 0;
  double nodes = 0.0;
  // With the int abs the iteration below stops after its first T build
  // whenever 0 < T(t0) < 2 T0 (the truncated |1 - T/T0| is 0 < sol_tol), and
  // the outputs come from the refined build at that first t0. That build also
  // yields T(t0) by the refined schedule; well inside the bound (< 1.5 T0) it
  // settles the stop without the first build (whose T differs from it only
  // by quadrature error).
  Accum a;
  bool settled = false;
  if (sp.abs_mode == 0 && sp.sol_tol > 0.0) {
    const double m = (T2 - T1) / (t2 - t1);
    const double tg = t1 + ((T0 - T1) / m);
    a = wave_build_sum<kBuildGrad>(sp, hints, tg, true, sp.compress != 0, x, y, z);
    const double Ts = sp.T_init + a.dT;
    if (Ts > 0.0 && Ts < 1.5 * T0) {
      t0 = tg;
      settled = true;
    }
  }
  while (!settled) {
    const double m = (T2 - T1) / (t2 - t1);
    t0 = t1 + ((T0 - T1) / m);
    // Grid.cpp:375 builds with isSol = false: with compression, the plain
    // list replaces the compressed one when it is no longer (Calc.cpp:49-58)
    bool comp = sp.compress != 0;
    if (comp) {
      const Accum cc = wave_build_sum<kBuildCount>(sp, hints, t0, false, true, x, y, z);
      const Accum cr = wave_build_sum<kBuildCount>(sp, hints, t0, false, false, x, y, z);
      comp = cr.nodes > cc.nodes;
    }
    const Accum at = wave_build_sum<kBuildT>(sp, hints, t0, false, comp, x, y, z);
    if (iter > 0) nodes += at.nodes;  // the first iteration's list is counted by k_solidify_count
    const double Tt = sp.T_init + at.dT;
    const double err = ref_abs(1.0 - Tt / T0, sp.abs_mode);
    ++iter;
    if ((err < sp.sol_tol) || (iter >= sp.sol_iter)) break;
    if (Tt > T0) {
      t1 = t0;
      T1 = Tt;
    } else {
      t2 = t0;
      T2 = Tt;
    }
  }
  if (!settled) a = wave_build_sum<kBuildGrad>(sp, hints, t0, true, sp.compress != 0, x, y, z);
  if (!lead) return 0.0;
  const double Gm = sqrt(a.gx * a.gx + a.gy * a.gy + a.gz * a.gz);
  const double rate = ref_abs(sp.alpha * a.lap + a.s, sp.abs_mode);
  const double V = rate / Gm;
  const double eq =
      (1.0 - exp((-4 * 3.14159265358979323846 * sp.cet_N0 / 3) *
                 pow(Gm * (sp.cet_n + 1) / (pow(sp.cet_a * V, 1 / sp.cet_n)), -3)));
  if (st.sol_log.n) {  // Grid.cpp:322, :547: every solidification appends (t_sol, dTdt)
    const int i = atomicAdd(st.sol_log.n, 1);
    if (i < st.sol_log.cap) {
      st.sol_log.p[i] = p;
      st.sol_log.step[i] = e.step;
      st.sol_log.a[i] = t0;
      st.sol_log.b[i] = rate;
    } else {
      atomicOr(dc.overflow, 1);
    }
  }
  // a later solidification (or, for G/V/eqFrac, a later liquid marking) of
  // this point has already written over these columns in the reference
  if (!st.stamp_sol || st.stamp_sol[p] == e.step) {
    if (st.tSol) st.tSol[p] = t0;
    if (st.Gx) st.Gx[p] = a.gx / Gm;
    if (st.Gy) st.Gy[p] = a.gy / Gm;
    if (st.Gz) st.Gz[p] = a.gz / Gm;
    if (st.dTdt) st.dTdt[p] = rate;
  }
  if (!st.stamp_liq || st.stamp_liq[p] == e.step) {
    if (st.G) st.G[p] = Gm;
    if (st.V) st.V[p] = V;
    if (st.eqFrac) st.eqFrac[p] = eq;
  }
  return nodes + a.nodes;
}

// Node count of one beam's list at time tq as the reference builds it (kept
// nodes only, the weight-0 one included), walked by a single lane.
struct LaneCount {
  const SolParams& sp;
  const PathSeg* P;
  const BeamQ& B;
  double tq;
  double nodes = 0.0;
  __device__ void instant(int seg) {
    RawNode n;
    if (instant_node(P, seg, B, sp.rb, tq, n)) nodes += 1.0;
  }
  // Every node of a segment lies between its end points (a spot: at its
  // point), so on a segment whose ends are inside the R_max box, by a margin
  // far above the rounding of an interpolated position, every node is kept
  // when the segment is powered and its weight is positive (t2 > t1), and
  // none when it is not.
  __device__ bool inside(int seg) const {
    constexpr double m = 1e-9;  // metres
    auto in = [&](const PathSeg& q) {
      return q.x >= sp.rb.xlo + m && q.x <= sp.rb.xhi - m && q.y >= sp.rb.ylo + m && q.y <= sp.rb.yhi - m;
    };
    return in(P[seg]) && (P[seg].mode || in(P[seg - 1]));
  }
  __device__ void interval(double t1, double t2, int seg, int order) {
    if (inside(seg)) {
      if (P[seg].qmod > 0.0 && B.beta > 0.0 && t2 > t1) nodes += order;
      return;
    }
    for (int a = order - 1; a >= 0; --a) {
      RawNode n;
      if (interval_node(P, seg, B, sp.rb, sp.alpha, tq, t1, t2, order, a, sp.gl, n)) nodes += 1.0;
    }
  }
  __device__ void merged(double t1, double t2, int order, double cx, double cy, double q) {
    for (int a = order - 1; a >= 0; --a) {
      q *= B.beta;  // compounds per node, as Calc.cpp:463 does
      RawNode n;
      if (merged_node(B, sp.alpha, tq, t1, t2, order, a, cx, cy, q, sp.gl, n)) nodes += 1.0;
    }
  }
};

// The reference's first regula-falsi iteration (Grid.cpp:366-378) builds the
// plain list (isSol false) at t0 = t1 + (T_L - T1)/m and sums it at the point,
// whatever the iteration does next. Its size, counted one lane per queue
// entry: with compression the smaller of the compressed and plain lists
// (Calc.cpp:49-58), once per wall image (Calc.cpp:62). k_solidify counts the
// builds after it.
// Blocks cblk0 .. gridDim.x - 1 of k_solidify's grid, one lane per entry.
template <int NT>
__device__ void solidify_count(const SolQueue& q, int q0, int q1, int s0, int nsteps, const double* step_t,
                               const int* step_hints, const SolParams& sp, const DevCounters& dc, int cblk0) {
  const int len = min(q1, q.cap);
  const int nblk = gridDim.x - cblk0;
  double n = 0.0;
  for (int w = q0 + (blockIdx.x - cblk0) * NT + threadIdx.x; w < len; w += nblk * NT) {
    const SolEntry e = q.v[w];
    if (e.step < s0 || e.step - s0 >= nsteps) continue;  // k_solidify reports it
    const double t2 = step_t[e.step - s0], t1 = t2 - sp.dt;
    const double m = (e.T - e.T_last) / (t2 - t1);
    const double t0 = t1 + ((sp.T_liq - e.T_last) / m);
    const int* hints = step_hints + (size_t)(e.step - s0) * sp.nbeams;
    double plain = 0.0, comp = 0.0;
    for (int b = 0; b < sp.nbeams; ++b) {
      const PathSeg* P = sp.segs + sp.seg_off[b];
      const int nseg = sp.seg_off[b + 1] - sp.seg_off[b];
      const int seg = locate_segment(P, nseg, t0, hints[b]);
      LaneCount cr{sp, P, sp.beams[b], t0};
      walk_intervals(P, nseg, sp.beams[b], t0, false, seg, cr);
      plain += cr.nodes;
      if (sp.compress) {
        LaneCount cc{sp, P, sp.beams[b], t0};
        walk_compressed(P, nseg, sp.beams[b], sp.rb, sp.alpha, t0, false, seg, cc);
        comp += cc.nodes;
      }
    }
    const double size = (sp.compress && plain > comp) ? comp : plain;
    n += size * (sp.nimages > 0 ? sp.nimages : 1);
  }
  unsigned long long c = (unsigned long long)n;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane_id() == 0 && c) atomicAdd(dc.evals, c);
}

// Grid::Solidify + Calc_Solidification_time + Set_Solidficiaton_Primary
// (Grid.cpp:316-404, :522-548) for queued points, one lane group per entry,
// in blocks 0 .. cblk0 - 1; blocks from cblk0 on count the first iteration's
// lists (solidify_count) at the same time, on other CUs. Blocks of NT = 512 lanes take all
// of a CU's registers.
template <int NT>
__global__ __launch_bounds__(NT) void k_solidify(SolQueue q, int q0, int q1, int s0, int nsteps,
                                                 const double* __restrict__ step_t, const int* __restrict__ step_hints,
                                                 GridGeom g, SolParams sp, int do_sol, PointState st, DevCounters dc,
                                                 int cblk0) {
  if ((int)blockIdx.x >= cblk0) {
    solidify_count<NT>(q, q0, q1, s0, nsteps, step_t, step_hints, sp, dc, cblk0);
    return;
  }
  constexpr int kWaves = NT / kWave;
  const int len = min(q1, q.cap);
  const int group = (blockIdx.x * kWaves + (threadIdx.x >> 6)) * kSolGroups + (lane_id() / kSolGroup);
  double evals = 0.0;  // counted on each group's lane 0, added once per wavefront
  for (int w = q0 + group; w < len; w += cblk0 * kWaves * kSolGroups) {
    const SolEntry e = q.v[w];
    if (e.step < s0 || e.step - s0 >= nsteps) {  // no time or hints for this step
      if (group_lane() == 0) atomicOr(dc.overflow, 1);
      continue;
    }
    // sp stays the kernel argument: a local copy would put its quadrature
    // table, indexed at run time, in scratch memory
    evals += solidify_one(e, step_t[e.step - s0], g, sp, step_hints + (size_t)(e.step - s0) * sp.nbeams, do_sol, st,
                          dc);
  }
  unsigned long long c = (unsigned long long)evals;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane_id() == 0 && c) atomicAdd(dc.evals, c);
}

__global__ __launch_bounds__(kBlock) void k_copy_list(DList src, DList dst, DevCounters dc) {
  const int n = list_len(src);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *dst.n = n;
    if (n > dst.cap) atomicOr(dc.overflow, 1);
  }
  const int m = min(n, dst.cap);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < m; q += gridDim.x * kBlock) dst.v[q] = src.v[q];
}

__global__ __launch_bounds__(kBlock) void k_none_apply(const double* __restrict__ Tnew, int pnum, double T_liq,
                                                       PointState st, SolQueue solid, int* liquid, DevCounters dc) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  bool is_liq = false;
  if (p < pnum) {
    const double Tl = st.T[p];
    const double T = Tnew[p];
    st.T_last[p] = Tl;
    st.T[p] = T;
    if (T >= T_liq) {
      is_liq = true;
      if (st.G) st.G[p] = DBL_MIN;
      if (st.V) st.V[p] = DBL_MIN;
      if (st.eqFrac) st.eqFrac[p] = -1.0;
      st.oflag[p] = 1;
      if (st.stamp_liq) st.stamp_liq[p] = st.step;
    } else if (Tl >= T_liq) {
      queue_solidify(solid, st, p, dc);
    }
  }
  const unsigned long long m = __ballot(is_liq);
  if (lane_id() == 0 && m) atomicAdd(liquid, __popcll(m));
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// n counters to pinned host memory, then the sequence number (system-scope
// release after the data), which the host polls.
__global__ __launch_bounds__(64) void k_export_counts(const int* __restrict__ src, int n, int* dst, unsigned* seq_h,
                                                      unsigned seq) {
  for (int i = threadIdx.x; i < n; i += 64) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(seq_h, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_fill_u8(uint8_t* p, uint8_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

constexpr int kMaxBlocks = 2048;
// Grid size for a list whose length is only bounded on the host (hint < 0:
// unknown). Kernels are grid-stride, so any size is correct.
inline int blocks_for(long long hint, int per_block) {
  if (hint < 0) return kMaxBlocks;
  const long long b = (hint + per_block - 1) / per_block;
  return (int)std::max<long long>(1, std::min<long long>(b, kMaxBlocks));
}

}  // namespace

void launch_reset(const DList& reset, int hint, const PointState& st, int reset_oflag, hipStream_t s) {
  if (hint == 0) return;
  hipLaunchKernelGGL(k_reset, dim3(blocks_for(hint, kBlock)), dim3(kBlock), 0, s, reset, st, reset_oflag);
  THESIS_HIP(hipGetLastError());
}

void launch_remark_liquid(const int* pts, int n, const PointState& st, hipStream_t s) {
  if (n <= 0 || !(st.G || st.V || st.eqFrac)) return;
  hipLaunchKernelGGL(k_remark_liquid, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, pts, n, st);
  THESIS_HIP(hipGetLastError());
}

// Blocks per CU of the LDS-staged list and BFS kernels (each block stages
// the node set once).
static int g_lds_blocks_per_cu = 1;
void set_lds_blocks_per_cu(int v) { g_lds_blocks_per_cu = std::max(1, std::min(v, 16)); }
// CUs held by a solidification range running beside the steps: the grids of
// the LDS-staged kernels leave them out, so that no block of a step waits for
// one to drain (the kernels are grid-stride, any grid is correct).
static int g_reserved_cus = 0;
void set_reserved_cus(int v) { g_reserved_cus = std::max(0, v); }
static int lds_grid(int ncu) { return g_lds_blocks_per_cu * std::max(ncu - g_reserved_cus, 1); }

void launch_eval_list(const NodeSetDev& ns, const GridGeom& g, const DList& pts, int hint, double T_init,
                      double T_liq, const PointState& st, const EvalAction& act, const DevCounters& dc, hipStream_t s) {
  if (hint == 0) return;
  // node sets that fit the LDS (beside the block's counters) are staged there
  static int lds_max = -1;
  if (lds_max < 0) {
    int dev = 0, v = 0;
    THESIS_HIP(hipGetDevice(&dev));
    THESIS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    lds_max = std::max(0, v - 1024);
    THESIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eval_list_lds),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  }
  const size_t bytes = ns.prec == kFP64 ? ns.n64 * sizeof(NodeF64) : ns.n32 * sizeof(NodeF32c);
  if (bytes <= static_cast<size_t>(lds_max)) {
    int dev = 0, ncu = 0;
    THESIS_HIP(hipGetDevice(&dev));
    THESIS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = std::min(blocks_for(hint, kListWaves * kListPPW), lds_grid(ncu));
    hipLaunchKernelGGL(k_eval_list_lds, dim3(blocks), dim3(kListBlock), bytes, s, ns, g, pts, T_init, T_liq, st, act,
                       dc);
  } else {
    hipLaunchKernelGGL(k_eval_list, dim3(blocks_for(hint, kListWaves)), dim3(kListBlock), 0, s, ns, g, pts, T_init,
                       T_liq, st, act, dc);
  }
  THESIS_HIP(hipGetLastError());
}

void launch_claim_neighbors(const DList& test, int hint, const GridGeom& g, int surface_only, const PointState& st,
                            const DList& cand, const DList& reset, const DevCounters& dc, hipStream_t s) {
  if (hint == 0) return;
  hipLaunchKernelGGL(k_claim_neighbors, dim3(blocks_for(hint, kBlock)), dim3(kBlock), 0, s, test, g, surface_only,
                     st, cand, reset, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_claim_seeds(const int* seeds, int count, const PointState& st, const DList& cand, const DList& reset,
                        const DevCounters& dc, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_claim_seeds, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s, seeds, count, st, cand,
                     reset, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_column_expand(const DList& cols, int hint, const NodeSetDev& cur, const NodeSetDev& last,
                          const GridGeom& g, double T_init, double T_liq, const PointState& st, const SolQueue& solid,
                          const DList& reset, const DevCounters& dc, hipStream_t s) {
  if (hint == 0) return;
  hipLaunchKernelGGL(k_column_expand, dim3(blocks_for(hint, kWavesPerBlock)), dim3(kBlock), 0, s, cols, cur, last, g,
                     T_init, T_liq, st, solid, reset, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_calc_depth(const DList& liq, int hint, const NodeSetDev& cur, const NodeSetDev& last, const GridGeom& g,
                       double T_init, double T_liq, const PointState& st, const SolQueue& solid, const DList& reset,
                       const DevCounters& dc, hipStream_t s) {
  if (hint == 0) return;
  hipLaunchKernelGGL(k_calc_depth, dim3(blocks_for(hint, kWavesPerBlock)), dim3(kBlock), 0, s, liq, cur, last, g,
                     T_init, T_liq, st, solid, reset, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_depth_max(const DList& liq, int hint, const GridGeom& g, double T_liq, const PointState& st,
                      hipStream_t s) {
  if (hint == 0) return;
  hipLaunchKernelGGL(k_depth_max, dim3(blocks_for(hint, kBlock)), dim3(kBlock), 0, s, liq, g, T_liq, st);
  THESIS_HIP(hipGetLastError());
}

void launch_solidify(const SolQueue& q, int q0, int q1, int s0, int nsteps, const double* step_t, const int* step_hints,
                     const GridGeom& g, const SolParams& sp, int do_sol, const PointState& st, const DevCounters& dc,
                     hipStream_t s, int max_blocks) {
  if (q1 <= q0) return;
  const int n = q1 - q0;
  const bool count = do_sol && dc.evals;
  if (max_blocks > 0) {
    // a few CU-filling blocks, a quarter of them counting
    constexpr int NT = 2 * kBlock;
    const int cblk = count ? std::min(std::max(max_blocks / 4, 1), blocks_for(n, NT)) : 0;
    const int sblk = std::min(std::max(max_blocks - cblk, 1), blocks_for(n, (NT / kWave) * kSolGroups));
    hipLaunchKernelGGL(k_solidify<NT>, dim3(sblk + cblk), dim3(NT), 0, s, q, q0, q1, s0, nsteps, step_t, step_hints,
                       g, sp, do_sol, st, dc, sblk);
  } else {
    const int sblk = blocks_for(n, kWavesPerBlock * kSolGroups);
    const int cblk = count ? blocks_for(n, kBlock) : 0;
    hipLaunchKernelGGL(k_solidify<kBlock>, dim3(sblk + cblk), dim3(kBlock), 0, s, q, q0, q1, s0, nsteps, step_t,
                       step_hints, g, sp, do_sol, st, dc, sblk);
  }
  THESIS_HIP(hipGetLastError());
}

void launch_copy_list(const DList& src, int hint, const DList& dst, const DevCounters& dc, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_list, dim3(blocks_for(hint, kBlock)), dim3(kBlock), 0, s, src, dst, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_none_apply(const double* Tnew, int pnum, double T_liq, const PointState& st, const SolQueue& solid,
                       int* liquid_count, const DevCounters& dc, hipStream_t s) {
  if (pnum <= 0) return;
  hipLaunchKernelGGL(k_none_apply, dim3((pnum + kBlock - 1) / kBlock), dim3(kBlock), 0, s, Tnew, pnum, T_liq, st,
                     solid, liquid_count, dc);
  THESIS_HIP(hipGetLastError());
}

// Node-set LDS bytes the LDS-staged BFS kernels accept and the CU count, read
// (and the kernels' dynamic LDS limits raised) once.
struct BfsLimits {
  int lds_max = 0, ncu = 1;
};
static const BfsLimits& bfs_limits() {
  static const BfsLimits lim = [] {
    BfsLimits l;
    int dev = 0, v = 0;
    THESIS_HIP(hipGetDevice(&dev));
    THESIS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    THESIS_HIP(hipDeviceGetAttribute(&l.ncu, hipDeviceAttributeMultiprocessorCount, dev));
    l.lds_max = std::max(0, v - 8192);  // beside the won-point list (4 KiB) and the counters
    for (const void* k : {reinterpret_cast<const void*>(&k_bfs_level_lds),
                          reinterpret_cast<const void*>(&k_recheck_bfs_lds)})
      THESIS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, l.lds_max));
    return l;
  }();
  return lim;
}

static size_t node_bytes(const NodeSetDev& ns) {
  return ns.prec == kFP64 ? ns.n64 * sizeof(NodeF64) : ns.n32 * sizeof(NodeF32c);
}

bool recheck_bfs_fits(const NodeSetDev& ns) { return node_bytes(ns) <= static_cast<size_t>(bfs_limits().lds_max); }

void launch_recheck_bfs(const NodeSetDev& ns, const GridGeom& g, const DList& last, int hint, double T_init,
                        double T_liq, const PointState& st, const DList& liq, const DList& next, const DList& reset,
                        const SolQueue& solid, int* zero, const int* seeds, int nseeds, const DevCounters& dc,
                        hipStream_t s) {
  const BfsLimits& lim = bfs_limits();
  const size_t bytes = node_bytes(ns);
  if (bytes > static_cast<size_t>(lim.lds_max))
    throw Error(kBadArgument, "node set too large for the fused step kernel");
  nseeds = std::max(nseeds, 0);
  const long long rounds = hint < 0 ? -1
                                    : (static_cast<long long>(hint) + kRecheckRound - 1) / kRecheckRound +
                                          (nseeds + kBfsLdsBlock - 1) / kBfsLdsBlock;
  const int blocks = std::min(blocks_for(rounds, 1), lds_grid(lim.ncu));
  hipLaunchKernelGGL(k_recheck_bfs_lds, dim3(blocks), dim3(kBfsLdsBlock), bytes, s, ns, g, last, T_init, T_liq, st, liq,
                     next, reset, solid, zero, seeds, nseeds, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_bfs_level(const NodeSetDev& ns, const GridGeom& g, const DList& test, int hint, int surface_only,
                      double T_init, double T_liq, const PointState& st, const DList& liq, const DList& next,
                      const DList& reset, int* zero, const DevCounters& dc, hipStream_t s, const int* seeds,
                      int nseeds) {
  const int lds_max = bfs_limits().lds_max, ncu = bfs_limits().ncu;
  const size_t bytes = node_bytes(ns);
  if (bytes <= static_cast<size_t>(lds_max)) {
    const long long items = hint < 0 ? -1 : static_cast<long long>(hint) * kBfsSlotLanes + std::max(nseeds, 0);
    const int blocks = std::min(blocks_for(items, kBfsLdsBlock), lds_grid(ncu));
    hipLaunchKernelGGL(k_bfs_level_lds, dim3(blocks), dim3(kBfsLdsBlock), bytes, s, ns, g, test, surface_only, T_init,
                       T_liq, st, liq, next, reset, zero, seeds, std::max(nseeds, 0), dc);
  } else {
    // the seeds first, on their own, as the reference claims them before the waves
    if (nseeds > 0) launch_seed_eval(seeds, nseeds, ns, g, T_init, T_liq, st, liq, next, reset, dc, s);
    hipLaunchKernelGGL(k_bfs_level, dim3(blocks_for(hint, kBfsWaves)), dim3(kBfsWaves * kWave), 0, s, ns, g, test,
                       surface_only, T_init, T_liq, st, liq, next, reset, zero, dc);
  }
  THESIS_HIP(hipGetLastError());
}

void launch_seed_eval(const int* seeds, int count, const NodeSetDev& ns, const GridGeom& g, double T_init,
                      double T_liq, const PointState& st, const DList& liq, const DList& next, const DList& reset,
                      const DevCounters& dc, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_seed_eval, dim3(blocks_for(count, kWavesPerBlock)), dim3(kBlock), 0, s, seeds, count, ns, g,
                     T_init, T_liq, st, liq, next, reset, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_step_prologue(const DList& old_reset, int hint, const DList& last, const DList& new_reset,
                          const DList& frontier, int* z0, int* z1, int* z2, int* z3, const PointState& st,
                          const DevCounters& dc, hipStream_t s, double tm_time, double T_liq, int preclaim,
                          const UploadRange* up_ranges, const unsigned short* up_map, unsigned up_blocks) {
  const int pb = blocks_for(hint, kBlock);
  if (up_blocks > 0)
    hipLaunchKernelGGL(k_step_start, dim3(up_blocks + pb), dim3(kBlock), 0, s, up_ranges, up_map, up_blocks, old_reset,
                       last, new_reset, frontier, z0, z1, z2, z3, st, dc, tm_time, T_liq, preclaim);
  else
    hipLaunchKernelGGL(k_step_prologue, dim3(pb), dim3(kBlock), 0, s, old_reset, last, new_reset, frontier, z0, z1, z2,
                       z3, st, dc, tm_time, T_liq, preclaim);
  THESIS_HIP(hipGetLastError());
}

void launch_stork_depth(const DList& liq, int hint, const NodeSetDev& ns, const GridGeom& g, double T_init,
                        double T_liq, const PointState& st, uint8_t* is_liq, const DevCounters& dc, hipStream_t s) {
  if (hint <= 0) return;
  hipLaunchKernelGGL(k_stork_depth, dim3(blocks_for(hint, kWavesPerBlock)), dim3(kBlock), 0, s, liq, ns, g, T_init,
                     T_liq, st, is_liq, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_stork_cells(const GridGeom& g, const uint8_t* is_liq, const uint8_t* c_prev, uint8_t* c_cur,
                        const DList& rel, const DevCounters& dc, hipStream_t s) {
  const long long n = (long long)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
  if (n <= 0) return;
  const int blocks = (int)std::min<long long>((n + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_stork_cells, dim3(blocks), dim3(kBlock), 0, s, g, is_liq, c_prev, c_cur, rel, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_stork_records(const int* rel, int n, const GridGeom& g, const NodeSetDev& prev, const NodeSetDev& cur,
                          double T_init, double* out, const DevCounters& dc, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_stork_records, dim3(blocks_for(16LL * n, kWavesPerBlock)), dim3(kBlock), 0, s, rel, n, g,
                     prev, cur, T_init, out, dc);
  THESIS_HIP(hipGetLastError());
}

void launch_export_counts(const int* src, int n, int* dst_h, unsigned* seq_h, unsigned seq, hipStream_t s) {
  hipLaunchKernelGGL(k_export_counts, dim3(1), dim3(64), 0, s, src, n, dst_h, seq_h, seq);
  THESIS_HIP(hipGetLastError());
}

void launch_set_int(int* p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v);
  THESIS_HIP(hipGetLastError());
}

void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks), dim3(kBlock), 0, s, p, v, n);
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
