// Device-side data layouts and kernel launch interfaces.
//
// Layout in HBM:
//   * FP64 node set: array-of-structs NodeF64 (64 B per node: centre, inverted
//     spreads, weight, log prefactor), so one wavefront-uniform node is one
//     scalar-cache line fetch.
//   * FP32-compensated node set: NodeF32c (40 B: FP64 centre for the
//     tile-relative subtraction, FP32 exponent coefficients pre-scaled to
//     base 2 with ln(dtau) folded into the constant). Weight-0 (instantaneous)
//     nodes are dropped: they add nothing to T.
//   * Points: implicit from a structured grid descriptor (index p = (i*ny + j)*nz + k),
//     or an explicit int32 list of grid indices; with a custom point file,
//     an FP64 x y z array in the descriptor.
//   * Per-point state (T, T_last, solidification columns) as flat FP64 arrays
//     of pnum entries; flags as uint8.
#pragma once

#include "common.hpp"

namespace thesis {

struct NodeF64 {
  double xb, yb, zb, ix, iy, iz, dtau, expmod;
};

struct NodeF32c {
  double xb, yb, zb;
  float nx, ny, nz, c;  // -3 log2(e)/phi per axis, log2(e)*(expmod + ln dtau)
};

struct GridGeom {
  int nx, ny, nz, pad;
  double xmin, ymin, zmin;
  double xmax, ymax, zmax;
  double sx, sy, sz;  // 0 when that axis has one point (coordinate = max)
  // Custom points (device, x y z per point): coordinates come from here and
  // the grid numbers are unused (Grid.cpp:17-27).
  const double* pts = nullptr;
};

THD double grid_x(const GridGeom& g, int i) { return g.nx > 1 ? g.xmin + double(i) * g.sx : g.xmax; }
THD double grid_y(const GridGeom& g, int j) { return g.ny > 1 ? g.ymin + double(j) * g.sy : g.ymax; }
THD double grid_z(const GridGeom& g, int k) { return g.nz > 1 ? g.zmin + double(k) * g.sz : g.zmax; }
THD void grid_ijk(const GridGeom& g, int p, int& i, int& j, int& k) {
  const int nyz = g.ny * g.nz;
  i = p / nyz;
  const int r = p - i * nyz;
  j = r / g.nz;
  k = r - j * g.nz;
}

// Coordinates of point p: the custom point, or the grid point (Grid.h:310-327).
THD void point_coords(const GridGeom& g, int p, double& x, double& y, double& z) {
  if (g.pts) {
    x = g.pts[3 * static_cast<size_t>(p)];
    y = g.pts[3 * static_cast<size_t>(p) + 1];
    z = g.pts[3 * static_cast<size_t>(p) + 2];
    return;
  }
  int i, j, k;
  grid_ijk(g, p, i, j, k);
  x = grid_x(g, i);
  y = grid_y(g, j);
  z = grid_z(g, k);
}

enum Precision : int { kFP64 = 0, kFP32c = 1 };

// Convert an FP64 node set into the FP32-compensated layout (host side).
// Returns the number of kept nodes.
size_t pack_f32c(const NodeF64* in, size_t n, NodeF32c* out);

// Launchers (stream-ordered, no synchronisation). T receives T_init + sum.
// Grid form: points p in [p_begin, p_begin + count); T indexed by p - p_begin.
void launch_tsum_grid_f64(const NodeF64* nodes, int n, const GridGeom& g, int p_begin, int count, double T0,
                          double* T, hipStream_t s);
void launch_tsum_grid_f32c(const NodeF32c* nodes, int n, const GridGeom& g, int p_begin, int count, double T0,
                           double* T, int points_per_thread, hipStream_t s);
// List form: T[q] for point index list[q].
void launch_tsum_list_f64(const NodeF64* nodes, int n, const GridGeom& g, const int* list, int count, double T0,
                          double* T, hipStream_t s);
void launch_tsum_list_f32c(const NodeF32c* nodes, int n, const GridGeom& g, const int* list, int count, double T0,
                           double* T, hipStream_t s);

// Allocation-free grid evaluation: nodes_dev is n node-major FP64 records in
// device memory. FP64 sums read them in place; FP32c packs them on the device
// into the library's file-scope scratch array (at most static_node_capacity()
// nodes) and sums from there. Structured grids only.
int static_node_capacity();
void launch_tsum_static(const double* nodes_dev, int n, int precision, const GridGeom& g, int p_begin, int count,
                        double T0, double* T, int points_per_thread, hipStream_t s);

}  // namespace thesis
