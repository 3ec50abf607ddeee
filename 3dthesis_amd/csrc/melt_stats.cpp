#include "melt_stats.hpp"

#include <cfloat>
#include <cmath>

namespace thesis {

MeltPoolStats::MeltPoolStats(const Problem& pb) : pb_(pb) {
  for (const Segment& s : pb.paths[0]) path_.push_back(PathSeg{s.x, s.y, s.z, s.qmod, s.param, s.time, s.mode, 0});
  const size_t n = static_cast<size_t>(pb.grid.pnum);
  mark_.assign(n, 0);
  w_.assign(n, 0.0);
  l_.assign(n, 0.0);
  d_.assign(n, 0.0);
}

// The 2x2 surface points at the corners of the cell holding (x, y)
// (Melt.cpp:384-409); the y clamp compares with xnum as the reference does.
void MeltPoolStats::cell_points(double x, double y, std::vector<int>& out) const {
  const GridSpec& g = pb_.grid;
  int xg = static_cast<int>(std::floor((x - g.xmin) / g.xres));
  int yg = static_cast<int>(std::floor((y - g.ymin) / g.yres));
  const int zg = g.znum - 1;
  if (xg < 0) xg = -1;
  if (xg > g.xnum - 1) xg = g.xnum - 1;
  if (yg < 0) yg = -1;
  if (yg > g.xnum - 1) yg = g.ynum - 1;
  for (int dx = 0; dx <= 1; ++dx)
    for (int dy = 0; dy <= 1; ++dy) {
      const int xx = xg + dx, yy = yg + dy;
      if (xx < 0 || xx > g.xnum - 1 || yy < 0 || yy > g.ynum - 1) continue;
      out.push_back(zg + g.znum * yy + g.znum * g.ynum * xx);
    }
}

void MeltPoolStats::step(double t, const std::vector<int>& liq, const std::vector<int>& depths) {
  const GridSpec& g = pb_.grid;
  const int npath = static_cast<int>(path_.size());
  while ((t > path_[seg_].time) && (seg_ + 1 < npath)) {
    const int s = ++seg_;
    if (!(path_[s].qmod > 0)) continue;
    Pool pl;
    const int sp = path_[s].mode == 0 ? s - 1 : s;
    cell_points(path_[sp].x, path_[sp].y, pl.test);
    pl.seg = s;
    const double angle =
        path_[s].mode == 1 ? 0.0 : std::atan2(path_[s].y - path_[s - 1].y, path_[s].x - path_[s - 1].x);
    pl.sn = std::sin(angle);
    pl.cs = std::cos(angle);
    pools_.push_back(std::move(pl));
  }
  if (pools_.empty()) return;
  if (pools_.back().seg == seg_ && path_[seg_].mode == 0) {
    const RBox everywhere{-HUGE_VAL, HUGE_VAL, -HUGE_VAL, HUGE_VAL};
    double xb, yb, zb, q;
    beam_at(path_.data(), seg_, t, everywhere, xb, yb, zb, q);
    cell_points(xb, yb, pools_.back().test);
  }
  for (int p : liq) mark_[p] = 1;
  int dmax = depths.empty() ? 0 : depths[0];
  for (int v : depths) dmax = dmax < v ? v : dmax;
  const int ny = g.ynum, nz = g.znum;
  std::vector<int> found, next;
  for (size_t e = pools_.size(); e-- > 0;) {
    Pool& pl = pools_[e];
    // Melt::local_neighbor_check (Melt.cpp:540-580): claim liquid 3x3
    // neighbours at the same layer, wave by wave
    found.clear();
    while (!pl.test.empty()) {
      next.clear();
      for (int p : pl.test) {
        const int i = p / (ny * nz), j = (p / nz) % ny, k = p % nz;
        for (int di = -1; di <= 1; ++di)
          for (int dj = -1; dj <= 1; ++dj) {
            const int ni = i + di, nj = j + dj;
            if (ni < 0 || ni > g.xnum - 1 || nj < 0 || nj > ny - 1) continue;
            const int q = (ni * ny + nj) * nz + k;
            if (mark_[q]) {
              mark_[q] = 0;
              next.push_back(q);
              found.push_back(q);
            }
          }
      }
      pl.test.swap(next);
    }
    if (found.empty()) {
      pools_.erase(pools_.begin() + static_cast<long>(e));
      continue;
    }
    pl.test = found;
    double minX = DBL_MAX, maxX = -DBL_MAX, minY = DBL_MAX, maxY = -DBL_MAX;
    for (int p : found) {
      const int i = p / (ny * nz), j = (p / nz) % ny;
      const double x = g.xnum > 1 ? g.xmin + double(i) * g.sx : g.xmax;
      const double y = g.ynum > 1 ? g.ymin + double(j) * g.sy : g.ymax;
      const double xr = x * pl.cs + y * pl.sn;
      minX = (minX < xr) ? minX : xr;  // std::min(x_rot, minRotX)
      maxX = (xr < maxX) ? maxX : xr;  // std::max(x_rot, maxRotX)
      const double yr = -x * pl.sn + y * pl.cs;
      minY = (minY < yr) ? minY : yr;
      maxY = (yr < maxY) ? maxY : yr;
    }
    const double width = maxY - minY, length = maxX - minX;
    const double depth = g.xres * dmax;  // the x spacing, as Melt.cpp:511 has it
    for (int p : found) {
      const int i = p / (ny * nz), j = (p / nz) % ny;
      for (int d = 0; d <= depths[i * ny + j]; ++d) {
        const int q = (i * ny + j) * nz + (nz - 1 - d);
        w_[q] = (width < w_[q]) ? w_[q] : width;  // std::max(new, old)
        l_[q] = (length < l_[q]) ? l_[q] : length;
        d_[q] = (depth < d_[q]) ? d_[q] : depth;
      }
    }
  }
  for (int p : liq) mark_[p] = 0;
}

}  // namespace thesis
