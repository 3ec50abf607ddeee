// Melt-pool statistics (MP_Stats output, Melt::calc_mp_info, Melt.cpp:359-580):
// after each Surface tracking step, the liquid surface points are grouped into
// local pools, one per powered path segment of the first beam (seeded at the
// segment's start cell and, for the segment under the beam, the beam's current
// cell); each pool's extent along and across its segment direction and the
// deepest liquid column of the grid give running maxima (width, length, depth)
// on every point of the pool's columns. Host code: per step it touches only
// the liquid surface set.
#pragma once

#include <vector>

#include "model.hpp"
#include "quadrature.hpp"

namespace thesis {

class MeltPoolStats {
 public:
  explicit MeltPoolStats(const Problem& pb);
  // One step at time t: liq = the liquid surface points of the step (in any
  // order, each once), depths = per-column depth indices (xnum*ynum).
  void step(double t, const std::vector<int>& liq, const std::vector<int>& depths);
  const std::vector<double>& width() const { return w_; }
  const std::vector<double>& length() const { return l_; }
  const std::vector<double>& depth() const { return d_; }

 private:
  struct Pool {
    std::vector<int> test;
    int seg;
    double sn, cs;
  };
  void cell_points(double x, double y, std::vector<int>& out) const;
  const Problem& pb_;
  std::vector<PathSeg> path_;
  int seg_ = 0;
  std::vector<Pool> pools_;  // oldest first
  std::vector<uint8_t> mark_;  // liquid surface points not yet claimed this step
  std::vector<double> w_, l_, d_;
};

}  // namespace thesis
