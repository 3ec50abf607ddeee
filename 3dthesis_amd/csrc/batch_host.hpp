#pragma once

#include <vector>

#include "batch.hpp"
#include "engine.hpp"

namespace thesis {

// Many single-beam parameter sets on the grid and material of a base problem.
class BatchRunner {
 public:
  BatchRunner(const Problem& base, hipStream_t stream);

  // beam5: per set Power, Efficiency, Width_X, Width_Y, Depth_Z. Paths: set s
  // owns Path-file rows [row_off[s], row_off[s+1]) of rows6. times: snapshot
  // time per set (nullptr: each set's scan end). Builds the node sets on the GPU.
  void set(int nsets, const double* beam5, const int* row_off, const double* rows6, const double* times);

  // One persistent launch: T_dev[s * pnum + p] for every set and grid point.
  void eval(int precision, double* T_dev, hipStream_t s, int ppt = 0, int blocks = 0);
  // Time the tile/occupancy candidates on this batch and keep the fastest.
  void autotune(int precision, double* T_dev, hipStream_t s, int* ppt_out, int* blocks_out);

  int nsets() const { return nsets_; }
  long long total_nodes() const { return total_nodes_; }
  double evals() const;
  void nodes_of(int set, std::vector<NodeF64>& out) const;

 private:
  BatchDev dev_view() const;
  void build_nodes();

  Problem base_;
  hipStream_t stream_;
  GridGeom geom_{};
  int cus_ = 256;
  int nsets_ = 0;
  long long total_nodes_ = 0;
  std::vector<int> counts_, node_off_;
  DevBuf<PathSeg> d_segs_;
  DevBuf<int> d_seg_off_, d_counts_, d_node_off_, d_order_, queue_;
  DevBuf<BeamQ> d_beams_;
  DevBuf<RBox> d_rbox_;
  DevBuf<double> d_times_;
  DevBuf<NodeF64> n64_;
  DevBuf<NodeF32c> n32_;
  int tuned_ppt_ = 2, tuned_blocks_ = 0;
};

}  // namespace thesis
