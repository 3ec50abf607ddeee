// Batched snapshots of many single-beam parameter sets on one grid.
#pragma once

#include <vector>

#include "device.hpp"
#include "model.hpp"
#include "quadrature.hpp"

namespace thesis {

// Device view of the sets' paths and beams.
struct BatchDev {
  int nsets;
  const PathSeg* segs;
  const int* seg_off;  // nsets+1
  const BeamQ* beams;
  const RBox* rbox;
  const double* times;
  double alpha;
};

struct BatchSumArgs {
  GridGeom g;
  int nsets;
  const int* order;     // work order of the sets (most nodes first)
  const int* node_off;  // nsets+1, node ranges per set
  const NodeF32c* n32;
  const NodeF64* n64;
  double T0;
  double* T;   // nsets x pnum
  int* queue;  // work counter, zero before the launch
};

void launch_count_nodes(const BatchDev& bd, int* counts, hipStream_t s);
void launch_emit_nodes(const BatchDev& bd, const int* node_off, NodeF64* o64, NodeF32c* o32, hipStream_t s);
void launch_tsum_batched(const BatchSumArgs& a, int precision, int ppt, int blocks, hipStream_t s);

// Paths of one set from Path-file rows (mode, x mm, y mm, z mm, Pmod, speed|time):
// the implicit start segment at the origin plus the rows, timed like Init.cpp:607-622.
std::vector<Segment> make_path(const double* rows6, int nrows);

}  // namespace thesis
