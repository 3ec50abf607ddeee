// Host side of the batched snapshot solve: per-set paths/beams from Path-file
// rows and beam parameters, device node build, persistent launch, autotuner.
#include "batch_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

namespace thesis {

BatchRunner::BatchRunner(const Problem& base, hipStream_t stream) : base_(base), stream_(stream) {
  const GridSpec& g = base_.grid;
  geom_ = GridGeom{g.xnum, g.ynum, g.znum, 0, g.xmin, g.ymin, g.zmin, g.xmax, g.ymax, g.zmax, g.sx, g.sy, g.sz};
  queue_.ensure(1);
  hipDeviceProp_t prop;
  int dev = 0;
  THESIS_HIP(hipGetDevice(&dev));
  THESIS_HIP(hipGetDeviceProperties(&prop, dev));
  cus_ = prop.multiProcessorCount;
}

void BatchRunner::set(int nsets, const double* beam5, const int* row_off, const double* rows6, const double* times) {
  if (nsets <= 0) throw Error(kBadArgument, "a batch needs at least one set");
  if (base_.compress || base_.walls.any())
    throw Error(kBadArgument, "batched snapshots build plain node lists: path compression and wall images are off here");
  if (base_.custom_points) throw Error(kBadArgument, "batched snapshots run on the deck's grid, not on custom points");
  nsets_ = nsets;
  std::vector<PathSeg> segs;
  std::vector<int> seg_off{0};
  std::vector<BeamQ> beams;
  std::vector<RBox> rbox;
  std::vector<double> tv;
  for (int s = 0; s < nsets; ++s) {
    const int r0 = row_off[s], r1 = row_off[s + 1];
    if (r1 <= r0) throw Error(kBadArgument, "every set needs at least one path row");
    const std::vector<Segment> path = make_path(rows6 + 6 * r0, r1 - r0);
    for (const Segment& g : path) {
      if (g.mode != 0 && g.mode != 1) throw Error(kBadArgument, "path row mode must be 0 or 1");
      segs.push_back(PathSeg{g.x, g.y, g.z, g.qmod, g.param, g.time, g.mode, 0});
    }
    seg_off.push_back(static_cast<int>(segs.size()));
    const double* b = beam5 + 5 * s;  // Power, Efficiency, Width_X, Width_Y, Depth_Z
    BeamSpec bs{};
    bs.power = b[0];
    bs.eff = b[1];
    bs.q = b[0] * b[1] * 2.0;
    bs.ax = b[2];
    bs.ay = b[3];
    bs.az = b[4];
    bs.nond_dt = bs.ax * bs.ax / base_.mat.alpha;
    if (!(bs.ax > 0 && bs.ay > 0 && bs.az > 0)) throw Error(kBadArgument, "beam widths must be positive");
    BeamQ q{};
    q.ax2 = bs.ax * bs.ax;
    q.ay2 = bs.ay * bs.ay;
    q.az2 = bs.az * bs.az;
    q.ax = bs.ax;
    q.az = bs.az;
    q.nond_dt = bs.nond_dt;
    q.beta = std::pow(3.0 / 3.14159265358979323846, 1.5) * bs.q / (base_.mat.rho * base_.mat.cps);
    beams.push_back(q);
    const double r = beam_radius_of_influence(base_, bs);
    rbox.push_back(RBox{base_.grid.xmin - r, base_.grid.xmax + r, base_.grid.ymin - r, base_.grid.ymax + r});
    tv.push_back(times ? times[s] : path.back().time);
  }
  d_segs_.ensure(segs.size());
  d_seg_off_.ensure(seg_off.size());
  d_beams_.ensure(beams.size());
  d_rbox_.ensure(rbox.size());
  d_times_.ensure(tv.size());
  THESIS_HIP(hipMemcpy(d_segs_.get(), segs.data(), segs.size() * sizeof(PathSeg), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_seg_off_.get(), seg_off.data(), seg_off.size() * sizeof(int), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_beams_.get(), beams.data(), beams.size() * sizeof(BeamQ), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_rbox_.get(), rbox.data(), rbox.size() * sizeof(RBox), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_times_.get(), tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice));
  build_nodes();
}

BatchDev BatchRunner::dev_view() const {
  return BatchDev{nsets_, d_segs_.get(), d_seg_off_.get(), d_beams_.get(), d_rbox_.get(), d_times_.get(),
                  base_.mat.alpha};
}

void BatchRunner::build_nodes() {
  const BatchDev bd = dev_view();
  d_counts_.ensure(static_cast<size_t>(nsets_));
  launch_count_nodes(bd, d_counts_.get(), stream_);
  counts_.resize(static_cast<size_t>(nsets_));
  THESIS_HIP(hipMemcpyAsync(counts_.data(), d_counts_.get(), nsets_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
  THESIS_HIP(hipStreamSynchronize(stream_));
  node_off_.assign(static_cast<size_t>(nsets_) + 1, 0);
  for (int s = 0; s < nsets_; ++s) node_off_[s + 1] = node_off_[s] + counts_[s];
  total_nodes_ = node_off_[nsets_];
  d_node_off_.ensure(node_off_.size());
  THESIS_HIP(hipMemcpy(d_node_off_.get(), node_off_.data(), node_off_.size() * sizeof(int), hipMemcpyHostToDevice));
  n64_.ensure(static_cast<size_t>(std::max<long long>(total_nodes_, 1)));
  n32_.ensure(static_cast<size_t>(std::max<long long>(total_nodes_, 1)));
  launch_emit_nodes(bd, d_node_off_.get(), n64_.get(), n32_.get(), stream_);
  // most nodes first, so the persistent grid ends on cheap work
  std::vector<int> order(static_cast<size_t>(nsets_));
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return counts_[a] > counts_[b]; });
  d_order_.ensure(order.size());
  THESIS_HIP(hipMemcpy(d_order_.get(), order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
  THESIS_HIP(hipStreamSynchronize(stream_));
}

double BatchRunner::evals() const { return static_cast<double>(total_nodes_) * static_cast<double>(base_.grid.pnum); }

void BatchRunner::nodes_of(int set, std::vector<NodeF64>& out) const {
  if (set < 0 || set >= nsets_) throw Error(kBadArgument, "set index out of range");
  const int n = counts_[set];
  out.resize(static_cast<size_t>(n));
  if (n)
    THESIS_HIP(hipMemcpy(out.data(), n64_.get() + node_off_[set], n * sizeof(NodeF64), hipMemcpyDeviceToHost));
}

void BatchRunner::eval(int precision, double* T_dev, hipStream_t s, int ppt, int blocks) {
  if (nsets_ == 0) throw Error(kBadArgument, "no batch set");
  if (ppt <= 0) ppt = tuned_ppt_;
  if (blocks <= 0) blocks = tuned_blocks_ > 0 ? tuned_blocks_ : 4 * cus_;
  BatchSumArgs a{};
  a.g = geom_;
  a.nsets = nsets_;
  a.order = d_order_.get();
  a.node_off = d_node_off_.get();
  a.n32 = n32_.get();
  a.n64 = n64_.get();
  a.T0 = base_.mat.T_init;
  a.T = T_dev;
  a.queue = queue_.get();
  launch_set_int(queue_.get(), 0, s);
  launch_tsum_batched(a, precision, ppt, blocks, s);
}

void BatchRunner::autotune(int precision, double* T_dev, hipStream_t s, int* ppt_out, int* blocks_out) {
  hipEvent_t e0, e1;
  THESIS_HIP(hipEventCreate(&e0));
  THESIS_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_ppt = 2, best_blocks = 4 * cus_;
  const int ppts[3] = {1, 2, 4};
  const int per_cu[3] = {2, 4, 8};
  for (int pi = 0; pi < (precision == kFP64 ? 1 : 3); ++pi) {
    for (int bi = 0; bi < 3; ++bi) {
      const int ppt = ppts[pi], blocks = per_cu[bi] * cus_;
      eval(precision, T_dev, s, ppt, blocks);  // warm
      THESIS_HIP(hipEventRecord(e0, s));
      eval(precision, T_dev, s, ppt, blocks);
      THESIS_HIP(hipEventRecord(e1, s));
      THESIS_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      THESIS_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_ppt = ppt;
        best_blocks = blocks;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tuned_ppt_ = best_ppt;
  tuned_blocks_ = best_blocks;
  if (ppt_out) *ppt_out = best_ppt;
  if (blocks_out) *blocks_out = best_blocks;
}

}  // namespace thesis
