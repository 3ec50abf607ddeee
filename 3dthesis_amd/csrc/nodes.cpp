#include "nodes.hpp"

#include <cmath>

namespace thesis {

PathData make_path_data(const Problem& pb) {
  PathData pd;
  pd.alpha = pb.mat.alpha;
  pd.off.push_back(0);
  for (size_t b = 0; b < pb.paths.size(); ++b) {
    for (const Segment& s : pb.paths[b]) {
      if (s.mode != 0 && s.mode != 1) throw Error(kBadInput, "path segment mode must be 0 (line) or 1 (spot)");
      pd.segs.push_back(PathSeg{s.x, s.y, s.z, s.qmod, s.param, s.time, s.mode, 0});
    }
    pd.off.push_back(static_cast<int>(pd.segs.size()));
    const BeamSpec& bs = pb.beams[b];
    BeamQ q{};
    q.ax2 = bs.ax * bs.ax;
    q.ay2 = bs.ay * bs.ay;
    q.az2 = bs.az * bs.az;
    q.ax = bs.ax;
    q.az = bs.az;
    q.nond_dt = bs.nond_dt;
    // Calc.cpp:96
    q.beta = std::pow(3.0 / 3.14159265358979323846, 1.5) * bs.q / (pb.mat.rho * pb.mat.cps);
    pd.beams.push_back(q);
  }
  pd.rbox = RBox{pb.grid.xmin - pb.r_max, pb.grid.xmax + pb.r_max, pb.grid.ymin - pb.r_max,
                 pb.grid.ymax + pb.r_max};
  return pd;
}

NodeF64 to_f64(const RawNode& r) {
  NodeF64 n;
  n.xb = r.xb;
  n.yb = r.yb;
  n.zb = r.zb;
  n.ix = 1.0 / r.px;
  n.iy = 1.0 / r.py;
  n.iz = 1.0 / r.pz;
  n.dtau = r.dtau;
  n.expmod = node_expmod(r);
  return n;
}

Quadrature::Quadrature(const Problem& pb) : pd_(make_path_data(pb)), gl_(gl_table()) { reset_hints(); }

void Quadrature::reset_hints() { hint_.assign(pd_.beams.size(), 1); }

namespace {
struct Collect {
  std::vector<NodeF64>& out;
  void operator()(const RawNode& r) { out.push_back(to_f64(r)); }
};
}  // namespace

void Quadrature::build(double t, bool is_sol, std::vector<NodeF64>& out) {
  for (size_t b = 0; b < pd_.beams.size(); ++b) {
    const PathSeg* P = pd_.segs.data() + pd_.off[b];
    const int nseg = pd_.off[b + 1] - pd_.off[b];
    const int seg = locate_segment(P, nseg, t, hint_[b]);
    if (!is_sol) hint_[b] = seg;
    Collect c{out};
    beam_nodes(P, nseg, pd_.beams[b], pd_.rbox, pd_.alpha, t, is_sol, seg, gl_, c);
  }
}

void Quadrature::build_from(double t, bool is_sol, const std::vector<int>& hints, std::vector<NodeF64>& out) const {
  for (size_t b = 0; b < pd_.beams.size(); ++b) {
    const PathSeg* P = pd_.segs.data() + pd_.off[b];
    const int nseg = pd_.off[b + 1] - pd_.off[b];
    const int seg = locate_segment(P, nseg, t, hints[b]);
    Collect c{out};
    beam_nodes(P, nseg, pd_.beams[b], pd_.rbox, pd_.alpha, t, is_sol, seg, gl_, c);
  }
}

size_t pack_f32c(const NodeF64* in, size_t n, NodeF32c* out) {
  const double log2e = 1.4426950408889634074;
  size_t m = 0;
  for (size_t i = 0; i < n; ++i) {
    const NodeF64& a = in[i];
    if (!(a.dtau > 0.0)) continue;  // weight-0 nodes add nothing to T
    NodeF32c& b = out[m++];
    b.xb = a.xb;
    b.yb = a.yb;
    b.zb = a.zb;
    b.nx = float(-3.0 * log2e * a.ix);
    b.ny = float(-3.0 * log2e * a.iy);
    b.nz = float(-3.0 * log2e * a.iz);
    b.c = float(log2e * (a.expmod + std::log(a.dtau)));
  }
  return m;
}

}  // namespace thesis
