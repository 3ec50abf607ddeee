#include "nodes.hpp"

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>

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
  pd.compress = pb.compress != 0;
  pd.images = wall_images(pb);
  return pd;
}

std::vector<WallImage> wall_images(const Problem& pb) {
  const WallSpec& w = pb.walls;
  if (!w.any()) return {};
  struct C3 {
    int v[3];
    bool operator==(const C3& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
  };
  // planes in the reference's order: X_min, X_max, Y_min, Y_max, Z_min, top surface
  const double plane[6] = {w.xmin, w.xmax, w.ymin, w.ymax, w.zmin, pb.grid.zmax};
  std::vector<C3> all{C3{{0, 0, 0}}}, check = all, fresh;
  const int rounds = w.reflections != INT_MAX ? w.reflections : 0;
  for (int it = 0; !check.empty() && it < rounds; ++it) {
    fresh.clear();
    for (const C3& c : check)
      for (int p = 0; p < 6; ++p) {
        if (plane[p] == DBL_MAX) continue;
        C3 r = c;
        const int ax = p / 2;
        r.v[ax] = (p % 2 == 0) ? (-1 - r.v[ax]) : (1 - r.v[ax]);
        if (p == 5 && r.v[2] == 1) continue;  // never above the top surface
        bool seen = false;
        for (const C3& a : all) seen = seen || a == r;
        if (seen) continue;
        all.push_back(r);
        fresh.push_back(r);
      }
    check.swap(fresh);
  }
  const double lo[3] = {w.xmin, w.ymin, w.zmin}, hi[3] = {w.xmax, w.ymax, pb.grid.zmax};
  std::vector<WallImage> out;
  for (const C3& c : all) {
    double o[3], sg[3];
    for (int ax = 0; ax < 3; ++ax) {
      const int r = c.v[ax], n = std::abs(r % 2);
      // an unset plane is DBL_MAX and only ever meets a zero factor here
      o[ax] = static_cast<double>(r + n) * hi[ax] - static_cast<double>(r - n) * lo[ax];
      sg[ax] = static_cast<double>(1 - 2 * n);
    }
    out.push_back(WallImage{o[0], sg[0], o[1], sg[1], o[2], sg[2]});
  }
  return out;
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

void Quadrature::build_at(double t, bool is_sol, const std::vector<int>& segs, std::vector<NodeF64>& out) const {
  const size_t base = out.size();
  auto beams = [&](bool compressed, std::vector<NodeF64>& dst) {
    for (size_t b = 0; b < pd_.beams.size(); ++b) {
      const PathSeg* P = pd_.segs.data() + pd_.off[b];
      const int nseg = pd_.off[b + 1] - pd_.off[b];
      Collect c{dst};
      beam_nodes(P, nseg, pd_.beams[b], pd_.rbox, pd_.alpha, t, is_sol, segs[b], gl_, c, compressed);
    }
  };
  beams(pd_.compress, out);
  if (pd_.compress && !is_sol) {
    std::vector<NodeF64> plain;
    beams(false, plain);
    if (plain.size() <= out.size() - base) {
      out.resize(base);
      out.insert(out.end(), plain.begin(), plain.end());
    }
  }
  if (!pd_.images.empty()) {
    const size_t n = out.size() - base;
    out.resize(base + n * pd_.images.size());
    // image 0 is the identity map; the reference still rewrites the centres
    for (size_t m = pd_.images.size(); m-- > 0;) {
      const WallImage& w = pd_.images[m];
      for (size_t i = 0; i < n; ++i) {
        NodeF64 nd = out[base + i];
        nd.xb = w.ox + w.sx * nd.xb;
        nd.yb = w.oy + w.sy * nd.yb;
        nd.zb = w.oz + w.sz * nd.zb;
        out[base + m * n + i] = nd;
      }
    }
  }
}

void Quadrature::build(double t, bool is_sol, std::vector<NodeF64>& out) {
  std::vector<int> segs(pd_.beams.size());
  for (size_t b = 0; b < pd_.beams.size(); ++b) {
    const PathSeg* P = pd_.segs.data() + pd_.off[b];
    segs[b] = locate_segment(P, pd_.off[b + 1] - pd_.off[b], t, hint_[b]);
    if (!is_sol) hint_[b] = segs[b];
  }
  build_at(t, is_sol, segs, out);
}

void Quadrature::build_from(double t, bool is_sol, const std::vector<int>& hints, std::vector<NodeF64>& out) const {
  std::vector<int> segs(pd_.beams.size());
  for (size_t b = 0; b < pd_.beams.size(); ++b) {
    const PathSeg* P = pd_.segs.data() + pd_.off[b];
    segs[b] = locate_segment(P, pd_.off[b + 1] - pd_.off[b], t, hints[b]);
  }
  build_at(t, is_sol, segs, out);
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
