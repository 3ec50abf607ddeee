// Reverse chain from node sensitivities to beam and path parameters.
//
// The node set depends on the design through the path timing (Init.cpp:607-622),
// the interval schedule (Calc.cpp:144-207, Util.cpp:192-209) and the node
// formulas (Calc.cpp:177-190, Util.cpp:211-237). With every branch decision
// frozen at its current value, that map is smooth; it is differentiated here in
// forward mode with dense dual numbers (value + gradient over the 5 + 2*rows
// design parameters). Every value is computed with the same operations in the
// same order as quadrature.hpp, so the values reproduce the FP64 builder's node
// list exactly (tests compare them) and the derivatives are those of that map.
#include <cmath>
#include <vector>

#include "adjoint.hpp"
#include "batch.hpp"
#include "quadrature.hpp"

namespace thesis {
namespace {

struct Dual {
  double v = 0.0;
  std::vector<double> g;  // empty: zero gradient
  Dual() = default;
  Dual(double x) : v(x) {}  // NOLINT: constants promote implicitly
};

void axpy(std::vector<double>& y, double a, const std::vector<double>& x) {
  if (x.empty() || a == 0.0) return;
  if (y.empty()) y.assign(x.size(), 0.0);
  for (size_t i = 0; i < x.size(); ++i) y[i] += a * x[i];
}

Dual operator+(const Dual& a, const Dual& b) {
  Dual r(a.v + b.v);
  axpy(r.g, 1.0, a.g);
  axpy(r.g, 1.0, b.g);
  return r;
}
Dual operator-(const Dual& a, const Dual& b) {
  Dual r(a.v - b.v);
  axpy(r.g, 1.0, a.g);
  axpy(r.g, -1.0, b.g);
  return r;
}
Dual operator*(const Dual& a, const Dual& b) {
  Dual r(a.v * b.v);
  axpy(r.g, b.v, a.g);
  axpy(r.g, a.v, b.g);
  return r;
}
Dual operator/(const Dual& a, const Dual& b) {
  Dual r(a.v / b.v);
  axpy(r.g, 1.0 / b.v, a.g);
  axpy(r.g, -a.v / (b.v * b.v), b.g);
  return r;
}
Dual dsqrt(const Dual& a) {
  Dual r(std::sqrt(a.v));
  axpy(r.g, 0.5 / r.v, a.g);
  return r;
}
Dual dsquare_pow(const Dual& a) {  // std::pow(a, 2) as in the path timing
  Dual r(std::pow(a.v, 2));
  axpy(r.g, 2.0 * a.v, a.g);
  return r;
}

struct DSeg {
  int mode;
  Dual x, y, z, time;
  double qmod, param;
};

struct DNode {
  Dual xb, yb, zb, px, py, pz, dtau, qb;
};

struct DBeam {
  Dual ax, ax2, ay2, az2, nond_dt, az, beta;
};

Dual var(double v, int idx, int G, double scale = 1.0) {
  Dual d(v);
  d.g.assign(static_cast<size_t>(G), 0.0);
  d.g[static_cast<size_t>(idx)] = scale;
  return d;
}

void beam_at(const std::vector<DSeg>& P, int s, const Dual& tm, const RBox& rb, Dual& xb, Dual& yb, Dual& zb,
             double& qmod) {
  if (P[s].mode) {
    xb = P[s].x;
    yb = P[s].y;
    zb = P[s].z;
  } else {
    const Dual dx = P[s].x - P[s - 1].x;
    const Dual dy = P[s].y - P[s - 1].y;
    const Dual dz = P[s].z - P[s - 1].z;
    const Dual tcur = tm - P[s - 1].time;
    const Dual span = P[s].time - P[s - 1].time;
    xb = P[s - 1].x + (tcur / span) * dx;
    yb = P[s - 1].y + (tcur / span) * dy;
    zb = P[s - 1].z + (tcur / span) * dz;
  }
  const bool inside = !((xb.v > rb.xhi) || (xb.v < rb.xlo)) && !((yb.v > rb.yhi) || (yb.v < rb.ylo));
  qmod = inside ? P[s].qmod : 0.0;
}

Dual step_cap(const std::vector<DSeg>& P, int s, const Dual& elapsed, const DBeam& b) {
  Dual spp = elapsed / b.nond_dt;
  if (!(spp.v > 0.0)) spp = Dual(0.0);
  if (P[s].mode == 0) return ((Dual(0.58870501125) * b.ax) / Dual(P[s].param)) * dsqrt(Dual(12.0) * spp + Dual(1.0));
  Dual r = P[s].time - P[s - 1].time;
  return r.v == 0 ? Dual(1.0e-9) : r;
}

}  // namespace

Problem apply_design(const Problem& base, const Design& d) {
  if (base.beams.size() != 1) throw Error(kBadArgument, "the design interface needs a single-beam problem");
  if (base.compress || base.walls.any())
    throw Error(kBadArgument, "the design interface differentiates plain node lists: path compression and wall "
                              "images are off here");
  if (base.custom_points) throw Error(kBadArgument, "the design interface runs on the deck's grid, not on custom points");
  if (d.nrows() < 1) throw Error(kBadArgument, "the design needs at least one path row");
  Problem pb = base;
  BeamSpec& bs = pb.beams[0];
  bs.power = d.beam[0];
  bs.eff = d.beam[1];
  bs.q = d.beam[0] * d.beam[1] * 2.0;
  bs.ax = d.beam[2];
  bs.ay = d.beam[3];
  bs.az = d.beam[4];
  if (!(bs.ax > 0 && bs.ay > 0 && bs.az > 0 && d.beam[0] > 0 && d.beam[1] > 0))
    throw Error(kBadArgument, "design beam parameters must be positive");
  pb.paths[0] = make_path(d.rows.data(), d.nrows());
  pb.path_rows[0] = d.rows;
  pb.r_max = beam_radius_of_influence(pb, bs);
  finalize_problem(pb);
  return pb;
}

std::vector<double> chain_to_design(const Problem& base, const Design& d, double t_in, bool t_is_scan_end,
                                    const std::vector<double>& S, size_t* nodes_built) {
  const Problem pb = apply_design(base, d);
  const int G = d.nparams();
  const double alpha = pb.mat.alpha;
  // design variables
  const Dual P = var(d.beam[0], 0, G), eff = var(d.beam[1], 1, G), ax = var(d.beam[2], 2, G),
             ay = var(d.beam[3], 3, G), az = var(d.beam[4], 4, G);
  DBeam b;
  b.ax = ax;
  b.ax2 = ax * ax;
  b.ay2 = ay * ay;
  b.az2 = az * az;
  b.az = az;
  b.nond_dt = (ax * ax) / Dual(alpha);
  const Dual q = (P * eff) * Dual(2.0);
  b.beta = (Dual(std::pow(3.0 / 3.14159265358979323846, 1.5)) * q) / Dual(pb.mat.rho * pb.mat.cps);
  // path with timing (make_path)
  std::vector<DSeg> path;
  path.push_back(DSeg{1, Dual(0.0), Dual(0.0), Dual(0.0), Dual(0.0), 0.0, 0.0});
  for (int r = 0; r < d.nrows(); ++r) {
    const double* v = d.rows.data() + 6 * r;
    DSeg s;
    s.mode = static_cast<int>(v[0]);
    s.x = var(v[1] * 1e-3, 5 + 2 * r, G, 1e-3);
    s.y = var(v[2] * 1e-3, 6 + 2 * r, G, 1e-3);
    s.z = Dual(v[3] * 1e-3);
    s.qmod = v[4];
    s.param = v[5];
    const DSeg& p = path.back();
    if (s.mode) {
      s.time = p.time + Dual(s.param);
    } else {
      const Dual dx = s.x - p.x, dy = s.y - p.y, dz = s.z - p.z;
      s.time = p.time + dsqrt(dsquare_pow(dx) + dsquare_pow(dy) + dsquare_pow(dz)) / Dual(s.param);
    }
    path.push_back(s);
  }
  const int nseg = static_cast<int>(path.size());
  const Dual t = t_is_scan_end ? path.back().time : Dual(t_in);
  const RBox rb{pb.grid.xmin - pb.r_max, pb.grid.xmax + pb.r_max, pb.grid.ymin - pb.r_max, pb.grid.ymax + pb.r_max};
  const GLTable gl = gl_table();

  // locate the segment from the start (a fresh build), Calc.cpp:99-101
  int seg = 1;
  while ((t.v > path[seg].time.v) && (seg + 1 < nseg)) ++seg;
  while ((t.v < path[seg - 1].time.v) && (seg - 1 > 0)) --seg;

  std::vector<double> grad(static_cast<size_t>(G), 0.0);
  size_t node = 0;
  auto consume = [&](const DNode& nd) {
    if (node * 7 + 7 > S.size()) throw Error(kBadArgument, "sensitivity array shorter than the node list");
    const double* s = S.data() + 7 * node;
    ++node;
    if (!(nd.dtau.v > 0.0)) return;  // weight-0 node: no contribution to T
    axpy(grad, s[0], nd.xb.g);
    axpy(grad, s[1], nd.yb.g);
    axpy(grad, s[2], nd.zb.g);
    axpy(grad, s[3], nd.px.g);
    axpy(grad, s[4], nd.py.g);
    axpy(grad, s[5], nd.pz.g);
    axpy(grad, s[6] / nd.dtau.v, nd.dtau.g);
    axpy(grad, s[6] / nd.qb.v, nd.qb.g);
  };

  // the schedule of walk_intervals (quadrature.hpp), in duals
  const Dual start = b.nond_dt;
  Dual cap = start;
  int order = 16;
  const Dual end = path[nseg - 1].time;
  if (t.v <= end.v) {
    DNode n;
    double qm;
    beam_at(path, seg, t, rb, n.xb, n.yb, n.zb, qm);
    n.px = b.ax2 + Dual(0.0);
    n.py = b.ay2 + Dual(0.0);
    n.pz = b.az2 + Dual(0.0);
    n.qb = Dual(qm) * b.beta;
    n.dtau = Dual(0.0);
    if (n.qb.v > 0.0) consume(n);
  }
  Dual t2 = t, elapsed(0.0);
  if (t.v > end.v) {
    elapsed = elapsed + (t - end);
    t2 = end;
  }
  while (elapsed.v >= (Dual(2.0) * cap - start).v) {
    cap = cap * Dual(2.0);
    if (order != 2) order /= 2;
  }
  bool more = true;
  while (more) {
    bool next_seg = false;
    const Dual rt = step_cap(path, seg, elapsed, b);
    Dual t1 = t2 - ((rt.v < cap.v) ? rt : cap);
    const Dual seg_start = path[seg - 1].time;
    if (t1.v <= seg_start.v) {
      t1 = seg_start;
      if (seg_start.v > 0.0) next_seg = true;
      else more = false;
    }
    for (int a = order - 1; a >= 0; --a) {
      const int k = gl_offset(order) + a;
      const Dual tp = Dual(0.5) * ((t2 - t1) * Dual(gl.x[k]) + (t2 + t1));
      const Dual ct = Dual(12.0 * alpha) * (t - tp);
      DNode n;
      double qm;
      beam_at(path, seg, tp, rb, n.xb, n.yb, n.zb, qm);
      n.px = b.ax2 + ct;
      n.py = b.ay2 + ct;
      n.pz = b.az2 + ct;
      n.qb = Dual(qm) * b.beta;
      n.dtau = (Dual(0.5) * (t2 - t1)) * Dual(gl.w[k]);
      if (n.qb.v > 0.0 && n.dtau.v > 0.0) consume(n);
    }
    if (next_seg) --seg;
    elapsed = elapsed + (t2 - t1);
    t2 = t1;
    if (elapsed.v >= (Dual(2.0) * cap - start).v) {
      cap = cap * Dual(2.0);
      if (order != 2) order /= 2;
    }
  }
  if (nodes_built) *nodes_built = node;
  return grad;
}

}  // namespace thesis
