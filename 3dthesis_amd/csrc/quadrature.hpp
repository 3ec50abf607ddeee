// Adaptive backward-in-time Gauss-Legendre quadrature over a scan path.
//
// The heat-source time integral of the moving-Gaussian solution is replaced by
// a sum over "nodes": Gaussian blobs at the beam position of each quadrature
// time, with spreads phi = a^2 + 12 alpha (t - t'), weight dtau and prefactor
// qmod*beta. The interval schedule follows Calc::GaussIntegrate
// (Calc.cpp:67-211): start step nond_dt (x Depth_Z for solidification builds),
// order 16, step doubled and order halved as elapsed time grows, steps capped by
// Util::GetRefTime (Util.cpp:192-209) and cut at segment ends, integration back
// to t = 0 (Util::t0calc returns 0, Util.cpp:187).
//
// Written once as __host__ __device__ code so the host builder and the GPU's
// per-point builders produce the same node sets. The schedule walk and the
// per-node evaluation are separate so a wavefront can enumerate the intervals
// on one pass and spread the nodes over its lanes on a second.
#pragma once

#include <cmath>

#include "common.hpp"

namespace thesis {

// Per-beam constants the quadrature needs.
struct BeamQ {
  double ax2, ay2, az2;  // squared widths
  double ax;             // Width_X (GetRefTime)
  double nond_dt;        // ax^2 / alpha
  double az;             // Depth_Z (solidification refinement factor)
  double beta;           // (3/pi)^1.5 q / (rho c)
};

// x/y box outside which a beam position contributes nothing (Util::InRMax,
// Util.cpp:168-172).
struct RBox {
  double xlo, xhi, ylo, yhi;
};

struct PathSeg {  // device copy of a Segment
  double x, y, z, qmod, param, time;
  int mode, pad;
};

// 8-digit Gauss-Legendre abscissae and weights the reference uses
// (Calc.cpp:73-86), stored per order in ascending abscissa. Offsets: order 2 at
// 0, 4 at 2, 8 at 6, 16 at 14.
struct GLTable {
  double x[30];
  double w[30];
};
THD int gl_offset(int order) { return order - 2; }

THD constexpr GLTable gl_table() {
  return GLTable{
      {-0.57735027, 0.57735027,
       -0.86113631, -0.33998104, 0.33998104, 0.86113631,
       -0.96028986, -0.79666648, -0.52553241, -0.18343464, 0.18343464, 0.52553241, 0.79666648, 0.96028986,
       -0.98940093, -0.94457502, -0.8656312, -0.75540441, -0.61787624, -0.45801678, -0.28160355,
       -0.09501251, 0.09501251, 0.28160355, 0.45801678, 0.61787624, 0.75540441, 0.8656312, 0.94457502,
       0.98940093},
      {1.0, 1.0,
       0.34785485, 0.65214515, 0.65214515, 0.34785485,
       0.10122854, 0.22238103, 0.31370665, 0.36268378, 0.36268378, 0.31370665, 0.22238103, 0.10122854,
       0.02715246, 0.06225352, 0.09515851, 0.12462897, 0.14959599, 0.16915652, 0.18260342, 0.18945061,
       0.18945061, 0.18260342, 0.16915652, 0.14959599, 0.12462897, 0.09515851, 0.06225352, 0.02715246}};
}

// Beam centre and effective qmod at time tm inside segment s (Util::GetBeamLoc).
THD void beam_at(const PathSeg* P, int s, double tm, const RBox& rb, double& xb, double& yb, double& zb,
                 double& qmod) {
  if (P[s].mode) {
    xb = P[s].x; yb = P[s].y; zb = P[s].z;
  } else {
    const double dx = P[s].x - P[s - 1].x;
    const double dy = P[s].y - P[s - 1].y;
    const double dz = P[s].z - P[s - 1].z;
    const double tcur = tm - P[s - 1].time;
    const double span = P[s].time - P[s - 1].time;
    xb = P[s - 1].x + (tcur / span) * dx;
    yb = P[s - 1].y + (tcur / span) * dy;
    zb = P[s - 1].z + (tcur / span) * dz;
  }
  const bool inside = !((xb > rb.xhi) || (xb < rb.xlo)) && !((yb > rb.yhi) || (yb < rb.ylo));
  qmod = inside ? P[s].qmod : 0.0;
}

// Largest allowed step inside segment s after `elapsed` seconds (GetRefTime).
THD double step_cap(const PathSeg* P, int s, double elapsed, const BeamQ& b) {
  const double spp = (elapsed / b.nond_dt) > 0.0 ? (elapsed / b.nond_dt) : 0.0;
  if (P[s].mode == 0) {
    return 0.58870501125 * b.ax / P[s].param * sqrt(12.0 * spp + 1.0);
  }
  double r = P[s].time - P[s - 1].time;
  return r == 0 ? 1.0e-9 : r;
}

// Segment holding time t, searched from `hint` the way Calc.cpp:99-101 does
// (up while t is past the segment end, then down while before its start).
THD int locate_segment(const PathSeg* P, int nseg, double t, int hint) {
  int s = hint;
  while ((t > P[s].time) && (s + 1 < nseg)) ++s;
  while ((t < P[s - 1].time) && (s - 1 > 0)) --s;
  return s;
}

// Walk the interval schedule for one beam. Calls
//   fn.instant(seg)                 if the source is on at time t (node of weight 0)
//   fn.interval(t1, t2, seg, order) for every quadrature interval, newest first.
// A non-finite t (a degenerate liquidus interpolation) yields no nodes; the
// reference's loop would not terminate there.
template <class Fn>
THD void walk_intervals(const PathSeg* P, int nseg, const BeamQ& b, double t, bool is_sol, int seg, Fn& fn) {
  if (!(t - t == 0.0)) return;
  const double t_stop = 0.0;
  const double start = is_sol ? b.nond_dt * b.az : b.nond_dt;
  double cap = start;
  int order = 16;
  const double end = P[nseg - 1].time;
  if (t <= end) fn.instant(seg);
  double t2 = t, elapsed = 0.0;
  if (t > end) {
    elapsed += t - end;
    t2 = end;
  }
  while (elapsed >= 2 * cap - start) {
    cap *= 2.0;
    if (order != 2) order /= 2;
  }
  bool more = true;
  while (more) {
    bool next_seg = false;
    const double rt = step_cap(P, seg, elapsed, b);
    double t1 = t2 - ((rt < cap) ? rt : cap);
    const double seg_start = P[seg - 1].time;
    if (t1 <= seg_start) {
      t1 = seg_start;
      if (seg_start > t_stop) next_seg = true;
      else more = false;
    }
    fn.interval(t1, t2, seg, order);
    if (next_seg) --seg;
    elapsed += (t2 - t1);
    t2 = t1;
    if (elapsed >= 2 * cap - start) {
      cap *= 2.0;
      if (order != 2) order /= 2;
    }
  }
}

THD bool in_rbox(const RBox& rb, double x, double y) {
  return !((x > rb.xhi) || (x < rb.xlo)) && !((y > rb.yhi) || (y < rb.ylo));
}

// Interval schedule of Calc::GaussCompressIntegrate (Calc.cpp:213-500). Where
// the path doubles back within the diffusion radius r2 of the beam position at
// the interval's newer end t2, the segments of one quadrature interval are
// merged into a single source at their power-weighted centre. Calls
//   fn.instant(seg)                          as walk_intervals does,
//   fn.interval(t1, t2, seg, order)          for an interval on one segment,
//   fn.merged(t1, t2, order, cx, cy, qavg)   for a merged interval: nodes at
//                                            (cx, cy, 0) whose qmod starts at
//                                            qavg and is multiplied by beta
//                                            before every node (Calc.cpp:463).
// The R_max skip of Calc.cpp:290-302 ends the walk at a vertex outside the box
// instead of looping on it; the input layer refuses paths that reach one.
template <class Fn>
THD void walk_compressed(const PathSeg* P, int nseg, const BeamQ& b, const RBox& rb, double alpha, double t,
                         bool is_sol, int seg, Fn& fn) {
  if (!(t - t == 0.0)) return;
  const double t_stop = 0.0;
  const double start = is_sol ? b.nond_dt * b.az : b.nond_dt;
  double cap = start;
  int order = 16;
  const double end = P[nseg - 1].time;
  if (t <= end) fn.instant(seg);
  double t2 = t, elapsed = 0.0;
  if (t > end) {
    elapsed += t - end;
    t2 = end;
  }
  while (elapsed >= 2 * cap - start) {
    cap *= 2.0;
    if (order != 2) order /= 2;
  }
  bool more = true;
  while (more) {
    if (P[seg].time <= t_stop || !in_rbox(rb, P[seg].x, P[seg].y)) break;
    while (elapsed >= 2 * cap - start) {
      cap *= 2.0;
      if (order != 2) order /= 2;
    }
    const double rt = step_cap(P, seg, elapsed, b);
    const double use = (rt < cap) ? rt : cap;
    double xp, yp, zq, qq;
    beam_at(P, seg, t2, rb, xp, yp, zq, qq);
    // log(2) / 8 * a^2 * (12 alpha (t - t2) / a^2 + 1)
    const double r2 =
        0.69314718055994530942 / 8.0 * (b.ax * b.ax) * (12.0 * (t - t2) * alpha / (b.ax * b.ax) + 1.0);
    int seg2 = seg;
    double t1 = P[seg - 1].time;
    int merged = 0;
    double sum_t = 0, sum_qt = 0, sum_qtx = 0, sum_qty = 0;
    if (t1 < t2 - use) {
      t1 = t2 - use;
    } else {
      while (true) {
        if (P[seg2 - 1].time <= t_stop) {
          more = false;
          break;
        }
        const PathSeg& A = P[seg2 - 1];
        const PathSeg& B = P[seg2];
        const double xs = A.x, ys = A.y, ts = A.time;
        const double dist2 = (xs - xp) * (xs - xp) + (ys - yp) * (ys - yp);
        if (!in_rbox(rb, xp, yp)) break;
        bool last = false;
        if ((dist2 > r2) || (B.qmod != A.qmod && (order > 2 || (t2 - ts) > (cap / 64.0)))) {
          // the segment reaching out of the radius closes the interval
          if (B.mode) {
            sum_qtx += B.x * B.qmod * (t1 - ts);
            sum_qty += B.y * B.qmod * (t1 - ts);
            sum_qt += B.qmod * (t1 - ts);
            sum_t += (t1 - ts);
            t1 = ts;
          } else {
            const double dx = B.x - A.x, dy = B.y - A.y, dt = B.time - ts;
            double t_int = ts + dt * (sqrt(((xp - xs) * (xp - xs) + (yp - ys) * (yp - ys)) / (dx * dx + dy * dy)) -
                                      sqrt(r2 / (dx * dx + dy * dy)));
            if (t_int < t2 - cap) t_int = t2 - cap;
            sum_qtx += (xs + dx * ((t1 + t_int) / 2.0 - ts) / dt) * B.qmod * (t1 - t_int);
            sum_qty += (ys + dy * ((t1 + t_int) / 2.0 - ts) / dt) * B.qmod * (t1 - t_int);
            sum_qt += B.qmod * (t1 - t_int);
            sum_t += (t1 - t_int);
            t1 = t_int;
          }
          last = true;
        } else {
          if (ts < t2 - cap) {
            last = true;
            t1 = t2 - cap;
          } else {
            t1 = ts;
            ++merged;
          }
          const double dt = B.time - t1;
          if (B.mode) {
            sum_qtx += B.x * B.qmod * dt;
            sum_qty += B.y * B.qmod * dt;
          } else {
            const double dx = B.x - A.x, dy = B.y - A.y;
            sum_qtx += (xs + dx * ((B.time + t1) / 2.0 - ts) / dt) * B.qmod * dt;
            sum_qty += (ys + dy * ((B.time + t1) / 2.0 - ts) / dt) * B.qmod * dt;
          }
          sum_qt += B.qmod * dt;
          sum_t += dt;
        }
        if (t1 == P[seg2 - 1].time) {
          --seg2;
          if (!in_rbox(rb, xs, ys)) break;
        }
        if (last) break;
      }
      if (!merged) {
        t1 = P[seg - 1].time;
        seg2 = seg - 1;
      }
    }
    if (merged) {
      if (sum_qt > 0.0) fn.merged(t1, t2, order, sum_qtx / sum_qt, sum_qty / sum_qt, sum_qt / sum_t);
    } else {
      fn.interval(t1, t2, seg, order);
    }
    elapsed += t2 - t1;
    t2 = t1;
    if (elapsed >= 2 * cap - start) {
      cap *= 2.0;
      if (order != 2) order /= 2;
    }
    seg = seg2;
    if (t1 <= t_stop) more = false;
  }
}

// Raw node: centre, spreads (not inverted), weight, qmod*beta.
struct RawNode {
  double xb, yb, zb, px, py, pz, dtau, qb;
};

// Node `a` (ascending abscissa index) of interval [t1, t2]. Returns whether the
// reference would keep it (qmod > 0 and dtau > 0).
THD bool interval_node(const PathSeg* P, int seg, const BeamQ& b, const RBox& rb, double alpha, double t,
                       double t1, double t2, int order, int a, const GLTable& gl, RawNode& n) {
  const int k = gl_offset(order) + a;
  const double tp = 0.5 * ((t2 - t1) * gl.x[k] + (t2 + t1));
  const double ct = 12.0 * alpha * (t - tp);
  double q;
  beam_at(P, seg, tp, rb, n.xb, n.yb, n.zb, q);
  n.px = b.ax2 + ct;
  n.py = b.ay2 + ct;
  n.pz = b.az2 + ct;
  n.qb = q * b.beta;
  n.dtau = 0.5 * (t2 - t1) * gl.w[k];
  return n.qb > 0.0 && n.dtau > 0.0;
}

// Node `a` of a merged interval; q is its qmod (the interval's average times
// beta^(k+1) for the k-th node emitted).
THD bool merged_node(const BeamQ& b, double alpha, double t, double t1, double t2, int order, int a, double cx,
                     double cy, double q, const GLTable& gl, RawNode& n) {
  const int k = gl_offset(order) + a;
  const double tp = 0.5 * ((t2 - t1) * gl.x[k] + (t2 + t1));
  const double ct = 12.0 * alpha * (t - tp);
  n.xb = cx;
  n.yb = cy;
  n.zb = 0.0;
  n.px = b.ax2 + ct;
  n.py = b.ay2 + ct;
  n.pz = b.az2 + ct;
  n.qb = q;
  n.dtau = 0.5 * (t2 - t1) * gl.w[k];
  return n.qb > 0.0 && n.dtau > 0.0;
}

// Instantaneous node at t (weight 0). Returns whether the source is on.
THD bool instant_node(const PathSeg* P, int seg, const BeamQ& b, const RBox& rb, double t, RawNode& n) {
  double q;
  beam_at(P, seg, t, rb, n.xb, n.yb, n.zb, q);
  n.px = b.ax2 + 0.0;
  n.py = b.ay2 + 0.0;
  n.pz = b.az2 + 0.0;
  n.qb = q * b.beta;
  n.dtau = 0.0;
  return n.qb > 0.0;
}

// Log prefactor stored with a node (Util::AddToNodes, Util.cpp:40-50).
THD double node_expmod(const RawNode& n) { return log(n.qb) - log(n.px * n.py * n.pz) / 2.0; }

// Emit every kept node of one beam in the reference's order.
template <class Emit>
struct EmitAll {
  const PathSeg* P;
  const BeamQ& b;
  const RBox& rb;
  double alpha, t;
  const GLTable& gl;
  Emit& emit;
  THD void instant(int seg) {
    RawNode n;
    if (instant_node(P, seg, b, rb, t, n)) emit(n);
  }
  THD void interval(double t1, double t2, int seg, int order) {
    for (int a = order - 1; a >= 0; --a) {
      RawNode n;
      if (interval_node(P, seg, b, rb, alpha, t, t1, t2, order, a, gl, n)) emit(n);
    }
  }
  THD void merged(double t1, double t2, int order, double cx, double cy, double qavg) {
    double q = qavg;
    for (int a = order - 1; a >= 0; --a) {
      q *= b.beta;
      RawNode n;
      if (merged_node(b, alpha, t, t1, t2, order, a, cx, cy, q, gl, n)) emit(n);
    }
  }
};

template <class Emit>
THD void beam_nodes(const PathSeg* P, int nseg, const BeamQ& b, const RBox& rb, double alpha, double t, bool is_sol,
                    int seg, const GLTable& gl, Emit& emit, bool compressed = false) {
  EmitAll<Emit> e{P, b, rb, alpha, t, gl, emit};
  if (compressed) walk_compressed(P, nseg, b, rb, alpha, t, is_sol, seg, e);
  else walk_intervals(P, nseg, b, t, is_sol, seg, e);
}

// One mirror image of the node set (Calc::AddBCs, Calc.cpp:502-590): a node
// centre x maps to ox + sx * x with sx = +-1, per axis.
struct WallImage {
  double ox, sx, oy, sy, oz, sz;
};

}  // namespace thesis
