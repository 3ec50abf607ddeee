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
};

template <class Emit>
THD void beam_nodes(const PathSeg* P, int nseg, const BeamQ& b, const RBox& rb, double alpha, double t, bool is_sol,
                    int seg, const GLTable& gl, Emit& emit) {
  EmitAll<Emit> e{P, b, rb, alpha, t, gl, emit};
  walk_intervals(P, nseg, b, t, is_sol, seg, e);
}

}  // namespace thesis
