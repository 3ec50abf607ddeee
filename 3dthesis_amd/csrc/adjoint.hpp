// Adjoint sensitivities of a snapshot T field to its node set (adjoint.hip) and
// their chain through the quadrature schedule to beam and path parameters
// (design.cpp).
#pragma once

#include <vector>

#include "device.hpp"
#include "model.hpp"

namespace thesis {

// S (n x 7, device): per node dL/dxb, dL/dyb, dL/dzb, dL/dphix, dL/dphiy,
// dL/dphiz, dL/d ln(amplitude) for w = dL/dT over the grid. partial must hold
// ceil(pnum/chunk) * n * 7 doubles.
void launch_adjoint(const NodeF64* nodes, int n, const GridGeom& g, int pnum, const double* w, int chunk,
                    double* partial, double* S, hipStream_t s);

// A single-beam design: the beam parameters and Path-file rows that the
// gradient is taken with respect to.
struct Design {
  double beam[5];            // Power, Efficiency, Width_X, Width_Y, Depth_Z
  std::vector<double> rows;  // Path-file rows, 6 per row (mode, x mm, y mm, z mm, Pmod, speed|time)
  int nrows() const { return static_cast<int>(rows.size() / 6); }
  int nparams() const { return 5 + 2 * nrows(); }
};

// The problem with its (single) beam and path replaced by the design.
Problem apply_design(const Problem& base, const Design& d);

// Gradient of L with respect to the design (5 beam parameters, then x and y in
// Path-file units (mm) of every row) from the node sensitivities S (n x 7, host)
// of the node list built at time t (t_is_scan_end: t follows the path's end).
// Also returns the node count it built, which must equal S's.
std::vector<double> chain_to_design(const Problem& base, const Design& d, double t, bool t_is_scan_end,
                                    const std::vector<double>& S, size_t* nodes_built);

}  // namespace thesis
