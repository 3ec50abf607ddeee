// Host-side node builder (the reference's Calc::Integrate_Serial,
// Calc.cpp:48-65, over Calc::GaussIntegrate) and the device copies of the
// path/beam data the GPU-side builders read.
#pragma once

#include <vector>

#include "device.hpp"
#include "model.hpp"
#include "quadrature.hpp"

namespace thesis {

struct PathData {
  std::vector<PathSeg> segs;  // all beams' segments back to back
  std::vector<int> off;       // beam b occupies [off[b], off[b+1])
  std::vector<BeamQ> beams;
  RBox rbox{};
  double alpha = 0.0;
  bool compress = false;          // compressed quadrature (Calc.cpp:49-58)
  std::vector<WallImage> images;  // wall images, empty without walls (Calc.cpp:62)
};

PathData make_path_data(const Problem& pb);

// Image list of the wall planes (Calc::AddBCs, Calc.cpp:502-570): integer
// reflection coordinates found breadth-first over `Reflections` rounds, each
// turned into a per-axis map x -> (r + n) hi - (r - n) lo + (1 - 2n) x,
// n = |r mod 2|. The first image is the identity. Empty without walls.
std::vector<WallImage> wall_images(const Problem& pb);

class Quadrature {
 public:
  explicit Quadrature(const Problem& pb);

  // Nodes of every beam at time t, appended to `out` in the reference's order
  // (Calc::Integrate_Serial, Calc.cpp:48-65): with compression, the compressed
  // list, or the plain one when that is no longer (main builds only); then the
  // wall images. Main builds (is_sol false) advance the per-beam segment hints
  // the way the reference's static start segments do (Calc.cpp:70, :104, :216).
  void build(double t, bool is_sol, std::vector<NodeF64>& out);
  // Same, starting from the given hints and leaving them untouched.
  void build_from(double t, bool is_sol, const std::vector<int>& hints, std::vector<NodeF64>& out) const;

  const PathData& data() const { return pd_; }
  std::vector<int>& hints() { return hint_; }
  void reset_hints();

 private:
  void build_at(double t, bool is_sol, const std::vector<int>& segs, std::vector<NodeF64>& out) const;
  PathData pd_;
  GLTable gl_;
  std::vector<int> hint_;
};

NodeF64 to_f64(const RawNode& r);

}  // namespace thesis
