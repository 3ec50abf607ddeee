// Problem description: the parsed Mode/Material/Beam/Path/Domain/Output/Settings
// inputs of a 3DThesis run (the reference keeps the same facts in Simdat,
// DataStructs.h:63-183). Plain data, filled by input.cpp or by the C-ABI's
// descriptor entry point.
#pragma once

#include <cfloat>
#include <climits>
#include <string>
#include <vector>

namespace thesis {

// One scan-path segment (reference path_seg, DataStructs.h:42-48).
struct Segment {
  int mode;        // 0 = line (param is speed), 1 = spot (param is dwell time)
  double x, y, z;  // end point, metres
  double qmod;     // power modulation
  double param;    // speed or dwell
  double time;     // cumulative end time (Init.cpp:607-622)
};

struct BeamSpec {
  double ax, ay, az;  // Width_X, Width_Y, Depth_Z
  double power;       // Power as read
  double eff;         // Efficiency
  double q;           // 2 * Power * Efficiency (half-space image, Init.cpp:862-865)
  double nond_dt;     // ax^2 / alpha (Util.cpp:138-143)
};

struct MaterialSpec {
  double T_init, T_liq, kon, cps, rho;
  double alpha;                      // kon / (rho cps)
  double cet_N0, cet_n, cet_a;       // Gaumann CET constants (DBL_MAX when unset)
};

enum class Mode { Snapshots, Solidification };
enum class Tracking { None, Volume, Surface, Stork };

struct GridSpec {
  int xnum = 0, ynum = 0, znum = 0;
  long long pnum = 0;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double xres, yres, zres;
  // coordinate step actually used for point positions: (max-min)/(num-1), or 0
  // when num == 1 (the coordinate is then max) -- Grid.h:310-327
  double sx, sy, sz;
};

struct OutputSpec {
  bool x = true, y = true, z = true, T = true, T_hist = false;
  bool tSol = false, G = false, Gx = false, Gy = false, Gz = false, V = false, dTdt = false,
       eqFrac = false, depth = false, numMelt = false, RDF = false, mp_stats = false;
  bool H = false, Hx = false, Hy = false, Hz = false;
};

// Insulated wall planes for the method of images (Domain BoundaryConditions,
// Init.cpp:679-694): DBL_MAX marks an unset plane, INT_MAX unset reflections.
struct WallSpec {
  double xmin = DBL_MAX, xmax = DBL_MAX, ymin = DBL_MAX, ymax = DBL_MAX, zmin = DBL_MAX;
  int reflections = INT_MAX;
  bool any() const { return xmin != DBL_MAX || xmax != DBL_MAX || ymin != DBL_MAX || ymax != DBL_MAX || zmin != DBL_MAX; }
};

// abs() semantics the reference gets from its compiler at Grid.cpp:381/:535:
// g++ binds int abs(int) (truncating), clang a floating overload.
enum class AbsMode { IntTrunc = 0, Fabs = 1 };

struct Problem {
  std::string name = "N/A";
  std::string data_dir = "Data";
  Mode mode = Mode::Snapshots;
  Tracking tracking = Tracking::None;
  std::vector<double> snapshot_times;
  double dt = 1e-5;
  int out_freq = INT_MAX;
  double radius_check = 1.0;
  bool secondary = false;

  MaterialSpec mat{};
  std::vector<BeamSpec> beams;
  std::vector<std::vector<Segment>> paths;
  std::vector<std::vector<double>> path_rows;  // Path-file rows as read (6 numbers each, file units)
  double scan_end = 0.0;  // Util::Calc_AllScansEndTime

  GridSpec grid{};
  // Domain Custom File (Init.cpp:816-857): explicit points (x, y, z in m, one
  // triple per point) replace the structured grid; pnum is their count.
  bool custom_points = false;
  std::vector<double> points;
  WallSpec walls{};
  OutputSpec out{};

  // settings
  double sol_tol = 1e-3;
  int sol_iter = 10;
  double t_hist_inv = 1e9;  // 1 / Cutoff_Peak after Util::Calc_RMax
  double p_hist = 1e-2;
  double r_max = -1.0;
  double buffer = -1.0;     // the Buffer setting as read (negative: derive r_max per beam)
  int compress = 0;          // Path Compression: merged quadrature intervals (Calc.cpp:213-500)
  int threads = 1;
  AbsMode abs_mode = AbsMode::IntTrunc;

  bool do_sol() const {  // any solidification column requested (Grid.h:176-241)
    return out.tSol || out.G || out.V || out.Gx || out.Gy || out.Gz || out.dTdt || out.eqFrac || out.RDF;
  }
};

// Parse a ParamInput file and every file it names. Throws thesis::Error.
Problem load_problem(const std::string& paraminput);

// Validate scope and derived quantities of a Problem built programmatically.
void finalize_problem(Problem& pb);

// Util::Calc_RMax for one beam under the problem's settings (Util.cpp:145-166).
double beam_radius_of_influence(const Problem& pb, const BeamSpec& b);

}  // namespace thesis
