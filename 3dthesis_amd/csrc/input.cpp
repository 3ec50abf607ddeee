// Input surface: the keyword-brace text files of 3DThesis.
//
// Reads ParamInput and the Mode, Material, Beam, Path, Domain, Output and
// Settings files it names, with the reference's defaults, wildcard rule and
// derived quantities (Init.cpp:33-908, Util.cpp:76-166). The file format and
// defaults have to match the reference exactly for the library to be a
// drop-in; the code below is this project's own reader.
#include <sys/stat.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "common.hpp"
#include "batch.hpp"
#include "model.hpp"

namespace thesis {
namespace {

// Cursor over a whole file with the stream semantics the reference relies on:
// a word read skips whitespace, a line skip discards through '\n', and any
// failed read (including one attempted after end-of-file) ends parsing.
class Cursor {
 public:
  explicit Cursor(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { failed_ = true; return; }
    std::ostringstream ss;
    ss << f.rdbuf();
    buf_ = ss.str();
    opened_ = true;
  }
  bool opened() const { return opened_; }

  bool word(std::string& w) {
    if (!guard()) return false;
    while (pos_ < buf_.size() && ws(buf_[pos_])) ++pos_;
    if (pos_ >= buf_.size()) { eof_ = failed_ = true; return false; }
    size_t s = pos_;
    while (pos_ < buf_.size() && !ws(buf_[pos_])) ++pos_;
    w.assign(buf_, s, pos_ - s);
    if (pos_ >= buf_.size()) eof_ = true;
    return true;
  }

  bool skip_line() {
    if (!guard()) return false;
    size_t n = 0;
    while (pos_ < buf_.size()) {
      ++n;
      if (buf_[pos_++] == '\n') return true;
    }
    eof_ = true;
    if (n == 0) { failed_ = true; return false; }
    return true;
  }

  bool number(double& v, bool integral) {
    if (!guard()) return false;
    while (pos_ < buf_.size() && ws(buf_[pos_])) ++pos_;
    if (pos_ >= buf_.size()) { eof_ = failed_ = true; return false; }
    const char* s = buf_.c_str() + pos_;
    char* e = nullptr;
    v = integral ? static_cast<double>(std::strtol(s, &e, 10)) : std::strtod(s, &e);
    if (e == s) { failed_ = true; return false; }
    pos_ += static_cast<size_t>(e - s);
    if (pos_ >= buf_.size()) eof_ = true;
    return true;
  }

 private:
  static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
  bool guard() {
    if (failed_) return false;
    if (eof_) { failed_ = true; return false; }
    return true;
  }
  std::string buf_;
  size_t pos_ = 0;
  bool eof_ = false, failed_ = false, opened_ = false;
};

// "Block { key value ... }" reader (Init.cpp:33-131). Values stay empty strings
// when absent. With count_only, only top-level block names are counted.
struct Blocks {
  std::map<std::string, std::map<std::string, std::string>> val;
  std::map<std::string, int> count;

  Blocks(const std::string& path, const std::map<std::string, std::vector<std::string>>& schema,
         bool count_only = false) {
    for (auto& kv : schema)
      for (auto& k : kv.second) val[kv.first][k] = "";
    Cursor c(path);
    if (!c.opened()) return;
    std::string w, v;
    while (true) {
      if (!c.word(w) || !c.skip_line()) return;
      auto it = schema.find(w);
      if (it == schema.end()) continue;
      const std::string block = w;
      count[block] += 1;
      if (!c.word(w) || !c.skip_line()) return;
      if (w != "{") continue;
      while (true) {
        if (!c.word(w)) return;
        if (w == "}") {
          if (!c.skip_line()) return;
          break;
        }
        if (!c.word(v) || !c.skip_line()) return;
        if (count_only) continue;
        for (auto& key : it->second)
          if (key == w) val[block][key] = v;
      }
    }
  }
  const std::string& get(const char* b, const char* k) { return val[b][k]; }
};

double as_double(const std::string& s, double def, const char* what) {
  if (s.empty()) return def;
  char* e = nullptr;
  double v = std::strtod(s.c_str(), &e);
  if (e == s.c_str()) throw Error(kBadInput, std::string("cannot read a number for ") + what + ": " + s);
  return v;
}
int as_int(const std::string& s, int def, const char* what) {
  if (s.empty()) return def;
  char* e = nullptr;
  long v = std::strtol(s.c_str(), &e, 10);
  if (e == s.c_str()) throw Error(kBadInput, std::string("cannot read an integer for ") + what + ": " + s);
  return static_cast<int>(v);
}
bool as_bool(const std::string& s, bool def, const char* what) { return as_int(s, def ? 1 : 0, what) != 0; }

bool exists(const std::string& p) {
  std::ifstream f(p);
  return static_cast<bool>(f);
}

// Beam/Path '*' wildcard: one file, or <before>1<after>, <before>2<after>, ...
// while they exist (Init.cpp:445-487, :517-559).
std::vector<std::string> expand(const std::string& name) {
  size_t star = name.find('*');
  if (star == std::string::npos) return {name};
  if (name.find('*', star + 1) != std::string::npos)
    throw Error(kBadInput, "Too many '*' in: " + name);
  std::vector<std::string> out;
  for (int n = 1;; ++n) {
    std::string f = name.substr(0, star) + std::to_string(n) + name.substr(star + 1);
    if (!exists(f)) break;
    out.push_back(f);
  }
  return out;
}

BeamSpec read_beam(const std::string& file) {  // Init.cpp:488-515
  Blocks b(file, {{"Shape", {"Width_X", "Width_Y", "Depth_Z"}}, {"Intensity", {"Power", "Efficiency"}}});
  BeamSpec s{};
  s.ax = as_double(b.get("Shape", "Width_X"), 10.0e-6, "Width_X");
  s.ay = as_double(b.get("Shape", "Width_Y"), 10.0e-6, "Width_Y");
  s.az = as_double(b.get("Shape", "Depth_Z"), 1.0e-6, "Depth_Z");
  s.power = as_double(b.get("Intensity", "Power"), 1200.0, "Power");
  s.eff = as_double(b.get("Intensity", "Efficiency"), 1.0, "Efficiency");
  if (s.eff == DBL_MAX) s.eff = 1.0;
  s.q = s.power * s.eff * 2.0;
  return s;
}

std::vector<Segment> read_path(const std::string& file, std::vector<double>& rows) {  // Init.cpp:560-625
  Cursor c(file);
  if (!c.opened()) throw Error(kBadInput, "cannot open path file: " + file);
  rows.clear();
  while (c.skip_line()) {
    double v[6];
    bool ok = true;
    for (int i = 0; i < 6 && ok; ++i) ok = c.number(v[i], i == 0);
    if (!ok) break;
    rows.insert(rows.end(), v, v + 6);
  }
  if (rows.empty()) throw Error(kBadInput, "path file has no segments: " + file);
  return make_path(rows.data(), static_cast<int>(rows.size() / 6));
}

// Init::FileRead_Points (Init.cpp:816-857): the first line is skipped, then
// x y z triples in mm (divided by 1000) until extraction fails.
std::vector<double> read_points(const std::string& file) {
  Cursor c(file);
  std::vector<double> pts;
  if (!c.opened()) return pts;
  while (c.skip_line()) {
    double v[3];
    bool ok = true;
    for (int i = 0; i < 3 && ok; ++i) ok = c.number(v[i], false);
    if (!ok) break;
    for (double x : v) pts.push_back(x / 1000.0);
  }
  return pts;
}

Tracking tracking_of(const std::string& s) {
  if (s.empty() || s == "None") return Tracking::None;
  if (s == "Volume") return Tracking::Volume;
  if (s == "Surface") return Tracking::Surface;
  if (s == "Stork") return Tracking::Stork;
  throw Error(kBadInput, "tracking mode '" + s + "' is outside this library's scope (None, Volume, Surface, Stork)");
}

std::vector<double> comma_list(const std::string& s) {
  std::vector<double> out;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) out.push_back(as_double(tok.empty() ? "x" : tok, 0.0, "time list"));
  return out;
}

// Util::Calc_ScanBounds (Util.cpp:82-136): fill unset domain bounds from the
// powered part of the paths (+-500 um in x/y, -250 um in z).
void scan_bounds(Problem& pb) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double prev[3] = {0, 0, 0}, prev_q = 0.0;
  bool first = true;
  auto take = [&](const double* v) {
    for (int a = 0; a < 3; ++a) {
      if (v[a] < lo[a]) lo[a] = v[a];
      if (v[a] > hi[a]) hi[a] = v[a];
    }
  };
  for (auto& path : pb.paths)
    for (auto& g : path) {
      const double v[3] = {g.x, g.y, g.z};
      if (g.qmod > 0.0) {
        if (first) {
          for (int a = 0; a < 3; ++a) lo[a] = hi[a] = v[a];
        } else {
          take(v);
          if (prev_q == 0.0 && g.mode == 0) take(prev);
        }
      }
      first = false;
      prev[0] = g.x; prev[1] = g.y; prev[2] = g.z; prev_q = g.qmod;
    }
  GridSpec& d = pb.grid;
  if (d.xmin == -DBL_MAX) d.xmin = lo[0] - 500e-6;
  if (d.xmax == DBL_MAX) d.xmax = hi[0] + 500e-6;
  if (d.ymin == -DBL_MAX) d.ymin = lo[1] - 500e-6;
  if (d.ymax == DBL_MAX) d.ymax = hi[1] + 500e-6;
  if (d.zmin == -DBL_MAX) d.zmin = lo[2] - 250e-6;
  if (d.zmax == DBL_MAX) d.zmax = hi[2];
}

// Init::SetDomainParams (Init.cpp:866-908), including its use of ynum when
// deciding the z resolution.
void size_domain(GridSpec& d) {
  auto axis = [](int& num, double lo, double& hi, double& res, bool single) {
    if (num == INT_MAX) {
      num = 1 + static_cast<int>(0.5 + (hi - lo) / res);
      hi = lo + (num - 1) * res;
    } else if (!single) {
      res = (hi - lo) / (num - 1);
    } else {
      res = DBL_MAX;
    }
  };
  axis(d.xnum, d.xmin, d.xmax, d.xres, d.xnum == 1);
  axis(d.ynum, d.ymin, d.ymax, d.yres, d.ynum == 1);
  axis(d.znum, d.zmin, d.zmax, d.zres, d.ynum == 1);
  d.pnum = static_cast<long long>(d.xnum) * d.ynum * d.znum;
}

// Util::Calc_RMax (Util.cpp:145-166): the last beam's radius wins.
void radius_of_influence(Problem& pb) {
  pb.t_hist_inv = 1.0 / pb.t_hist_inv;
  pb.buffer = pb.r_max;
  if (pb.r_max < 0.0) {
    for (const BeamSpec& b : pb.beams) {
      if (pb.t_hist_inv < std::exp(3.0 / 2.0)) pb.r_max = b.ax * std::sqrt(std::log(pb.t_hist_inv) / 3.0);
      else pb.r_max = b.ax * std::pow(pb.t_hist_inv, (1.0 / 3.0)) / std::sqrt(2.0 * std::exp(1.0));
      const double beta = std::pow(3.0 / 3.14159, 1.5) * b.q / (pb.mat.rho * pb.mat.cps);
      const double x = (pb.mat.T_liq - pb.mat.T_init) * pb.p_hist;
      const double ratio = beta / (x * b.ax * b.ax * b.ax);
      double r2;
      if (ratio < std::exp(3.0 / 2.0)) r2 = b.ax * std::sqrt(std::log(ratio) / 3.0);
      else r2 = std::pow(beta / x, (1.0 / 3.0)) / std::sqrt(2.0 * std::exp(1.0));
      pb.r_max = pb.r_max > r2 ? pb.r_max : r2;
    }
  }
}

}  // namespace

std::vector<Segment> make_path(const double* rows6, int nrows) {
  std::vector<Segment> path;
  path.reserve(static_cast<size_t>(nrows) + 1);
  path.push_back(Segment{1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0});  // implicit start at the origin (Init.cpp:572-579)
  for (int r = 0; r < nrows; ++r) {
    const double* v = rows6 + 6 * r;
    path.push_back(Segment{static_cast<int>(v[0]), v[1] * 1e-3, v[2] * 1e-3, v[3] * 1e-3, v[4], v[5], 0.0});
  }
  path[0].time = 0.0;
  for (size_t s = 1; s < path.size(); ++s) {
    Segment& g = path[s];
    const Segment& p = path[s - 1];
    if (g.mode) {
      g.time = p.time + g.param;
    } else {
      const double dx = g.x - p.x, dy = g.y - p.y, dz = g.z - p.z;
      g.time = p.time + std::sqrt(std::pow(dx, 2) + std::pow(dy, 2) + std::pow(dz, 2)) / g.param;
    }
  }
  return path;
}

double beam_radius_of_influence(const Problem& pb, const BeamSpec& b) {
  if (pb.buffer >= 0.0) return pb.buffer;
  double r;
  if (pb.t_hist_inv < std::exp(3.0 / 2.0)) r = b.ax * std::sqrt(std::log(pb.t_hist_inv) / 3.0);
  else r = b.ax * std::pow(pb.t_hist_inv, (1.0 / 3.0)) / std::sqrt(2.0 * std::exp(1.0));
  const double beta = std::pow(3.0 / 3.14159, 1.5) * b.q / (pb.mat.rho * pb.mat.cps);
  const double x = (pb.mat.T_liq - pb.mat.T_init) * pb.p_hist;
  const double ratio = beta / (x * b.ax * b.ax * b.ax);
  double r2;
  if (ratio < std::exp(3.0 / 2.0)) r2 = b.ax * std::sqrt(std::log(ratio) / 3.0);
  else r2 = std::pow(beta / x, (1.0 / 3.0)) / std::sqrt(2.0 * std::exp(1.0));
  return r > r2 ? r : r2;
}

void finalize_problem(Problem& pb) {
  if (pb.beams.empty()) throw Error(kBadInput, "no beams");
  if (pb.beams.size() != pb.paths.size())
    throw Error(kBadInput, "number of beam files and path files differ");
  for (auto& b : pb.beams) b.nond_dt = b.ax * b.ax / pb.mat.alpha;
  pb.scan_end = 0.0;
  for (auto& p : pb.paths) pb.scan_end = std::max(pb.scan_end, p.back().time);
  GridSpec& g = pb.grid;
  if (pb.custom_points) {
    // points replace the grid; the grid numbers stay as read and are not used
    if (pb.points.empty() || pb.points.size() % 3) throw Error(kBadInput, "custom points need x, y, z triples");
    if (pb.points.size() / 3 > static_cast<size_t>(INT_MAX)) throw Error(kBadInput, "more than 2^31-1 points");
    g.pnum = static_cast<long long>(pb.points.size() / 3);
  } else {
    if (g.xnum < 1 || g.ynum < 1 || g.znum < 1) throw Error(kBadInput, "grid dimensions must be positive");
    if (g.xnum > 65535 || g.ynum > 65535 || g.znum > 65535)
      throw Error(kBadInput, "grid dimensions above 65535 (the reference stores them as uint16)");
    g.pnum = static_cast<long long>(g.xnum) * g.ynum * g.znum;
    if (g.pnum > INT_MAX) throw Error(kBadInput, "more than 2^31-1 grid points");
    g.sx = g.xnum > 1 ? (g.xmax - g.xmin) / (g.xnum - 1) : 0.0;
    g.sy = g.ynum > 1 ? (g.ymax - g.ymin) / (g.ynum - 1) : 0.0;
    g.sz = g.znum > 1 ? (g.zmax - g.zmin) / (g.znum - 1) : 0.0;
  }
  // Run.cpp:58-77 has no Stork snapshots: the reference prints an error and runs nothing
  if (pb.mode == Mode::Snapshots && pb.tracking == Tracking::Stork)
    throw Error(kBadInput, "Stork tracking needs Solidification mode");
  if (pb.out.T_hist && pb.tracking == Tracking::Stork && !pb.custom_points)
    throw Error(kBadInput, "T_hist with Stork tracking on a grid is outside this library's scope");
  if (pb.secondary) throw Error(kBadInput, "secondary solidification is outside this library's scope");
  if (pb.compress) {
    // The compressed quadrature's R_max skip (Calc.cpp:290-302) rereads one
    // path vertex until it lies inside the R_max box: the reference does not
    // terminate once its walk meets a vertex (after t = 0) outside the box.
    for (const auto& path : pb.paths)
      for (size_t s = 1; s < path.size(); ++s) {
        const Segment& v = path[s];
        const bool inside = !(v.x > g.xmax + pb.r_max || v.x < g.xmin - pb.r_max) &&
                            !(v.y > g.ymax + pb.r_max || v.y < g.ymin - pb.r_max);
        if (v.time > 0.0 && !inside)
          throw Error(kBadInput,
                      "path compression with a path vertex outside the R_max box (the reference does not terminate)");
      }
  }
  if (pb.mode == Mode::Solidification && !(pb.dt > 0.0)) throw Error(kBadInput, "Timestep must be positive");
}

Problem load_problem(const std::string& paraminput) {
  Problem pb;
  {
    std::ifstream probe(paraminput);
    if (!probe) throw Error(kBadInput, "cannot open " + paraminput);
  }
  size_t slash = paraminput.find_last_of('/');
  pb.data_dir = slash == std::string::npos ? "Data" : paraminput.substr(0, slash) + "/Data";
  mkdir(pb.data_dir.c_str(), 0777);

  Blocks top(paraminput, {{"Simulation", {"Name", "Mode", "Material", "Beam", "Path"}},
                          {"Options", {"Domain", "Output", "Settings"}}});
  const std::string nm = top.get("Simulation", "Name");
  pb.name = nm.empty() ? "N/A" : nm;
  const std::string f_mode = top.get("Simulation", "Mode"), f_mat = top.get("Simulation", "Material"),
                    f_beam = top.get("Simulation", "Beam"), f_path = top.get("Simulation", "Path"),
                    f_dom = top.get("Options", "Domain"), f_out = top.get("Options", "Output"),
                    f_set = top.get("Options", "Settings");

  for (auto& f : expand(f_beam)) pb.beams.push_back(read_beam(f));

  Blocks mat(f_mat, {{"Constants", {"T_0", "T_L", "k", "c", "p"}}, {"CET", {"N0", "n", "a"}}});
  pb.mat.T_init = as_double(mat.get("Constants", "T_0"), 1273.0, "T_0");
  pb.mat.T_liq = as_double(mat.get("Constants", "T_L"), 1610.0, "T_L");
  pb.mat.kon = as_double(mat.get("Constants", "k"), 26.6, "k");
  pb.mat.cps = as_double(mat.get("Constants", "c"), 600.0, "c");
  pb.mat.rho = as_double(mat.get("Constants", "p"), 7451.0, "p");
  pb.mat.cet_N0 = as_double(mat.get("CET", "N0"), DBL_MAX, "N0");
  pb.mat.cet_n = as_double(mat.get("CET", "n"), DBL_MAX, "n");
  pb.mat.cet_a = as_double(mat.get("CET", "a"), DBL_MAX, "a");
  pb.mat.alpha = pb.mat.kon / (pb.mat.rho * pb.mat.cps);

  for (auto& f : expand(f_path)) {
    std::vector<double> rows;
    pb.paths.push_back(read_path(f, rows));
    pb.path_rows.push_back(rows);
  }
  pb.scan_end = 0.0;
  for (auto& p : pb.paths) pb.scan_end = std::max(pb.scan_end, p.back().time);

  Blocks modes(f_mode, {{"Snapshots", {}}, {"Solidification", {}}}, true);
  const int n_snap = modes.count["Snapshots"], n_sol = modes.count["Solidification"];
  if (n_snap + n_sol > 1) throw Error(kBadInput, "Too many modes in file " + f_mode);
  if (n_snap + n_sol == 0) throw Error(kBadInput, "No modes in file " + f_mode);
  if (n_snap) {
    pb.mode = Mode::Snapshots;
    Blocks m(f_mode, {{"Snapshots", {"Times", "ScanFracs", "Tracking"}}});
    const std::string times = m.get("Snapshots", "Times"), fracs = m.get("Snapshots", "ScanFracs");
    if (!times.empty() && !fracs.empty()) throw Error(kBadInput, "Too many time types for snapshot in " + f_mode);
    for (double v : comma_list(times)) pb.snapshot_times.push_back(v);
    for (double v : comma_list(fracs)) pb.snapshot_times.push_back(v / 100.0 * pb.scan_end);
    pb.tracking = tracking_of(m.get("Snapshots", "Tracking"));
  } else {
    pb.mode = Mode::Solidification;
    Blocks m(f_mode, {{"Solidification", {"Tracking", "Timestep", "OutputFrequency", "CheckRadius", "Secondary"}}});
    pb.tracking = tracking_of(m.get("Solidification", "Tracking"));
    pb.dt = as_double(m.get("Solidification", "Timestep"), 1e-5, "Timestep");
    pb.out_freq = as_int(m.get("Solidification", "OutputFrequency"), INT_MAX, "OutputFrequency");
    pb.radius_check = as_double(m.get("Solidification", "CheckRadius"), 1.0, "CheckRadius");
    pb.secondary = as_bool(m.get("Solidification", "Secondary"), false, "Secondary");
  }

  Blocks dom(f_dom, {{"X", {"Min", "Max", "Num", "Res"}}, {"Y", {"Min", "Max", "Num", "Res"}},
                     {"Z", {"Min", "Max", "Num", "Res"}},
                     {"BoundaryConditions", {"X_min", "X_max", "Y_min", "Y_max", "Z_min", "Reflections"}},
                     {"Custom", {"File"}}});
  GridSpec& g = pb.grid;
  g.xmin = as_double(dom.get("X", "Min"), -DBL_MAX, "X Min");
  g.xmax = as_double(dom.get("X", "Max"), DBL_MAX, "X Max");
  g.xnum = as_int(dom.get("X", "Num"), INT_MAX, "X Num");
  g.xres = as_double(dom.get("X", "Res"), 50e-6, "X Res");
  g.ymin = as_double(dom.get("Y", "Min"), -DBL_MAX, "Y Min");
  g.ymax = as_double(dom.get("Y", "Max"), DBL_MAX, "Y Max");
  g.ynum = as_int(dom.get("Y", "Num"), INT_MAX, "Y Num");
  g.yres = as_double(dom.get("Y", "Res"), 50e-6, "Y Res");
  g.zmin = as_double(dom.get("Z", "Min"), -DBL_MAX, "Z Min");
  g.zmax = as_double(dom.get("Z", "Max"), DBL_MAX, "Z Max");
  g.znum = as_int(dom.get("Z", "Num"), INT_MAX, "Z Num");
  g.zres = as_double(dom.get("Z", "Res"), 50e-6, "Z Res");
  WallSpec& w = pb.walls;
  w.xmin = as_double(dom.get("BoundaryConditions", "X_min"), DBL_MAX, "BC X min");
  w.xmax = as_double(dom.get("BoundaryConditions", "X_max"), DBL_MAX, "BC X max");
  w.ymin = as_double(dom.get("BoundaryConditions", "Y_min"), DBL_MAX, "BC Y min");
  w.ymax = as_double(dom.get("BoundaryConditions", "Y_max"), DBL_MAX, "BC Y max");
  w.zmin = as_double(dom.get("BoundaryConditions", "Z_min"), DBL_MAX, "BC Z min");
  w.reflections = as_int(dom.get("BoundaryConditions", "Reflections"), INT_MAX, "BC reflections");
  const std::string f_pts = dom.get("Custom", "File");
  if (!f_pts.empty()) {  // Init.cpp:696-703
    pb.custom_points = true;
    pb.points = read_points(f_pts);
    if (pb.points.empty()) throw Error(kBadInput, "ERROR: No points found in " + f_pts);
  }
  scan_bounds(pb);
  if (!pb.custom_points) size_domain(g);  // Init.cpp:869

  Blocks out(f_out, {{"Grid", {"x", "y", "z"}}, {"Temperature", {"T", "T_hist"}},
                     {"Solidification", {"tSol", "G", "Gx", "Gy", "Gz", "V", "dTdt", "eqFrac", "depth",
                                         "numMelt", "RDF", "MP_Stats"}},
                     {"Solidification+", {"H", "Hx", "Hy", "Hz"}}});
  OutputSpec& o = pb.out;
  o.x = as_bool(out.get("Grid", "x"), true, "x");
  o.y = as_bool(out.get("Grid", "y"), true, "y");
  o.z = as_bool(out.get("Grid", "z"), true, "z");
  o.T = as_bool(out.get("Temperature", "T"), true, "T");
  o.T_hist = as_bool(out.get("Temperature", "T_hist"), false, "T_hist");
  o.tSol = as_bool(out.get("Solidification", "tSol"), false, "tSol");
  o.G = as_bool(out.get("Solidification", "G"), false, "G");
  o.Gx = as_bool(out.get("Solidification", "Gx"), false, "Gx");
  o.Gy = as_bool(out.get("Solidification", "Gy"), false, "Gy");
  o.Gz = as_bool(out.get("Solidification", "Gz"), false, "Gz");
  o.V = as_bool(out.get("Solidification", "V"), false, "V");
  o.dTdt = as_bool(out.get("Solidification", "dTdt"), false, "dTdt");
  o.eqFrac = as_bool(out.get("Solidification", "eqFrac"), false, "eqFrac");
  o.depth = as_bool(out.get("Solidification", "depth"), false, "depth");
  o.numMelt = as_bool(out.get("Solidification", "numMelt"), false, "numMelt");
  o.RDF = as_bool(out.get("Solidification", "RDF"), false, "RDF");
  o.mp_stats = as_bool(out.get("Solidification", "MP_Stats"), false, "MP_Stats");
  o.H = as_bool(out.get("Solidification+", "H"), false, "H");
  o.Hx = as_bool(out.get("Solidification+", "Hx"), false, "Hx");
  o.Hy = as_bool(out.get("Solidification+", "Hy"), false, "Hy");
  o.Hz = as_bool(out.get("Solidification+", "Hz"), false, "Hz");

  Blocks set(f_set, {{"Temperature", {"Sol_Tol", "Sol_Iter", "Cutoff_Peak", "Cutoff_T0TL"}},
                     {"Path", {"Buffer", "Compression"}}, {"Compute", {"MaxThreads", "PINT"}},
                     {"MPI", {"Overlap"}}});
  pb.sol_tol = as_double(set.get("Temperature", "Sol_Tol"), 1e-3, "Sol_Tol");
  pb.sol_iter = as_int(set.get("Temperature", "Sol_Iter"), 10, "Sol_Iter");
  pb.t_hist_inv = as_double(set.get("Temperature", "Cutoff_Peak"), 1e-9, "Cutoff_Peak");
  pb.p_hist = as_double(set.get("Temperature", "Cutoff_T0TL"), 1e-2, "Cutoff_T0TL");
  pb.r_max = as_double(set.get("Path", "Buffer"), -1.0, "Buffer");
  pb.compress = as_int(set.get("Path", "Compression"), 0, "Compression");
  pb.threads = as_int(set.get("Compute", "MaxThreads"), 1, "MaxThreads");
  radius_of_influence(pb);
  finalize_problem(pb);
  return pb;
}

}  // namespace thesis
