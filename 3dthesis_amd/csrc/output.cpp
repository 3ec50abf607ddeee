#include "output.hpp"

#include <cstdio>
#include <cstring>

#include "common.hpp"
#include "device.hpp"

namespace thesis {

enum ColId { cX, cY, cZ, cT, cTSol, cG, cV, cGx, cGy, cGz, cDTdt, cEq, cDepth, cNumMelt, cMpW, cMpL, cMpD };

std::vector<Column> output_columns(const Problem& pb) {
  std::vector<Column> c;
  const OutputSpec& o = pb.out;
  if (o.x) c.push_back({"x", cX});
  if (o.y) c.push_back({"y", cY});
  if (o.z) c.push_back({"z", cZ});
  if (o.T) c.push_back({"T", cT});
  if (pb.mode == Mode::Solidification) {
    if (o.tSol) c.push_back({"tSol", cTSol});
    if (o.G) c.push_back({"G", cG});
    if (o.V) c.push_back({"V", cV});
    if (o.Gx) c.push_back({"Gx", cGx});
    if (o.Gy) c.push_back({"Gy", cGy});
    if (o.Gz) c.push_back({"Gz", cGz});
    if (o.dTdt) c.push_back({"dTdt", cDTdt});
    if (o.eqFrac) c.push_back({"eqFrac", cEq});
    if (o.depth && pb.tracking == Tracking::Surface) c.push_back({"depth", cDepth});
    if (o.numMelt) c.push_back({"numMelt", cNumMelt});
    if (o.mp_stats) {
      c.push_back({"MP_width", cMpW});
      c.push_back({"MP_length", cMpL});
      c.push_back({"MP_depth", cMpD});
    }
  }
  return c;
}

namespace {

class Writer {
 public:
  explicit Writer(const std::string& path) : f_(std::fopen(path.c_str(), "w")) { buf_.reserve(1 << 20); }
  ~Writer() {
    flush();
    if (f_) std::fclose(f_);
  }
  bool ok() const { return f_ != nullptr; }
  void put(const char* s) { buf_.append(s); maybe_flush(); }
  void put(char ch) { buf_.push_back(ch); }
  void num(double v, int digits) {
    char tmp[64];
    std::snprintf(tmp, sizeof tmp, "%.*g", digits, v);
    buf_.append(tmp);
    maybe_flush();
  }
  void flush() {
    if (f_ && !buf_.empty()) std::fwrite(buf_.data(), 1, buf_.size(), f_);
    buf_.clear();
  }

 private:
  void maybe_flush() {
    if (buf_.size() > (1u << 20)) flush();
  }
  std::FILE* f_;
  std::string buf_;
};

void coords(const Problem& pb, const GridGeom& gg, long long p, double& x, double& y, double& z) {
  if (pb.custom_points) {
    x = pb.points[3 * p];
    y = pb.points[3 * p + 1];
    z = pb.points[3 * p + 2];
    return;
  }
  int i, j, k;
  grid_ijk(gg, static_cast<int>(p), i, j, k);
  x = grid_x(gg, i);
  y = grid_y(gg, j);
  z = grid_z(gg, k);
}

GridGeom host_geom(const Problem& pb) {
  const GridSpec& g = pb.grid;
  return GridGeom{g.xnum, g.ynum, g.znum, 0, g.xmin, g.ymin, g.zmin, g.xmax, g.ymax, g.zmax, g.sx, g.sy, g.sz};
}

}  // namespace

void write_rdf(const Problem& pb, const RdfEvents& ev, const std::vector<uint8_t>& out_flag,
               const std::string& full_dir) {
  const GridGeom gg = host_geom(pb);
  Writer w(pb.data_dir + "/" + pb.name + ".RDF.Final.csv");
  if (!w.ok()) {
    std::fprintf(stderr, "Exception writing data file, check that Data directory exists\n");
    return;
  }
  Writer full_w(full_dir.empty() ? std::string("/dev/null") : full_dir + "/" + pb.name + ".RDF.full.txt");
  Writer* full = (!full_dir.empty() && full_w.ok()) ? &full_w : nullptr;
  w.put("x,y,z,tm,tl,cr\n");
  char tmp[32];
  for (long long p = 0; p < pb.grid.pnum; ++p) {
    if (!out_flag.empty() && !out_flag[p]) continue;
    const std::vector<double>& tm = ev.tm[p];
    if (tm.empty()) continue;
    double x, y, z;
    coords(pb, gg, p, x, y, z);
    for (size_t e = 0; e < tm.size(); ++e) {
      const double tl = e < ev.tl[p].size() ? ev.tl[p][e] : 0.0;
      const double cr = e < ev.cr[p].size() ? ev.cr[p][e] : 0.0;
      const double v[6] = {x, y, z, tm[e], tl, cr};
      for (int c = 0; c < 6; ++c) {
        if (c) w.put(',');
        w.num(v[c], 6);
      }
      w.put('\n');
      if (full) {
        std::snprintf(tmp, sizeof tmp, "%lld", p);
        full->put(tmp);
        for (int c = 3; c < 6; ++c) {
          full->put(' ');
          full->num(v[c], 17);
        }
        full->put('\n');
      }
    }
  }
}

void write_rrdf(const Problem& pb, const std::vector<uint32_t>& idx, const std::vector<double>& ts,
                const std::vector<double>& Ts, const std::string& full_dir) {
  Writer w(pb.data_dir + "/" + pb.name + ".RRDF.csv");
  if (!w.ok()) {
    std::fprintf(stderr, "Exception writing data file, check that Data directory exists\n");
    return;
  }
  Writer full_w(full_dir.empty() ? std::string("/dev/null") : full_dir + "/" + pb.name + ".RRDF.full.txt");
  Writer* full = (!full_dir.empty() && full_w.ok()) ? &full_w : nullptr;
  w.put("x,y,z,t_prev,t_cur");
  w.put(",T000_prev,T001_prev,T010_prev,T011_prev,T100_prev,T101_prev,T110_prev,T111_prev");
  w.put(",T000_cur,T001_cur,T010_cur,T011_cur,T100_cur,T101_cur,T110_cur,T111_cur\n");
  char tmp[48];
  const size_t n = ts.size() / 2;
  for (size_t r = 0; r < n; ++r) {
    std::snprintf(tmp, sizeof tmp, "%u,%u,%u,", idx[3 * r], idx[3 * r + 1], idx[3 * r + 2]);
    w.put(tmp);
    w.num(ts[2 * r], 6);
    w.put(',');
    w.num(ts[2 * r + 1], 6);
    for (int m = 0; m < 16; ++m) {
      w.put(',');
      w.num(Ts[16 * r + m], 6);
    }
    w.put('\n');
    if (full) {
      std::snprintf(tmp, sizeof tmp, "%u %u %u", idx[3 * r], idx[3 * r + 1], idx[3 * r + 2]);
      full->put(tmp);
      for (int m = 0; m < 2; ++m) {
        full->put(' ');
        full->num(ts[2 * r + m], 17);
      }
      for (int m = 0; m < 16; ++m) {
        full->put(' ');
        full->num(Ts[16 * r + m], 17);
      }
      full->put('\n');
    }
  }
}

void write_thist(const Problem& pb, const std::vector<double>& t, const std::vector<double>& T,
                 const std::vector<uint8_t>& out_flag, const std::string& full_dir) {
  const long long np = pb.grid.pnum;
  const size_t ne = t.size();
  const GridGeom gg = host_geom(pb);
  char tmp[64];
  auto fixed = [&](Writer& w, double v) {  // std::to_string(double)
    std::snprintf(tmp, sizeof tmp, "%f", v);
    w.put(tmp);
  };
  if (pb.tracking == Tracking::None) {
    Writer w(pb.data_dir + "/" + pb.name + "T.hist.csv");
    if (!w.ok()) {
      std::fprintf(stderr, "Exception writing data file, check that Data directory exists\n");
    } else {
      w.put("x,y,z");
      for (size_t e = 0; e < ne; ++e) {
        w.put(',');
        fixed(w, t[e]);
      }
      w.put('\n');
      for (long long p = 0; p < np; ++p) {
        double x, y, z;
        coords(pb, gg, p, x, y, z);
        w.num(x, 6);
        w.put(',');
        w.num(y, 6);
        w.put(',');
        w.num(z, 6);
        for (size_t e = 0; e < ne; ++e) {
          w.put(',');
          fixed(w, T[e * np + p]);
        }
        w.put('\n');
      }
    }
  } else {
    auto flagged = [&](long long p) { return out_flag.empty() || out_flag[p]; };
    {
      Writer w(pb.data_dir + "/" + pb.name + ".Points.csv");
      if (w.ok()) {
        w.put("p,x,y,z\n");
        for (long long p = 0; p < np; ++p) {
          if (!flagged(p)) continue;
          double x, y, z;
          coords(pb, gg, p, x, y, z);
          std::snprintf(tmp, sizeof tmp, "%lld,", p);
          w.put(tmp);
          w.num(x, 6);
          w.put(',');
          w.num(y, 6);
          w.put(',');
          w.num(z, 6);
          w.put('\n');
        }
      }
    }
    for (long long p = 0; p < np; ++p) {
      if (!flagged(p)) continue;
      Writer w(pb.data_dir + "/" + pb.name + ".T_hist." + std::to_string(p) + ".csv");
      if (!w.ok()) continue;
      w.put("t,T\n");
      for (size_t e = 0; e < ne; ++e) {
        w.num(t[e], 6);
        w.put(',');
        w.num(T[e * np + p], 6);
        w.put('\n');
      }
    }
  }
  if (!full_dir.empty()) {
    Writer w(full_dir + "/" + pb.name + ".T_hist.full.txt");
    if (w.ok())
      for (long long p = 0; p < np; ++p)
        for (size_t e = 0; e < ne; ++e) {
          std::snprintf(tmp, sizeof tmp, "%lld ", p);
          w.put(tmp);
          w.num(t[e], 17);
          w.put(' ');
          w.num(T[e * np + p], 17);
          w.put('\n');
        }
  }
}

void write_thist_tracked(const Problem& pb, const HistRecords& h, const std::vector<uint8_t>& out_flag,
                         const std::string& full_dir) {
  const long long np = pb.grid.pnum;
  const GridGeom gg = host_geom(pb);
  char tmp[64];
  auto flagged = [&](long long p) { return out_flag.empty() || out_flag[p]; };
  {
    Writer w(pb.data_dir + "/" + pb.name + ".Points.csv");
    if (w.ok()) {
      w.put("p,x,y,z\n");
      for (long long p = 0; p < np; ++p) {
        if (!flagged(p)) continue;
        double x, y, z;
        coords(pb, gg, p, x, y, z);
        std::snprintf(tmp, sizeof tmp, "%lld,", p);
        w.put(tmp);
        w.num(x, 6);
        w.put(',');
        w.num(y, 6);
        w.put(',');
        w.num(z, 6);
        w.put('\n');
      }
    }
  }
  size_t e = 0;
  const size_t ne = h.p.size();
  for (long long p = 0; p < np; ++p) {
    const size_t e0 = e;
    while (e < ne && h.p[e] == p) ++e;
    if (!flagged(p)) continue;
    Writer w(pb.data_dir + "/" + pb.name + ".T_hist." + std::to_string(p) + ".csv");
    if (!w.ok()) continue;
    w.put("t,T\n");
    for (size_t k = e0; k < e; ++k) {
      w.num(h.t[k], 6);
      w.put(',');
      w.num(h.T[k], 6);
      w.put('\n');
    }
  }
  if (!full_dir.empty()) {
    Writer w(full_dir + "/" + pb.name + ".T_hist.full.txt");
    if (w.ok())
      for (size_t k = 0; k < ne; ++k) {
        std::snprintf(tmp, sizeof tmp, "%d ", h.p[k]);
        w.put(tmp);
        w.num(h.t[k], 17);
        w.put(' ');
        w.num(h.T[k], 17);
        w.put('\n');
      }
  }
}

void write_output(const Problem& pb, const HostState& st, const std::string& tag, const std::string& full_dir) {
  const std::vector<Column> cols = output_columns(pb);
  if (cols.empty()) return;
  Writer csv(pb.data_dir + "/" + pb.name + "." + tag + ".csv");
  if (!csv.ok()) {
    std::fprintf(stderr, "Exception writing data file, check that Data directory exists\n");
    return;
  }
  Writer* full = nullptr;
  Writer full_w(full_dir.empty() ? std::string("/dev/null") : full_dir + "/" + pb.name + "." + tag + ".full.txt");
  if (!full_dir.empty() && full_w.ok()) full = &full_w;

  for (size_t c = 0; c < cols.size(); ++c) {
    if (c) csv.put(',');
    csv.put(cols[c].name);
  }
  csv.put('\n');
  const GridSpec& g = pb.grid;
  const GridGeom gg = host_geom(pb);
  // Grid.h:340-342 (with custom points the depth column stays 0 whatever this is)
  const double zstep = (g.zmax - g.zmin) / (static_cast<uint16_t>(g.znum) - 1);
  for (long long p = 0; p < g.pnum; ++p) {
    if (!st.out_flag.empty() && !st.out_flag[p]) continue;
    double x, y, z;
    coords(pb, gg, p, x, y, z);
    if (full) {
      char tmp[32];
      std::snprintf(tmp, sizeof tmp, "%lld", p);
      full->put(tmp);
    }
    for (size_t c = 0; c < cols.size(); ++c) {
      double v = 0.0;
      switch (cols[c].id) {
        case cX: v = x; break;
        case cY: v = y; break;
        case cZ: v = z; break;
        case cT: v = st.T[p]; break;
        case cTSol: v = st.tSol[p]; break;
        case cG: v = st.G[p]; break;
        case cV: v = st.V[p]; break;
        case cGx: v = st.Gx[p]; break;
        case cGy: v = st.Gy[p]; break;
        case cGz: v = st.Gz[p]; break;
        case cDTdt: v = st.dTdt[p]; break;
        case cEq: v = st.eqFrac[p]; break;
        case cDepth: v = st.depth[p] * zstep; break;
        case cNumMelt: v = st.numMelt[p]; break;
        case cMpW: v = st.mp_w.empty() ? 0.0 : st.mp_w[p]; break;
        case cMpL: v = st.mp_l.empty() ? 0.0 : st.mp_l[p]; break;
        case cMpD: v = st.mp_d.empty() ? 0.0 : st.mp_d[p]; break;
      }
      if (c) csv.put(',');
      csv.num(v, 6);
      if (full) {
        full->put(' ');
        full->num(v, 17);
      }
    }
    csv.put('\n');
    if (full) full->put('\n');
  }
}

}  // namespace thesis
