// extern "C" boundary: status codes, thread-local error text, no exceptions out.
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/thesis_amd.h"
#include "adjoint.hpp"
#include "batch_host.hpp"
#include "engine.hpp"

using namespace thesis;

struct thesis_ctx {
  Problem pb;
  std::unique_ptr<Quadrature> quad;  // host node builder (no GPU needed)
  std::unique_ptr<Engine> eng;       // opened on first GPU use
  std::unique_ptr<BatchRunner> batch;
  Design design{};
  Options opt;
  Engine& engine() {
    if (!eng) {
      eng.reset(new Engine(pb));
      eng->opt = opt;
    }
    return *eng;
  }
  void sync_options() {
    if (eng) eng->opt = opt;
  }
};

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    g_err.clear();
    f();
    return THESIS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return THESIS_E_CAPACITY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return THESIS_E_ARG;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw Error(kBadArgument, msg);
}
}  // namespace

extern "C" {

const char* thesis_last_error(void) { return g_err.c_str(); }

int thesis_ctx_create(const char* paraminput, thesis_ctx** out) {
  return guarded([&] {
    need(paraminput && out, "null argument");
    *out = nullptr;
    std::unique_ptr<thesis_ctx> c(new thesis_ctx);
    c->pb = load_problem(paraminput);
    c->quad.reset(new Quadrature(c->pb));
    if (c->pb.beams.size() == 1) {  // the deck's own beam and path as the initial design
      const BeamSpec& b = c->pb.beams[0];
      c->design.beam[0] = b.power;
      c->design.beam[1] = b.eff;
      c->design.beam[2] = b.ax;
      c->design.beam[3] = b.ay;
      c->design.beam[4] = b.az;
      c->design.rows = c->pb.path_rows[0];
    }
    *out = c.release();
  });
}

int thesis_ctx_destroy(thesis_ctx* ctx) {
  return guarded([&] { delete ctx; });
}

int thesis_set_option(thesis_ctx* ctx, const char* key, double value) {
  return guarded([&] {
    need(ctx && key, "null argument");
    const std::string k(key);
    if (k == "precision") {
      need(value == 0 || value == 1, "precision must be 0 (FP64) or 1 (FP32c)");
      ctx->opt.precision = static_cast<int>(value);
    } else if (k == "abs_mode") {
      need(value == 0 || value == 1, "abs_mode must be 0 or 1");
      ctx->pb.abs_mode = value == 0 ? AbsMode::IntTrunc : AbsMode::Fabs;
      if (ctx->eng) throw Error(kBadArgument, "abs_mode must be set before the first GPU call");
      ctx->quad.reset(new Quadrature(ctx->pb));
    } else if (k == "ppt") {
      need(value == 1 || value == 2 || value == 4, "ppt must be 1, 2 or 4");
      ctx->opt.ppt = static_cast<int>(value);
    } else if (k == "progress") {
      ctx->opt.progress = value != 0;
    } else if (k == "outputs") {
      ctx->opt.outputs = value != 0;
    } else if (k == "timing") {
      ctx->opt.timing = value != 0;
    } else if (k == "phase_times") {
      ctx->opt.phase_times = value != 0;
    } else if (k == "tracker_lds") {
      ctx->opt.tracker_lds = value != 0;
    } else if (k == "async_sol") {
      need(value == 0 || value == 1, "async_sol must be 0 or 1");
      ctx->opt.async_sol = static_cast<int>(value);
    } else if (k == "volume_switch") {
      need(value >= 0 && value <= 1e6, "volume_switch must be 0..1e6 windows");
      ctx->opt.volume_switch = static_cast<int>(value);
    } else if (k == "box_margin") {
      need(value >= 0 && value <= 64, "box_margin must be 0..64 cells");
      ctx->opt.box_margin = static_cast<int>(value);
    } else if (k == "depth_margin") {
      need(value >= 0 && value <= 64, "depth_margin must be 0..64 layers");
      ctx->opt.depth_margin = static_cast<int>(value);
    } else if (k == "pipeline") {
      need(value >= 1 && value <= kPipe, "pipeline must be 1 to 6 windows");
      ctx->opt.pipeline = static_cast<int>(value);
    } else if (k == "lds_blocks_per_cu") {
      need(value >= 1 && value <= 16, "lds_blocks_per_cu must be 1..16");
      set_lds_blocks_per_cu(static_cast<int>(value));
    } else if (k == "window_sol_batch") {
      need(value >= 1 && value <= (1 << 30), "window_sol_batch must be 1 .. 2^30 entries");
      ctx->opt.window_sol_batch = static_cast<int>(value);
    } else if (k == "step_sol_cus") {
      need(value >= 0 && value <= 1024, "step_sol_cus must be 0..1024");
      ctx->opt.step_sol_cus = static_cast<int>(value);
    } else if (k == "step_sol_batch") {
      need(value >= 1 && value <= (1 << 30), "step_sol_batch must be 1 .. 2^30 entries");
      ctx->opt.step_sol_batch = static_cast<int>(value);
    } else if (k == "sol_stream") {
      need(value >= 0 && value <= 3, "sol_stream must be 0 .. 3");
      ctx->opt.sol_stream = static_cast<int>(value);
    } else if (k == "surface_rows") {
      need(value == 0 || value == 1 || value == 2, "surface_rows must be 0, 1 or 2");
      ctx->opt.surface_rows = static_cast<int>(value);
    } else if (k == "surface_threads") {
      need(value == 0 || value == 64 || value == 256 || value == 1024, "surface_threads must be 0, 64, 256 or 1024");
      ctx->opt.surface_threads = static_cast<int>(value);
    } else if (k == "window") {
      need(value >= 0 && value <= 4096, "window must be 0..4096 steps");
      ctx->opt.window = static_cast<int>(value);
    } else {
      throw Error(kBadArgument, "unknown option " + k);
    }
    ctx->sync_options();
  });
}

int thesis_set_full_dump(thesis_ctx* ctx, const char* dir) {
  return guarded([&] {
    need(ctx, "null context");
    ctx->opt.full_dir = dir ? dir : "";
    ctx->sync_options();
  });
}

int thesis_get_info(const thesis_ctx* ctx, thesis_info* out) {
  return guarded([&] {
    need(ctx && out, "null argument");
    const Problem& p = ctx->pb;
    out->mode = p.mode == Mode::Snapshots ? 0 : 1;
    out->tracking = static_cast<int>(p.tracking);
    out->nbeams = static_cast<int>(p.beams.size());
    out->xnum = p.grid.xnum;
    out->ynum = p.grid.ynum;
    out->znum = p.grid.znum;
    out->pnum = p.grid.pnum;
    out->nsnap = static_cast<int>(p.snapshot_times.size());
    out->scan_end = p.scan_end;
    out->dt = p.dt;
    out->T_init = p.mat.T_init;
    out->T_liq = p.mat.T_liq;
    out->alpha = p.mat.alpha;
    out->xmin = p.grid.xmin;
    out->xmax = p.grid.xmax;
    out->ymin = p.grid.ymin;
    out->ymax = p.grid.ymax;
    out->zmin = p.grid.zmin;
    out->zmax = p.grid.zmax;
  });
}

int thesis_get_snapshot_times(const thesis_ctx* ctx, double* out, int cap) {
  return guarded([&] {
    need(ctx && (out || cap == 0), "null argument");
    const auto& v = ctx->pb.snapshot_times;
    if (static_cast<int>(v.size()) > cap) throw Error(kCapacity, "snapshot time buffer too small");
    std::copy(v.begin(), v.end(), out);
  });
}

int thesis_build_nodes(thesis_ctx* ctx, double t, int is_sol, double* out, size_t cap, size_t* n_nodes) {
  return guarded([&] {
    need(ctx && n_nodes, "null argument");
    static_assert(sizeof(NodeF64) == 8 * sizeof(double), "node layout");
    std::vector<NodeF64> v;
    ctx->quad->build(t, is_sol != 0, v);
    *n_nodes = v.size();
    if (v.size() > cap) throw Error(kCapacity, "node buffer too small");
    if (!v.empty()) {
      need(out != nullptr, "null node buffer");
      std::memcpy(out, v.data(), v.size() * sizeof(NodeF64));
    }
  });
}

int thesis_reset_hints(thesis_ctx* ctx) {
  return guarded([&] {
    need(ctx, "null context");
    ctx->quad->reset_hints();
  });
}

int thesis_eval_T(thesis_ctx* ctx, const double* nodes, size_t n_nodes, int precision, long long p_begin,
                  long long count, double* T_out) {
  return guarded([&] {
    need(ctx && T_out && (nodes || n_nodes == 0), "null argument");
    need(precision == 0 || precision == 1, "bad precision");
    need(p_begin >= 0 && count >= 0 && p_begin + count <= ctx->pb.grid.pnum, "point range outside the grid");
    Engine& e = ctx->engine();
    DevBuf<double> d;
    d.ensure(static_cast<size_t>(count));
    e.eval_grid(reinterpret_cast<const NodeF64*>(nodes), static_cast<int>(n_nodes), precision,
                static_cast<int>(p_begin), static_cast<int>(count), d.get());
    THESIS_HIP(hipMemcpyAsync(T_out, d.get(), count * sizeof(double), hipMemcpyDeviceToHost, e.stream()));
    THESIS_HIP(hipStreamSynchronize(e.stream()));
  });
}

int thesis_eval_T_points(thesis_ctx* ctx, const double* nodes, size_t n_nodes, int precision, const int* points,
                         size_t count, double* T_out) {
  return guarded([&] {
    need(ctx && T_out && points && (nodes || n_nodes == 0), "null argument");
    need(precision == 0 || precision == 1, "bad precision");
    for (size_t i = 0; i < count; ++i)
      need(points[i] >= 0 && points[i] < ctx->pb.grid.pnum, "point index outside the grid");
    Engine& e = ctx->engine();
    DevBuf<double> d;
    DevBuf<int> pl;
    d.ensure(count);
    pl.ensure(count);
    THESIS_HIP(hipMemcpy(pl.get(), points, count * sizeof(int), hipMemcpyHostToDevice));
    e.eval_list(reinterpret_cast<const NodeF64*>(nodes), static_cast<int>(n_nodes), precision, pl.get(),
                static_cast<int>(count), d.get());
    THESIS_HIP(hipMemcpyAsync(T_out, d.get(), count * sizeof(double), hipMemcpyDeviceToHost, e.stream()));
    THESIS_HIP(hipStreamSynchronize(e.stream()));
  });
}

int thesis_upload_nodes(thesis_ctx* ctx, const double* nodes, size_t n_nodes) {
  return guarded([&] {
    need(ctx && (nodes || n_nodes == 0), "null argument");
    Engine& e = ctx->engine();
    e.upload_nodes(reinterpret_cast<const NodeF64*>(nodes), static_cast<int>(n_nodes));
    THESIS_HIP(hipStreamSynchronize(e.stream()));
  });
}

int thesis_eval_T_dev(thesis_ctx* ctx, int precision, long long p_begin, long long count, double* T_dev,
                      void* stream) {
  return guarded([&] {
    need(ctx && T_dev, "null argument");
    need(precision == 0 || precision == 1, "bad precision");
    need(p_begin >= 0 && count >= 0 && p_begin + count <= ctx->pb.grid.pnum, "point range outside the grid");
    Engine& e = ctx->engine();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e.stream();
    e.eval_grid_resident(precision, static_cast<int>(p_begin), static_cast<int>(count), T_dev, s);
  });
}

int thesis_eval_T_static(const thesis_ctx* ctx, const double* nodes_dev, size_t n_nodes, int precision,
                         long long p_begin, long long count, double* T_dev, void* stream) {
  return guarded([&] {
    need(ctx && T_dev && (nodes_dev || n_nodes == 0), "null argument");
    need(precision == 0 || precision == 1, "bad precision");
    need(p_begin >= 0 && count >= 0 && p_begin + count <= ctx->pb.grid.pnum, "point range outside the grid");
    if (n_nodes > static_cast<size_t>(static_node_capacity()))
      throw Error(kCapacity, "allocation-free evaluation takes at most " + std::to_string(static_node_capacity()) +
                                 " nodes; use thesis_upload_nodes + thesis_eval_T_dev");
    if (ctx->pb.custom_points)
      throw Error(kBadArgument, "allocation-free evaluation needs a structured grid; use thesis_eval_T_dev");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device visible");
    const GridSpec& g = ctx->pb.grid;
    const GridGeom geom{g.xnum, g.ynum, g.znum, 0, g.xmin, g.ymin, g.zmin, g.xmax, g.ymax, g.zmax, g.sx, g.sy, g.sz};
    launch_tsum_static(nodes_dev, static_cast<int>(n_nodes), precision, geom, static_cast<int>(p_begin),
                       static_cast<int>(count), ctx->pb.mat.T_init, T_dev, ctx->opt.ppt,
                       static_cast<hipStream_t>(stream));
  });
}

int thesis_eval_sol(thesis_ctx* ctx, double t, const int* points, size_t n, const double* T, const double* T_last,
                    double* out7) {
  return guarded([&] {
    need(ctx && points && T && T_last && out7, "null argument");
    ctx->engine().solidify_points(t, points, static_cast<int>(n), T, T_last, out7);
  });
}

int thesis_run(thesis_ctx* ctx) {
  return guarded([&] {
    need(ctx, "null context");
    Engine& e = ctx->engine();
    e.opt = ctx->opt;
    e.run();
  });
}

int thesis_get_window_stats(const thesis_ctx* ctx, double* out, int cap) {
  return guarded([&] {
    need(ctx && out && cap >= 0, "null argument");
    for (int q = 0; q < cap && q < kWindowFields; ++q) out[q] = ctx->eng ? ctx->eng->stats.window[q] : 0.0;
  });
}

int thesis_get_stats(const thesis_ctx* ctx, thesis_stats* out) {
  return guarded([&] {
    need(ctx && out, "null argument");
    if (!ctx->eng) {
      *out = thesis_stats{};
      return;
    }
    const RunStats& s = ctx->eng->stats;
    out->evals = s.evals;
    out->wall_s = s.wall_s;
    out->steps = static_cast<double>(s.steps);
    out->nodes_built = s.nodes_built;
    out->evals_window = s.evals_window;
    for (int c = 0; c < 3; ++c) {
      out->cat_evals[c] = s.cat_evals[c];
      out->cat_ms[c] = s.cat_ms[c];
      out->cat_launches[c] = static_cast<double>(s.cat_launches[c]);
    }
  });
}

int thesis_batch_set(thesis_ctx* ctx, int nsets, const double* beam5, const int* row_off, const double* rows6,
                     const double* times) {
  return guarded([&] {
    need(ctx && beam5 && row_off && rows6, "null argument");
    Engine& e = ctx->engine();
    if (!ctx->batch) ctx->batch.reset(new BatchRunner(ctx->pb, e.stream()));
    ctx->batch->set(nsets, beam5, row_off, rows6, times);
  });
}

int thesis_batch_eval_dev(thesis_ctx* ctx, int precision, double* T_dev, void* stream) {
  return guarded([&] {
    need(ctx && T_dev, "null argument");
    need(ctx->batch != nullptr, "no batch set");
    need(precision == 0 || precision == 1, "bad precision");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->engine().stream();
    ctx->batch->eval(precision, T_dev, s);
  });
}

int thesis_batch_autotune(thesis_ctx* ctx, int precision, double* T_dev, int* ppt, int* blocks) {
  return guarded([&] {
    need(ctx && T_dev, "null argument");
    need(ctx->batch != nullptr, "no batch set");
    ctx->batch->autotune(precision, T_dev, ctx->engine().stream(), ppt, blocks);
  });
}

int thesis_batch_info(const thesis_ctx* ctx, long long* total_nodes, double* evals) {
  return guarded([&] {
    need(ctx && ctx->batch, "no batch set");
    if (total_nodes) *total_nodes = ctx->batch->total_nodes();
    if (evals) *evals = ctx->batch->evals();
  });
}

int thesis_batch_nodes(thesis_ctx* ctx, int set, double* out, size_t cap, size_t* n_nodes) {
  return guarded([&] {
    need(ctx && ctx->batch && n_nodes, "no batch set");
    std::vector<NodeF64> v;
    ctx->batch->nodes_of(set, v);
    *n_nodes = v.size();
    if (v.size() > cap) throw Error(kCapacity, "node buffer too small");
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(NodeF64));
  });
}

int thesis_design_get(const thesis_ctx* ctx, double* beam5, double* rows6, int cap_rows, int* nrows) {
  return guarded([&] {
    need(ctx && nrows, "null argument");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    *nrows = ctx->design.nrows();
    if (beam5) std::copy(ctx->design.beam, ctx->design.beam + 5, beam5);
    if (ctx->design.nrows() > cap_rows) throw Error(kCapacity, "row buffer too small");
    if (rows6) std::copy(ctx->design.rows.begin(), ctx->design.rows.end(), rows6);
  });
}

int thesis_design_set(thesis_ctx* ctx, const double* beam5, const double* rows6, int nrows) {
  return guarded([&] {
    need(ctx && beam5 && rows6 && nrows > 0, "null argument");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    Design d;
    std::copy(beam5, beam5 + 5, d.beam);
    d.rows.assign(rows6, rows6 + 6 * nrows);
    (void)apply_design(ctx->pb, d);  // validate
    ctx->design = d;
  });
}

namespace {
void design_nodes(thesis_ctx* ctx, double t, std::vector<NodeF64>& nodes, double* t_used) {
  const Problem pd = apply_design(ctx->pb, ctx->design);
  const double tt = t < 0 ? pd.scan_end : t;
  Quadrature q(pd);
  q.build(tt, false, nodes);
  if (t_used) *t_used = tt;
}
}  // namespace

int thesis_forward_dev(thesis_ctx* ctx, int precision, double t, double* T_dev, void* stream) {
  return guarded([&] {
    need(ctx && T_dev, "null argument");
    need(precision == 0 || precision == 1, "bad precision");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    std::vector<NodeF64> nodes;
    design_nodes(ctx, t, nodes, nullptr);
    Engine& e = ctx->engine();
    e.upload_nodes(nodes.data(), static_cast<int>(nodes.size()));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e.stream();
    if (s != e.stream()) THESIS_HIP(hipStreamSynchronize(e.stream()));
    e.eval_grid_resident(precision, 0, static_cast<int>(ctx->pb.grid.pnum), T_dev, s);
  });
}

int thesis_adjoint_dev(thesis_ctx* ctx, double t, const double* dLdT_dev, double* grad) {
  return guarded([&] {
    need(ctx && dLdT_dev && grad, "null argument");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    std::vector<NodeF64> nodes;
    design_nodes(ctx, t, nodes, nullptr);
    Engine& e = ctx->engine();
    e.upload_nodes(nodes.data(), static_cast<int>(nodes.size()));
    const int n = static_cast<int>(nodes.size());
    const int pnum = static_cast<int>(ctx->pb.grid.pnum);
    const int chunk = 4096;
    const int nchunks = (pnum + chunk - 1) / chunk;
    DevBuf<double> partial, S;
    partial.ensure(static_cast<size_t>(nchunks) * std::max(n, 1) * 7);
    S.ensure(static_cast<size_t>(std::max(n, 1)) * 7);
    THESIS_HIP(hipDeviceSynchronize());  // dL/dT may come from another stream
    launch_adjoint(e.resident_f64(), n, e.geom(), pnum, dLdT_dev, chunk, partial.get(), S.get(), e.stream());
    std::vector<double> Sh(static_cast<size_t>(n) * 7);
    if (n) THESIS_HIP(hipMemcpyAsync(Sh.data(), S.get(), Sh.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream()));
    THESIS_HIP(hipStreamSynchronize(e.stream()));
    size_t built = 0;
    const std::vector<double> g = chain_to_design(ctx->pb, ctx->design, t < 0 ? 0.0 : t, t < 0, Sh, &built);
    if (built != nodes.size()) throw Error(kDevice, "adjoint node lists disagree");
    std::copy(g.begin(), g.end(), grad);
  });
}

int thesis_design_chain(thesis_ctx* ctx, double t, const double* S, size_t n_nodes, double* grad) {
  return guarded([&] {
    need(ctx && grad && (S || n_nodes == 0), "null argument");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    const std::vector<double> Sh(S, S + 7 * n_nodes);
    size_t built = 0;
    const std::vector<double> g = chain_to_design(ctx->pb, ctx->design, t < 0 ? 0.0 : t, t < 0, Sh, &built);
    if (built != n_nodes) throw Error(kBadArgument, "S does not match the design's node count");
    std::copy(g.begin(), g.end(), grad);
  });
}

int thesis_design_nodes(thesis_ctx* ctx, double t, double* out, size_t cap, size_t* n_nodes) {
  return guarded([&] {
    need(ctx && n_nodes, "null argument");
    need(ctx->pb.beams.size() == 1, "the design interface needs a single-beam deck");
    std::vector<NodeF64> v;
    design_nodes(ctx, t, v, nullptr);
    *n_nodes = v.size();
    if (v.size() > cap) throw Error(kCapacity, "node buffer too small");
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(NodeF64));
  });
}

}  // extern "C"
