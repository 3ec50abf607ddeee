#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace thesis {

void hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", static_cast<int>(e), hipGetErrorString(e), what,
                file, line);
  throw Error(kDevice, buf);
}

namespace {
std::string zero_pad(int num, int width) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%0*d", width, num);
  return buf;
}
}  // namespace

Engine::Engine(Problem pb) : pb_(std::move(pb)), quad_(pb_) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device visible");
  const GridSpec& g = pb_.grid;
  geom_ = GridGeom{g.xnum, g.ynum, g.znum, 0, g.xmin, g.ymin, g.zmin, g.xmax, g.ymax, g.zmax, g.sx, g.sy, g.sz};
  THESIS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  counters_.ensure(32);
  h_counters_.ensure(64);
  THESIS_HIP(hipMemsetAsync(counters_.get(), 0, 32 * sizeof(int), stream_));
  for (int c = 0; c < kNumCats; ++c) {
    dc_cat_[c].overflow = counters_.get() + cOverflow;
    dc_cat_[c].evals = reinterpret_cast<unsigned long long*>(counters_.get() + 16 + 2 * c);
  }
  dc_ = dc_cat_[kCatSum];
  const PathData& pd = quad_.data();
  d_segs_.ensure(pd.segs.size());
  d_seg_off_.ensure(pd.off.size());
  d_beams_.ensure(pd.beams.size());
  d_hints_.ensure(pd.beams.size());
  h_hints_.ensure(pd.beams.size());
  THESIS_HIP(hipMemcpy(d_segs_.get(), pd.segs.data(), pd.segs.size() * sizeof(PathSeg), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_seg_off_.get(), pd.off.data(), pd.off.size() * sizeof(int), hipMemcpyHostToDevice));
  THESIS_HIP(hipMemcpy(d_beams_.get(), pd.beams.data(), pd.beams.size() * sizeof(BeamQ), hipMemcpyHostToDevice));
}

Engine::~Engine() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

void Engine::sync() { THESIS_HIP(hipStreamSynchronize(stream_)); }

void Engine::stage_nodes(const NodeF64* nodes, int n, NodeSlot& slot) {
  sync();  // staging buffers are reused: the previous copy must have landed
  h64_.ensure(static_cast<size_t>(n));
  h32_.ensure(static_cast<size_t>(n));
  if (n) std::memcpy(h64_.get(), nodes, sizeof(NodeF64) * n);
  const size_t n32 = pack_f32c(h64_.get(), static_cast<size_t>(n), h32_.get());
  slot.f64.ensure(static_cast<size_t>(n));
  slot.f32.ensure(n32);
  slot.n64 = n;
  slot.n32 = static_cast<int>(n32);
  if (n) THESIS_HIP(hipMemcpyAsync(slot.f64.get(), h64_.get(), sizeof(NodeF64) * n, hipMemcpyHostToDevice, stream_));
  if (n32) THESIS_HIP(hipMemcpyAsync(slot.f32.get(), h32_.get(), sizeof(NodeF32c) * n32, hipMemcpyHostToDevice, stream_));
  stats.nodes_built += n;
}

NodeSetDev Engine::view(const NodeSlot& slot) const {
  NodeSetDev v;
  v.f64 = slot.f64.get();
  v.f32 = slot.f32.get();
  v.n64 = slot.n64;
  v.n32 = slot.n32;
  v.prec = opt.precision;
  return v;
}

// ------------------------------------------------------------ stateless entry points

void Engine::eval_grid(const NodeF64* nodes, int n, int precision, int p_begin, int count, double* T_dev) {
  upload_nodes(nodes, n);
  eval_grid_resident(precision, p_begin, count, T_dev, stream_);
}

void Engine::upload_nodes(const NodeF64* nodes, int n) { stage_nodes(nodes, n, slots_[0]); }

int Engine::resident_count(int precision) const { return precision == kFP64 ? slots_[0].n64 : slots_[0].n32; }

void Engine::eval_grid_resident(int precision, int p_begin, int count, double* T_dev, hipStream_t s) {
  if (p_begin < 0 || count < 0 || static_cast<long long>(p_begin) + count > pb_.grid.pnum)
    throw Error(kBadArgument, "point range outside the grid");
  const NodeSlot& n = slots_[0];
  if (precision == kFP64) launch_tsum_grid_f64(n.f64.get(), n.n64, geom_, p_begin, count, pb_.mat.T_init, T_dev, s);
  else launch_tsum_grid_f32c(n.f32.get(), n.n32, geom_, p_begin, count, pb_.mat.T_init, T_dev, opt.ppt, s);
}

void Engine::eval_list(const NodeF64* nodes, int n, int precision, const int* list_dev, int count, double* T_dev) {
  upload_nodes(nodes, n);
  const NodeSlot& s = slots_[0];
  if (precision == kFP64) launch_tsum_list_f64(s.f64.get(), s.n64, geom_, list_dev, count, pb_.mat.T_init, T_dev, stream_);
  else launch_tsum_list_f32c(s.f32.get(), s.n32, geom_, list_dev, count, pb_.mat.T_init, T_dev, stream_);
}

// ------------------------------------------------------------------ run state

void Engine::alloc_state() {
  const size_t n = static_cast<size_t>(pb_.grid.pnum);
  const bool sol = pb_.mode == Mode::Solidification;
  T_.ensure(n);
  st_.T = T_.get();
  const bool flags = sol || pb_.tracking != Tracking::None;
  if (flags) {
    flag_.ensure(n);
    oflag_.ensure(n);
    st_.flag = flag_.get();
    st_.oflag = oflag_.get();
    THESIS_HIP(hipMemsetAsync(flag_.get(), 0, n * sizeof(uint32_t), stream_));
    THESIS_HIP(hipMemsetAsync(oflag_.get(), 0, n, stream_));
  }
  THESIS_HIP(hipMemsetAsync(T_.get(), 0, n * sizeof(double), stream_));
  if (sol) {
    T_last_.ensure(n);
    st_.T_last = T_last_.get();
    THESIS_HIP(hipMemsetAsync(T_last_.get(), 0, n * sizeof(double), stream_));
    const OutputSpec& o = pb_.out;
    auto col = [&](bool on, DevBuf<double>& b, double*& dst) {
      if (!on) { dst = nullptr; return; }
      b.ensure(n);
      dst = b.get();
      THESIS_HIP(hipMemsetAsync(dst, 0, n * sizeof(double), stream_));
    };
    col(o.tSol, tSol_, st_.tSol);
    col(o.G, G_, st_.G);
    col(o.V, V_, st_.V);
    col(o.Gx, Gx_, st_.Gx);
    col(o.Gy, Gy_, st_.Gy);
    col(o.Gz, Gz_, st_.Gz);
    col(o.dTdt, dTdt_, st_.dTdt);
    col(o.eqFrac, eqFrac_, st_.eqFrac);
    col(o.depth && pb_.tracking == Tracking::Surface, depth_, st_.depth);
    if (o.numMelt) {
      numMelt_.ensure(n);
      st_.numMelt = numMelt_.get();
      THESIS_HIP(hipMemsetAsync(st_.numMelt, 0, n * sizeof(uint16_t), stream_));
    }
    const size_t ncol = static_cast<size_t>(pb_.grid.xnum) * pb_.grid.ynum;
    depths_.ensure(ncol);
    depth_max_.ensure(ncol);
    st_.depths = depths_.get();
    st_.depth_max = depth_max_.get();
    THESIS_HIP(hipMemsetAsync(st_.depths, 0, ncol * sizeof(int), stream_));
    THESIS_HIP(hipMemsetAsync(st_.depth_max, 0, ncol * sizeof(double), stream_));
  }
  if (flags) {
    for (int l = 0; l < 8; ++l) {
      const size_t cap = (l == cReset) ? 2 * n + 1024 : n + 1024;
      lst_[l].ensure(cap);
      L_[l] = DList{lst_[l].get(), counters_.get() + l, static_cast<int>(std::min<size_t>(cap, INT_MAX))};
    }
  }
  THESIS_HIP(hipMemsetAsync(counters_.get(), 0, 32 * sizeof(int), stream_));
}

void Engine::read_counts() {
  THESIS_HIP(hipMemcpyAsync(h_counters_.get(), counters_.get(), 32 * sizeof(int), hipMemcpyDeviceToHost, stream_));
  sync();
  if (h_counters_.get()[cOverflow]) throw Error(kCapacity, "a device work list overflowed its capacity");
  if (opt.timing) timer_.collect(stats);
  gather_evals();
}

int Engine::count_of(int idx) { return h_counters_.get()[idx]; }

void Engine::gather_evals() {
  const unsigned long long* ev = reinterpret_cast<const unsigned long long*>(h_counters_.get() + 16);
  stats.evals = host_grid_evals_;
  for (int c = 0; c < kNumCats; ++c) {
    stats.cat_evals[c] = static_cast<double>(ev[c]) + (c == kCatSum ? host_grid_evals_ : 0.0);
    stats.evals += static_cast<double>(ev[c]);
  }
}

LaunchTimer::~LaunchTimer() {
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t LaunchTimer::get() {
  if (used_ == pool_.size()) {
    hipEvent_t e;
    THESIS_HIP(hipEventCreate(&e));
    pool_.push_back(e);
  }
  return pool_[used_++];
}

void LaunchTimer::begin(int cat, hipStream_t s) {
  cats_.push_back(cat);
  THESIS_HIP(hipEventRecord(get(), s));
}

void LaunchTimer::end(hipStream_t s) { THESIS_HIP(hipEventRecord(get(), s)); }

void LaunchTimer::collect(RunStats& st) {
  for (size_t i = 0; i < cats_.size(); ++i) {
    float ms = 0.f;
    THESIS_HIP(hipEventElapsedTime(&ms, pool_[2 * i], pool_[2 * i + 1]));
    st.cat_ms[cats_[i]] += ms;
    st.cat_launches[cats_[i]] += 1;
  }
  cats_.clear();
  used_ = 0;
}

int Engine::count(const DList& L) { return h_counters_.get()[L.n - counters_.get()]; }

void Engine::set_count(const DList& L, int v) { launch_set_int(L.n, v, stream_); }

void Engine::download_state(HostState& hs) {
  const size_t n = static_cast<size_t>(pb_.grid.pnum);
  auto get = [&](const double* src, std::vector<double>& dst) {
    if (!src) return;
    dst.resize(n);
    THESIS_HIP(hipMemcpyAsync(dst.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  };
  get(st_.T, hs.T);
  get(st_.tSol, hs.tSol);
  get(st_.G, hs.G);
  get(st_.V, hs.V);
  get(st_.Gx, hs.Gx);
  get(st_.Gy, hs.Gy);
  get(st_.Gz, hs.Gz);
  get(st_.dTdt, hs.dTdt);
  get(st_.eqFrac, hs.eqFrac);
  get(st_.depth, hs.depth);
  if (st_.numMelt) {
    hs.numMelt.resize(n);
    THESIS_HIP(hipMemcpyAsync(hs.numMelt.data(), st_.numMelt, n * sizeof(uint16_t), hipMemcpyDeviceToHost, stream_));
  }
  if (st_.oflag) {
    hs.out_flag.resize(n);
    THESIS_HIP(hipMemcpyAsync(hs.out_flag.data(), st_.oflag, n, hipMemcpyDeviceToHost, stream_));
  } else {
    hs.out_flag.clear();
  }
  sync();
}

void Engine::output(const std::string& tag) {
  if (!opt.outputs) return;
  HostState hs;
  download_state(hs);
  write_output(pb_, hs, tag, opt.full_dir);
}

SolParams Engine::sol_params() {
  const PathData& pd = quad_.data();
  SolParams sp{};
  sp.segs = d_segs_.get();
  sp.seg_off = d_seg_off_.get();
  sp.beams = d_beams_.get();
  sp.hints = d_hints_.get();
  sp.nbeams = static_cast<int>(pd.beams.size());
  sp.rb = pd.rbox;
  sp.alpha = pd.alpha;
  sp.T_init = pb_.mat.T_init;
  sp.T_liq = pb_.mat.T_liq;
  sp.dt = pb_.dt;
  sp.sol_tol = pb_.sol_tol;
  sp.sol_iter = pb_.sol_iter;
  sp.abs_mode = static_cast<int>(pb_.abs_mode);
  sp.cet_N0 = pb_.mat.cet_N0;
  sp.cet_n = pb_.mat.cet_n;
  sp.cet_a = pb_.mat.cet_a;
  sp.gl = gl_table();
  return sp;
}

void Engine::upload_hints() {
  const std::vector<int>& h = quad_.hints();
  std::copy(h.begin(), h.end(), h_hints_.get());
  THESIS_HIP(hipMemcpyAsync(d_hints_.get(), h_hints_.get(), h.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
}

// Seed candidates of Melt::beam_trace (Melt.cpp:20-183): surface points at the
// 2x2 corners of cells holding path vertices in [t0, t1] and the beam position
// at t1, plus perimeter points for beams outside the grid. The device claim
// flags drop points already evaluated this step.
void Engine::seeds_between(double t0, double t1, std::vector<int>& out) const {
  const GridSpec& g = pb_.grid;
  const int zk = g.znum - 1;
  auto pid = [&](int i, int j) { return (i * g.ynum + j) * g.znum + zk; };
  const PathData& pd = quad_.data();
  for (size_t b = 0; b < pd.beams.size(); ++b) {  // perimeter (Melt.cpp:20-96)
    const PathSeg* P = pd.segs.data() + pd.off[b];
    const int nseg = pd.off[b + 1] - pd.off[b];
    const double rc = pb_.radius_check, ax = pb_.beams[b].ax;
    const double t_sub = (rc * rc - 1) * (ax * ax) / (12.0 * pb_.mat.alpha);
    const double ts = (t1 - t_sub < 0.0) ? 0.0 : t1 - t_sub;
    const double rc2 = rc * rc * ax * ax;
    int s0 = 0, s1 = 0;
    while (P[s0].time < ts && (s0 + 1) != nseg) ++s0;
    while (P[s1].time < t1 && (s1 + 1) != nseg) ++s1;
    auto perim = [&](int fixed, bool x_fixed, int lim, double X, double Y) {
      for (int v = 0; v < lim; ++v) {
        const int i = x_fixed ? fixed : v, j = x_fixed ? v : fixed;
        const double dx = g.xmin + i * g.xres - X, dy = g.ymin + j * g.yres - Y;
        if (dx * dx + dy * dy < rc2) out.push_back(pid(i, j));
      }
    };
    auto around = [&](double X, double Y) {
      const int xg = static_cast<int>(std::floor((X - g.xmin) / g.xres));
      const int yg = static_cast<int>(std::floor((Y - g.ymin) / g.yres));
      if (xg < 0) perim(0, true, g.ynum, X, Y);
      if (xg > g.xnum - 1) perim(g.xnum - 1, true, g.ynum, X, Y);
      if (yg < 0) perim(0, false, g.xnum, X, Y);
      if (yg > g.ynum - 1) perim(g.ynum - 1, false, g.xnum, X, Y);
    };
    for (int s = std::max(s0 - 1, 0); s <= s1; ++s) around(P[s].x, P[s].y);
    if (P[s1].mode == 0 && t1 < P[nseg - 1].time) {
      double xb, yb, zb, q;
      beam_at(P, s1, t1, pd.rbox, xb, yb, zb, q);
      around(xb, yb);
    }
  }
  for (size_t b = 0; b < pd.beams.size(); ++b) {  // vertices and current position (Melt.cpp:98-183)
    const PathSeg* P = pd.segs.data() + pd.off[b];
    const int nseg = pd.off[b + 1] - pd.off[b];
    int s0 = 0, s1 = 0;
    while (P[s0].time < t0 && (s0 + 1) != nseg) ++s0;
    while (P[s1].time < t1 && (s1 + 1) != nseg) ++s1;
    auto corners = [&](double X, double Y) {
      const int xg = static_cast<int>(std::floor((X - g.xmin) / g.xres));
      const int yg = static_cast<int>(std::floor((Y - g.ymin) / g.yres));
      if (xg < 0 || xg > g.xnum - 1 || yg < 0 || yg > g.ynum - 1) return;
      for (int dx = 0; dx <= 1; ++dx)
        for (int dy = 0; dy <= 1; ++dy) {
          const int i = xg + dx, j = yg + dy;
          if (i < 0 || i > g.xnum - 1 || j < 0 || j > g.ynum - 1) continue;
          out.push_back(pid(i, j));
        }
    };
    for (int s = std::max(s0 - 1, 0); s <= s1; ++s) corners(P[s].x, P[s].y);
    if (P[s1].mode == 0 && t1 < P[nseg - 1].time) {
      double xb, yb, zb, q;
      beam_at(P, s1, t1, pd.rbox, xb, yb, zb, q);
      corners(xb, yb);
    }
  }
}

// ------------------------------------------------------------------ drivers

void Engine::run() {
  const auto t0 = std::chrono::steady_clock::now();
  stats = RunStats{};
  host_grid_evals_ = 0;
  quad_.reset_hints();
  alloc_state();
  if (pb_.mode == Mode::Snapshots) {
    if (pb_.tracking == Tracking::None) snapshots_none();
    else snapshots_tracked();
  } else {
    if (pb_.tracking == Tracking::None) solidify_none();
    else solidify_tracked(pb_.tracking == Tracking::Surface);
  }
  read_counts();
  const auto t1 = std::chrono::steady_clock::now();
  stats.wall_s = std::chrono::duration<double>(t1 - t0).count();
  if (pb_.mode == Mode::Solidification) output("Solidification.Final");  // Main.cpp:104-107
}

// Run::Snapshots_NoTracking (Run.cpp:80-124)
void Engine::snapshots_none() {
  const int pnum = static_cast<int>(pb_.grid.pnum);
  for (size_t s = 0; s < pb_.snapshot_times.size(); ++s) {
    const double t = pb_.snapshot_times[s];
    if (opt.progress) std::printf("Time: %g\n", t);
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[0];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    timed(kCatSum, [&] {
      if (opt.precision == kFP64)
        launch_tsum_grid_f64(slot.f64.get(), slot.n64, geom_, 0, pnum, pb_.mat.T_init, T_.get(), stream_);
      else
        launch_tsum_grid_f32c(slot.f32.get(), slot.n32, geom_, 0, pnum, pb_.mat.T_init, T_.get(), opt.ppt, stream_);
    });
    host_grid_evals_ += static_cast<double>(pnum) * slot.n64;
    ++stats.steps;
    if (st_.oflag) launch_fill_u8(st_.oflag, 1, static_cast<size_t>(pnum), stream_);
    output("Snapshot." + zero_pad(static_cast<int>(s), 2));
  }
}

// One BFS of Melt::neighbor_check waves until the frontier is empty.
void Engine::bfs(const NodeSetDev& ns, bool surface_only, int n_test) {
  DList& test = L_[cTest];
  DList& next = L_[cNext];
  while (n_test > 0) {
    set_count(L_[cCand], 0);
    // next's counter slot is whichever list currently plays "next"
    set_count(next, 0);
    launch_claim_neighbors(test, n_test, geom_, surface_only ? 1 : 0, st_, L_[cCand], L_[cReset], dc_, stream_);
    EvalAction act;
    act.liq = L_[cLiq];
    act.next = next;
    timed(kCatSum, [&] {
      launch_eval_list(ns, geom_, L_[cCand], n_test * (surface_only ? 9 : 18), pb_.mat.T_init, pb_.mat.T_liq, st_,
                       act, dc_, stream_);
    });
    std::swap(test, next);
    read_counts();
    n_test = count(test);
  }
}

// Run::Snapshots_Volume (Run.cpp:126-193); Surface tracking uses it too (Run.cpp:67-69).
void Engine::snapshots_tracked() {
  int n_reset = 0;
  std::vector<int> seeds, liq_host;
  for (size_t s = 0; s < pb_.snapshot_times.size(); ++s) {
    launch_reset(L_[cReset], n_reset, st_, 1, stream_);
    set_count(L_[cReset], 0);
    const double t = pb_.snapshot_times[s];
    if (opt.progress) std::printf("Time: %g\n", t);
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[0];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    const NodeSetDev ns = view(slot);
    // test = liquid points of earlier snapshots (the reference never clears
    // them; duplicates only repeat BFS sources, so they are dropped here)
    read_counts();
    const int n_liq = count(L_[cLiq]);
    liq_host.resize(n_liq);
    if (n_liq)
      THESIS_HIP(hipMemcpy(liq_host.data(), L_[cLiq].v, n_liq * sizeof(int), hipMemcpyDeviceToHost));
    std::sort(liq_host.begin(), liq_host.end());
    liq_host.erase(std::unique(liq_host.begin(), liq_host.end()), liq_host.end());
    const int n_u = static_cast<int>(liq_host.size());
    if (n_u) {
      THESIS_HIP(hipMemcpy(L_[cLiq].v, liq_host.data(), n_u * sizeof(int), hipMemcpyHostToDevice));
      THESIS_HIP(hipMemcpy(L_[cTest].v, liq_host.data(), n_u * sizeof(int), hipMemcpyHostToDevice));
    }
    set_count(L_[cLiq], n_u);
    set_count(L_[cTest], n_u);
    // seeds over the whole path so far
    seeds.clear();
    seeds_between(0.0, t, seeds);
    seeds_.ensure(seeds.size());
    h_seeds_.ensure(seeds.size());
    std::copy(seeds.begin(), seeds.end(), h_seeds_.get());
    if (!seeds.empty())
      THESIS_HIP(hipMemcpyAsync(seeds_.get(), h_seeds_.get(), seeds.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    set_count(L_[cCand], 0);
    launch_claim_seeds(seeds_.get(), static_cast<int>(seeds.size()), st_, L_[cCand], L_[cReset], dc_, stream_);
    EvalAction act;
    act.liq = L_[cLiq];
    act.next = L_[cTest];
    timed(kCatSum, [&] {
      launch_eval_list(ns, geom_, L_[cCand], static_cast<int>(seeds.size()), pb_.mat.T_init, pb_.mat.T_liq, st_, act,
                       dc_, stream_);
    });
    read_counts();
    bfs(ns, false, count(L_[cTest]));
    read_counts();
    n_reset = count(L_[cReset]);
    ++stats.steps;
    output("Snapshot." + zero_pad(static_cast<int>(s), 2));
  }
}

// Run::Solidify_NoTracking (Run.cpp:409-467)
void Engine::solidify_none() {
  const int pnum = static_cast<int>(pb_.grid.pnum);
  {
    std::vector<double> init(static_cast<size_t>(pnum), pb_.mat.T_init);
    THESIS_HIP(hipMemcpy(T_.get(), init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  Tnew_.ensure(static_cast<size_t>(pnum));
  const SolParams sp = sol_params();
  double t = 0.0;
  int itert = 0;
  while (true) {
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[cur_];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    upload_hints();
    timed(kCatSum, [&] {
      if (opt.precision == kFP64)
        launch_tsum_grid_f64(slot.f64.get(), slot.n64, geom_, 0, pnum, pb_.mat.T_init, Tnew_.get(), stream_);
      else
        launch_tsum_grid_f32c(slot.f32.get(), slot.n32, geom_, 0, pnum, pb_.mat.T_init, Tnew_.get(), opt.ppt,
                              stream_);
    });
    host_grid_evals_ += static_cast<double>(pnum) * slot.n64;
    set_count(L_[cSolid], 0);
    set_count(DList{nullptr, counters_.get() + cLiquid, 0}, 0);
    launch_none_apply(Tnew_.get(), pnum, pb_.mat.T_liq, st_, L_[cSolid], counters_.get() + cLiquid, dc_, stream_);
    timed(kCatSol,
          [&] { launch_solidify(L_[cSolid], -1, t, geom_, sp, pb_.do_sol() ? 1 : 0, st_, dc_cat_[kCatSol], stream_); });
    read_counts();
    const int liq = count_of(cLiquid);
    ++stats.steps;
    if (itert && (itert % pb_.out_freq == 0)) output(zero_pad(itert, 7));
    if (t > pb_.scan_end && liq == 0) {
      if (pb_.out_freq != INT_MAX) {
        itert = ((itert / pb_.out_freq) + 1) * pb_.out_freq;
        output(zero_pad(itert, 7));
      }
      break;
    }
    t += pb_.dt;
    ++itert;
  }
}

// Run::Solidify_Volume (Run.cpp:469-573) and Run::Solidify_Surface (Run.cpp:575-709)
void Engine::solidify_tracked(bool surface) {
  const SolParams sp = sol_params();
  const std::string prefix = surface ? "Solidification." : "";
  double t = 0.0;
  int itert = 0;
  int n_liq = 0, n_reset = 0;
  std::vector<int> seeds;
  while (true) {
    // nodes of this step; the other slot keeps the previous step's
    cur_ ^= 1;
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[cur_];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    upload_hints();
    const NodeSetDev cur = view(slot);
    const NodeSetDev last = view(slots_[cur_ ^ 1]);

    launch_reset(L_[cReset], n_reset, st_, 0, stream_);
    // last = liq, reset = liq, liq = {}
    std::swap(L_[cLiq], L_[cLast]);
    const int n_last = n_liq;
    launch_copy_list(L_[cLast], n_last, L_[cReset], dc_, stream_);
    set_count(L_[cLiq], 0);
    set_count(L_[cSolid], 0);
    set_count(L_[cCols], 0);

    // recheck the previous liquid points
    EvalAction act;
    act.liq = L_[cLiq];
    act.solid = L_[cSolid];
    if (surface) act.solid2 = L_[cCols];
    timed(kCatSum,
          [&] { launch_eval_list(cur, geom_, L_[cLast], n_last, pb_.mat.T_init, pb_.mat.T_liq, st_, act, dc_, stream_); });
    if (surface)
      timed(kCatCol, [&] {
        launch_column_expand(L_[cCols], n_last, cur, last, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, L_[cSolid],
                             L_[cReset], dc_cat_[kCatCol], stream_);
      });
    // BFS sources: all previous liquid points (Surface) or the still-liquid ones (Volume)
    launch_copy_list(surface ? L_[cLast] : L_[cLiq], n_last, L_[cTest], dc_, stream_);

    // seeds from the beam's travel during this step
    seeds.clear();
    if (itert && t < pb_.scan_end + pb_.dt) seeds_between(t - pb_.dt, t, seeds);
    if (!seeds.empty()) {
      seeds_.ensure(seeds.size());
      h_seeds_.ensure(seeds.size());
      std::copy(seeds.begin(), seeds.end(), h_seeds_.get());
      THESIS_HIP(hipMemcpyAsync(seeds_.get(), h_seeds_.get(), seeds.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
      set_count(L_[cCand], 0);
      launch_claim_seeds(seeds_.get(), static_cast<int>(seeds.size()), st_, L_[cCand], L_[cReset], dc_, stream_);
      EvalAction sa;
      sa.liq = L_[cLiq];
      sa.next = L_[cTest];
      timed(kCatSum, [&] {
        launch_eval_list(cur, geom_, L_[cCand], static_cast<int>(seeds.size()), pb_.mat.T_init, pb_.mat.T_liq, st_, sa,
                         dc_, stream_);
      });
    }
    read_counts();
    bfs(cur, surface, count(L_[cTest]));
    if (surface) {
      timed(kCatCol, [&] {
        launch_calc_depth(L_[cLiq], -1, cur, last, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, L_[cSolid], L_[cReset],
                          dc_cat_[kCatCol], stream_);
      });
      if (pb_.out.depth && pb_.grid.znum > 1) launch_depth_max(L_[cLiq], -1, geom_, pb_.mat.T_liq, st_, stream_);
    }
    timed(kCatSol,
          [&] { launch_solidify(L_[cSolid], -1, t, geom_, sp, pb_.do_sol() ? 1 : 0, st_, dc_cat_[kCatSol], stream_); });
    read_counts();
    n_liq = count(L_[cLiq]);
    n_reset = count(L_[cReset]);
    ++stats.steps;
    if (opt.progress && itert % 1000 == 0) std::printf("Time step: %d\n", itert);
    if (itert && (itert % pb_.out_freq == 0)) output(prefix + zero_pad(itert, 7));
    if (t > pb_.scan_end && n_liq == 0) {
      if (pb_.out_freq != INT_MAX) {
        itert = ((itert / pb_.out_freq) + 1) * pb_.out_freq;
        output(prefix + zero_pad(itert, 7));
      }
      break;
    }
    t += pb_.dt;
    ++itert;
  }
}

void Engine::solidify_points(double t, const int* pts, int n, const double* T, const double* T_last, double* out7) {
  if (pb_.mode != Mode::Solidification) throw Error(kBadArgument, "solidification quantities need Solidification mode");
  if (!st_.T) alloc_state();
  for (int i = 0; i < n; ++i)
    if (pts[i] < 0 || pts[i] >= pb_.grid.pnum) throw Error(kBadArgument, "point index outside the grid");
  std::vector<int> list(pts, pts + n);
  // place T and T_last at the points, run the per-point kernel, gather
  std::vector<double> Tfull, Tl;
  host_nodes_.clear();
  quad_.build(t, false, host_nodes_);
  upload_hints();
  DevBuf<int> dl;
  dl.ensure(static_cast<size_t>(n));
  THESIS_HIP(hipMemcpy(dl.get(), list.data(), n * sizeof(int), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i) {
    THESIS_HIP(hipMemcpy(st_.T + pts[i], T + i, sizeof(double), hipMemcpyHostToDevice));
    THESIS_HIP(hipMemcpy(st_.T_last + pts[i], T_last + i, sizeof(double), hipMemcpyHostToDevice));
  }
  DList L{dl.get(), counters_.get() + cSolid, n};
  set_count(L, n);
  PointState st = st_;
  // the API always returns every column, whatever the output selection
  DevBuf<double> cols[7];
  double** dst[7] = {&st.tSol, &st.G, &st.Gx, &st.Gy, &st.Gz, &st.dTdt, &st.V};
  for (int c = 0; c < 7; ++c) {
    cols[c].ensure(static_cast<size_t>(pb_.grid.pnum));
    *dst[c] = cols[c].get();
  }
  st.eqFrac = nullptr;
  st.numMelt = nullptr;
  launch_solidify(L, n, t, geom_, sol_params(), 1, st, dc_, stream_);
  sync();
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 7; ++c)
      THESIS_HIP(hipMemcpy(out7 + 7 * i + c, *dst[c] + pts[i], sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace thesis
