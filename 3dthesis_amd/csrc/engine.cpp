#include "engine.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iterator>
#include <thread>

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
  bool pts_pending = false;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device visible");
  const GridSpec& g = pb_.grid;
  geom_ = GridGeom{g.xnum, g.ynum, g.znum, 0, g.xmin, g.ymin, g.zmin, g.xmax, g.ymax, g.zmax, g.sx, g.sy, g.sz};
  if (pb_.custom_points) {
    // point p is a list entry (the reference sets i = p, j = k = 0, Grid.cpp:19-21)
    d_pts_.ensure(pb_.points.size());
    pts_pending = true;
    geom_.nx = static_cast<int>(g.pnum);
    geom_.ny = geom_.nz = 1;
    geom_.pts = d_pts_.get();
  }
  THESIS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // every transfer runs on the engine's own streams: the legacy null stream
  // would hold one of the device's few hardware queues
  if (pts_pending) copy_sync(d_pts_.get(), pb_.points.data(), pb_.points.size() * sizeof(double), hipMemcpyHostToDevice);
  counters_.ensure(32);
  h_counters_.ensure(64);
  h_seq_.ensure(2);
  *h_seq_.get() = 0u;
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
  d_hints_.ensure(pd.beams.size() + 4);  // uploads move whole 16-byte units
  for (Staging& sg : stage_) {
    sg.hints.ensure(pd.beams.size());
    THESIS_HIP(hipEventCreateWithFlags(&sg.done, hipEventDisableTiming));
  }
  copy_sync(d_segs_.get(), pd.segs.data(), pd.segs.size() * sizeof(PathSeg), hipMemcpyHostToDevice);
  copy_sync(d_seg_off_.get(), pd.off.data(), pd.off.size() * sizeof(int), hipMemcpyHostToDevice);
  copy_sync(d_beams_.get(), pd.beams.data(), pd.beams.size() * sizeof(BeamQ), hipMemcpyHostToDevice);
  if (!pd.images.empty()) {
    d_images_.ensure(pd.images.size());
    copy_sync(d_images_.get(), pd.images.data(), pd.images.size() * sizeof(WallImage), hipMemcpyHostToDevice);
  }
}

Engine::~Engine() {
  for (hipStream_t s : {stream_eval_, stream_copy_, stream_solo_, stream_})
    if (s) (void)hipStreamSynchronize(s);
  if (ev_step_) (void)hipEventDestroy(ev_step_);
  if (ev_sol_end_) (void)hipEventDestroy(ev_sol_end_);
  if (cus_reserved_) set_reserved_cus(0);
  for (int b = 0; b < kPipe; ++b)
    for (hipEvent_t e : {w_evald_[b], w_tracked_[b], w_free_[b], w_copied_[b]})
      if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : {stream_eval_, stream_copy_, stream_solo_})
    if (s) (void)hipStreamDestroy(s);
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    for (Staging& sg : stage_)
      if (sg.done) (void)hipEventDestroy(sg.done);
    (void)hipStreamDestroy(stream_);
  }
}

Engine::Staging& Engine::next_staging() {
  stage_cur_ ^= 1;
  Staging& sg = stage_[stage_cur_];
  THESIS_HIP(hipEventSynchronize(sg.done));  // its previous uploads have been consumed
  return sg;
}

void Engine::sync() { THESIS_HIP(hipStreamSynchronize(stream_)); }

void Engine::copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  THESIS_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream_));
  sync();
}

void Engine::sync_all() {
  sync();
  if (stream_eval_) THESIS_HIP(hipStreamSynchronize(stream_eval_));
  if (stream_copy_) THESIS_HIP(hipStreamSynchronize(stream_copy_));
  if (stream_solo_) THESIS_HIP(hipStreamSynchronize(stream_solo_));
}

void Engine::stage_nodes(const NodeF64* nodes, int n, NodeSlot& slot) {
  Staging& sg = next_staging();
  sg.h64.ensure(static_cast<size_t>(n));
  sg.h32.ensure(static_cast<size_t>(n));
  if (n) std::memcpy(sg.h64.get(), nodes, sizeof(NodeF64) * n);
  const size_t n32 = pack_f32c(sg.h64.get(), static_cast<size_t>(n), sg.h32.get());
  upload_staged(sg, sg.h64.dev(), n, sg.h32.dev(), n32, nullptr, quad_.hints(), nullptr, slot);
}

// A node set built ahead on the workers (its FP64 records and their
// FP32-compensated form, in pinned memory), the build's segment hints and,
// with `seeds`, the step's seed candidates.
void Engine::stage_step_nodes(const StepData& sd, NodeSlot& slot, const std::vector<int>* seeds,
                              DeferredUpload* defer) {
  Staging& sg = next_staging();
  if (sd.p64.size() != sd.n64.size()) throw Error(kBadArgument, "step node set without its pinned FP64 copy");
  // the per-step kernels sum in the run's precision: FP32c runs need no FP64 records on the device
  const bool f32 = opt.precision == kFP32c;
  upload_staged(sg, f32 ? nullptr : sd.p64.data(), static_cast<int>(sd.n64.size()), sd.n32.data(), sd.n32.size(),
                defer, sd.hints, seeds, slot);
}

// Uploads run as one kernel on the engine's stream (reading pinned host
// memory), not as copy-engine transfers, so that the kernels behind them start
// at once. The staging set's event marks when its buffers may be rewritten.
void Engine::upload_staged(Staging& sg, const NodeF64* n64, int n, const NodeF32c* n32p, size_t n32,
                           DeferredUpload* defer, const std::vector<int>& hv, const std::vector<int>* seeds,
                           NodeSlot& slot) {
  const size_t ns = seeds ? seeds->size() : 0;
  if (slot.f64.capacity() < static_cast<size_t>(n) || slot.f32.capacity() < n32 ||
      (ns && seeds_.capacity() < ns + 4))
    sync();  // realloc: the buffers are idle
  slot.f64.ensure(static_cast<size_t>(n));
  slot.f32.ensure(n32);
  slot.n64 = n;
  slot.n32 = static_cast<int>(n32);
  sg.hints.ensure(hv.size() + 4);
  sg.ranges.ensure(UploadList::storage_for(4, (n64 ? sizeof(NodeF64) * static_cast<size_t>(n) : 0) +
                                                  sizeof(NodeF32c) * n32 + (hv.size() + ns + 8) * sizeof(int), 0));
  std::copy(hv.begin(), hv.end(), sg.hints.get());
  UploadList ul{reinterpret_cast<UploadRange*>(sg.ranges.get())};
  if (n64) ul.add(n64, slot.f64.get(), sizeof(NodeF64) * static_cast<size_t>(n));
  ul.add(n32p, slot.f32.get(), sizeof(NodeF32c) * n32);
  ul.add(sg.hints.dev(), d_hints_.get(), hv.size() * sizeof(int));
  if (ns) {
    sg.seeds.ensure(ns + 4);
    std::copy(seeds->begin(), seeds->end(), sg.seeds.get());
    seeds_.ensure(ns + 4);
    ul.add(sg.seeds.dev(), seeds_.get(), ns * sizeof(int));
  }
  const UploadRange* r_dev = reinterpret_cast<const UploadRange*>(sg.ranges.dev());
  stats.nodes_built += n;
  if (defer) {  // the caller launches it and records sg.done after it
    *defer = DeferredUpload{r_dev, ul.write_map(r_dev), ul.blocks, sg.done};
    return;
  }
  ul.launch(r_dev, stream_);
  THESIS_HIP(hipEventRecord(sg.done, stream_));
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
    if (o.numMelt && !o.RDF) {  // with RDF the column counts melt times (Grid.h:232)
      numMelt_.ensure(n);
      st_.numMelt = numMelt_.get();
      THESIS_HIP(hipMemsetAsync(st_.numMelt, 0, n * sizeof(uint16_t), stream_));
    }
    stamp_liq_.ensure(n);
    stamp_sol_.ensure(n);
    st_.stamp_liq = stamp_liq_.get();
    st_.stamp_sol = stamp_sol_.get();
    THESIS_HIP(hipMemsetAsync(st_.stamp_liq, 0xff, n * sizeof(int), stream_));  // -1: never written
    THESIS_HIP(hipMemsetAsync(st_.stamp_sol, 0xff, n * sizeof(int), stream_));
    if (pb_.out.RDF) {
      // event logs, drained to the host whenever they pass half full
      const size_t lcap = std::min<size_t>(2 * n + (1u << 16), (size_t)1 << 26);
      nsol_.ensure(n);
      ntm_.ensure(n);
      rdf_ip_.ensure(3 * lcap);
      rdf_dp_.ensure(3 * lcap);
      THESIS_HIP(hipMemsetAsync(nsol_.get(), 0, n * sizeof(int), stream_));
      THESIS_HIP(hipMemsetAsync(ntm_.get(), 0, n * sizeof(int), stream_));
      st_.nsol = nsol_.get();
      st_.ntm = ntm_.get();
      const int cap = static_cast<int>(lcap);
      st_.tm_log = RdfLog{counters_.get() + cTmLog, cap, rdf_ip_.get(), nullptr, rdf_dp_.get(), nullptr};
      st_.sol_log = RdfLog{counters_.get() + cSolLog, cap, rdf_ip_.get() + lcap, rdf_ip_.get() + 2 * lcap,
                           rdf_dp_.get() + lcap, rdf_dp_.get() + 2 * lcap};
      rdf_.tm.assign(n, {});
      rdf_.tl.assign(n, {});
      rdf_.cr.assign(n, {});
      rdf_sol_.assign(n, {});
    }
    const size_t qcap = 2 * n + (1u << 20);
    solq_.ensure(qcap);
    SQ_ = SolQueue{solq_.get(), counters_.get() + cSolQ, static_cast<int>(std::min<size_t>(qcap, INT_MAX))};
    step_t_.clear();
    step_hints_.clear();
    step0_ = 0;
    // per-column depth state of the tracked drivers (custom points never run them)
    const size_t ncol = pb_.custom_points ? 1 : static_cast<size_t>(pb_.grid.xnum) * pb_.grid.ynum;
    depths_.ensure(ncol);
    depth_max_.ensure(ncol);
    st_.depths = depths_.get();
    st_.depth_max = depth_max_.get();
    THESIS_HIP(hipMemsetAsync(st_.depths, 0, ncol * sizeof(int), stream_));
    THESIS_HIP(hipMemsetAsync(st_.depth_max, 0, ncol * sizeof(double), stream_));
  }
  hist_on_ = pb_.out.T_hist && pb_.tracking != Tracking::None && !pb_.custom_points;
  hist_p_.clear();
  hist_s_.clear();
  hist_T_.clear();
  hist_step_t_.clear();
  hist_pending_ = 0;
  if (hist_on_) {
    const size_t hcap = std::min<size_t>(std::max<size_t>(4 * n, (size_t)1 << 20), (size_t)1 << 26);
    hist_ip_.ensure(2 * hcap);
    hist_dp_.ensure(hcap);
    st_.hist_log = RdfLog{counters_.get() + cHistLog, static_cast<int>(hcap), hist_ip_.get(), hist_ip_.get() + hcap,
                          hist_dp_.get(), nullptr};
    launch_set_int(st_.hist_log.n, 0, stream_);
  } else {
    st_.hist_log = RdfLog{};
  }
  if (flags) {
    for (int l = 0; l < kNumLists; ++l) {
      const size_t cap = (l == cReset || l == cReset2) ? 2 * n + 1024 : n + 1024;
      lst_[l].ensure(cap);
      L_[l] = DList{lst_[l].get(), counters_.get() + l, static_cast<int>(std::min<size_t>(cap, INT_MAX))};
    }
  }
  THESIS_HIP(hipMemsetAsync(counters_.get(), 0, 32 * sizeof(int), stream_));
}

void Engine::read_counts() {
  THESIS_HIP(hipMemcpyAsync(h_counters_.get(), counters_.get(), 32 * sizeof(int), hipMemcpyDeviceToHost, stream_));
  sync();
  after_counts();
}

// read_counts for the per-step loop: a kernel writes the counters and a
// sequence number to pinned host memory and the host polls the number (the
// copy and stream synchronisation cost ~15 us more per read), blocking on the
// stream instead after 2 ms.
void Engine::read_counts_polled() {
  const unsigned seq = ++count_seq_;
  launch_export_counts(counters_.get(), 32, h_counters_.dev(), h_seq_.dev(), seq, stream_);
  volatile unsigned* flag = h_seq_.get();
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1; *flag != seq; ++it) {
    __builtin_ia32_pause();
    if ((it & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      sync();  // a long wait: block on the stream (which also reports any device error)
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  after_counts();
}

// Host side of read_counts once the counter copy has landed.
void Engine::after_counts() {
  if (h_counters_.get()[cOverflow]) throw Error(kCapacity, "a device work list overflowed its capacity");
  if (opt.timing) {
    sync_all();  // timed launches may sit on the evaluation stream too
    timer_.collect(stats);
  }
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

void LaunchTimer::begin(int cat, hipStream_t s, int kern) {
  cats_.push_back(cat);
  kerns_.push_back(kern);
  THESIS_HIP(hipEventRecord(get(), s));
}

void LaunchTimer::end(hipStream_t s) { THESIS_HIP(hipEventRecord(get(), s)); }

void LaunchTimer::collect(RunStats& st) {
  for (size_t i = 0; i < cats_.size(); ++i) {
    float ms = 0.f;
    THESIS_HIP(hipEventElapsedTime(&ms, pool_[2 * i], pool_[2 * i + 1]));
    st.cat_ms[cats_[i]] += ms;
    st.cat_launches[cats_[i]] += 1;
    st.window[29 + kerns_[i]] += ms;
    st.window[29 + kNumKern + kerns_[i]] += 1;
  }
  cats_.clear();
  kerns_.clear();
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
  } else if (st_.ntm && pb_.out.numMelt) {  // Grid.h:343-350: the count of recorded melt times
    std::vector<int> c(n);
    THESIS_HIP(hipMemcpyAsync(c.data(), st_.ntm, n * sizeof(int), hipMemcpyDeviceToHost, stream_));
    sync();
    hs.numMelt.assign(c.begin(), c.end());
  }
  if (mp_) {
    hs.mp_w = mp_->width();
    hs.mp_l = mp_->length();
    hs.mp_d = mp_->depth();
  }
  if (st_.oflag) {
    hs.out_flag.resize(n);
    THESIS_HIP(hipMemcpyAsync(hs.out_flag.data(), st_.oflag, n, hipMemcpyDeviceToHost, stream_));
  } else {
    hs.out_flag.clear();
  }
  sync();
}

void Engine::output(const std::string& tag, const int* liq, int n_liq) {
  if (pb_.mode == Mode::Solidification) {
    flush_solidify(true);
    if (sol_async_ran_) {
      // a range solved beside the tracking may have stored G, V and eqFrac
      // over the mark of a point that melted again while it ran; the points
      // liquid now hold their marks in the reference
      if (liq) launch_remark_liquid(liq, n_liq, st_, stream_);
      sol_async_ran_ = false;
    }
  }
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
  sp.images = d_images_.get();
  sp.nimages = static_cast<int>(pd.images.size());
  sp.compress = pd.compress ? 1 : 0;
  sp.gl = gl_table();
  return sp;
}

void Engine::upload_hints() {
  Staging& sg = next_staging();
  const std::vector<int>& h = quad_.hints();
  std::copy(h.begin(), h.end(), sg.hints.get());
  THESIS_HIP(hipMemcpyAsync(d_hints_.get(), sg.hints.get(), h.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
  THESIS_HIP(hipEventRecord(sg.done, stream_));
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

// Record the step's time and per-beam segment hints for its queued solidifications.
void Engine::note_step(int step, double t, const std::vector<int>& hints) {
  if (step_t_.empty()) step0_ = step;
  step_t_.push_back(t);
  step_hints_.insert(step_hints_.end(), hints.begin(), hints.end());
  st_.step = step;
  note_hist_time(step, t);
}

// The T_hist records of the run in the reference's layout: per point in
// ascending order, its evaluations in step order, with the step times.
HistRecords Engine::hist_records() const {
  HistRecords h;
  const size_t n = hist_p_.size();
  std::vector<size_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return hist_p_[a] != hist_p_[b] ? hist_p_[a] < hist_p_[b] : hist_s_[a] < hist_s_[b];
  });
  h.p.resize(n);
  h.t.resize(n);
  h.T.resize(n);
  for (size_t k = 0; k < n; ++k) {
    const size_t i = idx[k];
    const size_t s = static_cast<size_t>(hist_s_[i]);
    if (s >= hist_step_t_.size()) throw Error(kDevice, "a T_hist entry names a step without a recorded time");
    h.p[k] = hist_p_[i];
    h.t[k] = hist_step_t_[s];
    h.T[k] = hist_T_[i];
  }
  return h;
}

void Engine::note_hist_time(int step, double t) {
  if (!hist_on_) return;
  if (static_cast<size_t>(step) >= hist_step_t_.size()) hist_step_t_.resize(static_cast<size_t>(step) + 1, 0.0);
  hist_step_t_[static_cast<size_t>(step)] = t;
}

// Move the T_hist log to the host when it is half full (or always, with force).
void Engine::drain_hist(bool force) {
  if (!hist_on_) return;
  read_counts();
  const int n = count_of(cHistLog);
  if (n > st_.hist_log.cap)
    throw Error(kCapacity, "the T_hist log overflowed its capacity (" + std::to_string(st_.hist_log.cap) + " entries)");
  if (!force && n < st_.hist_log.cap / 4) return;
  if (n > 0) {
    const size_t at = hist_p_.size();
    hist_p_.resize(at + n);
    hist_s_.resize(at + n);
    hist_T_.resize(at + n);
    THESIS_HIP(hipMemcpyAsync(hist_p_.data() + at, st_.hist_log.p, n * sizeof(int), hipMemcpyDeviceToHost, stream_));
    THESIS_HIP(hipMemcpyAsync(hist_s_.data() + at, st_.hist_log.step, n * sizeof(int), hipMemcpyDeviceToHost,
                              stream_));
    THESIS_HIP(hipMemcpyAsync(hist_T_.data() + at, st_.hist_log.a, n * sizeof(double), hipMemcpyDeviceToHost,
                              stream_));
    launch_set_int(st_.hist_log.n, 0, stream_);
    sync();
  }
  hist_pending_ = 0;
}

// Solve the queued solidifications in one launch. Their results are read only
// by the outputs, so the queue is drained when it is large, when the step
// table is long, and before every output.
void Engine::flush_solidify(bool force) {
  if (!SQ_.v) return;
  if (force) read_counts();
  const int queued = h_counters_.get()[cSolQ];  // as of the last read_counts
  const long n = static_cast<long>(pb_.grid.pnum);
  if (!force && queued < std::min<long>(n / 2 + 1, 1 << 20) && step_t_.size() < kStepTable) return;
  // ranges solved on the side stream first: a point's later entry must write after its earlier one
  if (stream_sol_) THESIS_HIP(hipStreamSynchronize(stream_sol_));
  release_cus();
  if (!step_t_.empty() && queued > sol_done_) {
    if (d_step_t_.capacity() < step_t_.size() || d_step_hints_.capacity() < step_hints_.size()) sync();
    d_step_t_.ensure(step_t_.size());
    d_step_hints_.ensure(step_hints_.size());
    THESIS_HIP(hipMemcpyAsync(d_step_t_.get(), step_t_.data(), step_t_.size() * sizeof(double),
                              hipMemcpyHostToDevice, stream_));
    THESIS_HIP(hipMemcpyAsync(d_step_hints_.get(), step_hints_.data(), step_hints_.size() * sizeof(int),
                              hipMemcpyHostToDevice, stream_));
    const SolParams sp = sol_params();
    timed(kCatSol, [&] {
      launch_solidify(SQ_, sol_done_, std::min(queued, SQ_.cap), step0_, static_cast<int>(step_t_.size()),
                      d_step_t_.get(), d_step_hints_.get(), geom_, sp, pb_.do_sol() ? 1 : 0, st_, dc_cat_[kCatSol],
                      stream_);
    }, kKSolidify);
  }
  if (queued > 0) {
    set_count(DList{nullptr, SQ_.n, 0}, 0);
    sync();  // the host tables are rewritten from the next step on
  }
  sol_done_ = 0;
  sol_steps_up_ = 0;
  step_t_.clear();
  step_hints_.clear();
}

// Streams and events of the windowed driver.
void Engine::make_window_streams() {
  if (!stream_eval_) THESIS_HIP(hipStreamCreateWithFlags(&stream_eval_, hipStreamNonBlocking));
  // solidification ranges follow the box sums on the same stream, or run on
  // a stream of their own
  // the windows' replays run there too, after their column walks
  if (!stream_solo_) THESIS_HIP(hipStreamCreateWithFlags(&stream_solo_, hipStreamNonBlocking));
  stream_sol_ = opt.sol_stream == 1 ? stream_solo_ : stream_eval_;
}

// The LDS-staged kernels take the whole device again.
void Engine::release_cus() {
  if (!cus_reserved_) return;
  set_reserved_cus(0);
  cus_reserved_ = false;
}

// Streams and step tables for solidification ranges solved while tracking
// continues.
void Engine::init_async_sol() {
  make_window_streams();
  // steps between drains stay below kStepTable + the windows in flight
  sol_steps_cap_ = kStepTable + (kPipe + 2) * static_cast<size_t>(std::max(opt.window, 1)) + 64;
  const size_t nb = quad_.data().beams.size();
  sol_t_h_.ensure(sol_steps_cap_);
  sol_hints_h_.ensure(sol_steps_cap_ * nb);
  d_step_t_.ensure(sol_steps_cap_);
  d_step_hints_.ensure(sol_steps_cap_ * nb);
  sol_done_ = 0;
  sol_steps_up_ = 0;
}

// Launch the queue entries [sol_done_, queued) on the side stream once
// `ready` (the replay that appended them) has completed.
void Engine::solidify_async(int queued, hipEvent_t ready, int min_batch, int max_blocks) {
  if (queued - sol_done_ < min_batch || queued > SQ_.cap) return;
  const size_t nb = quad_.data().beams.size();
  if (step_t_.size() > sol_steps_cap_) return;  // past the mirror: the next drain solves them
  // new step-table entries: host mirror, then the device table
  const size_t a = sol_steps_up_, b = step_t_.size();
  if (b > a) {
    std::copy(step_t_.begin() + static_cast<long>(a), step_t_.end(), sol_t_h_.get() + a);
    std::copy(step_hints_.begin() + static_cast<long>(a * nb), step_hints_.end(), sol_hints_h_.get() + a * nb);
    THESIS_HIP(hipMemcpyAsync(d_step_t_.get() + a, sol_t_h_.get() + a, (b - a) * sizeof(double),
                              hipMemcpyHostToDevice, stream_sol_));
    THESIS_HIP(hipMemcpyAsync(d_step_hints_.get() + a * nb, sol_hints_h_.get() + a * nb, (b - a) * nb * sizeof(int),
                              hipMemcpyHostToDevice, stream_sol_));
    sol_steps_up_ = b;
  }
  THESIS_HIP(hipStreamWaitEvent(stream_sol_, ready, 0));
  const SolParams sp = sol_params();
  timed_on(kCatSol, stream_sol_, [&] {
    launch_solidify(SQ_, sol_done_, queued, step0_, static_cast<int>(sol_steps_up_), d_step_t_.get(),
                    d_step_hints_.get(), geom_, sp, pb_.do_sol() ? 1 : 0, st_, dc_cat_[kCatSol], stream_sol_,
                    max_blocks);
  }, kKSolidify);
  sol_done_ = queued;
  sol_async_ran_ = true;
}

// ------------------------------------------------------------------ drivers

void Engine::run() {
  const auto t0 = std::chrono::steady_clock::now();
  auto fresh_state = [this] {
    stats = RunStats{};
    host_grid_evals_ = 0;
    sol_done_ = 0;
    sol_steps_up_ = 0;
    sol_async_ran_ = false;
    quad_.reset_hints();
    alloc_state();
    thist_t_.clear();
    thist_T_.clear();
  };
  fresh_state();
  // custom point files always take the untracked drivers (Run.cpp:61, :391)
  const bool untracked = pb_.tracking == Tracking::None || pb_.custom_points;
  // melt-pool statistics follow the step-by-step liquid sets
  const bool per_step = pb_.out.mp_stats && pb_.tracking == Tracking::Surface;
  mp_.reset();
  if (pb_.out.mp_stats && pb_.mode == Mode::Solidification && !untracked && pb_.tracking == Tracking::Surface)
    mp_ = std::make_unique<MeltPoolStats>(pb_);
  if (pb_.mode == Mode::Snapshots) {
    if (untracked) snapshots_none();
    else snapshots_tracked();
  } else {
    if (untracked) solidify_none();
    else if (pb_.tracking == Tracking::Stork) stork();
    else if (opt.window > 0 && !per_step) {
      if (!solidify_windowed(pb_.tracking == Tracking::Surface)) {
        // windows too large for the LDS tracker: the whole run again, one step
        // at a time (the outputs written so far are rewritten unchanged)
        double window[kWindowFields];
        std::copy(std::begin(stats.window), std::end(stats.window), window);
        if (opt.timing) timer_.collect(stats);  // the abandoned windows' launch times are dropped with stats
        fresh_state();
        std::copy(std::begin(window), std::end(window), stats.window);
        stats.window[28] = 1;
        std::fill(stats.window + 29, stats.window + 42, 0.0);  // per-kernel times and windowed evals
        solidify_tracked(pb_.tracking == Tracking::Surface);
      }
    } else {
      solidify_tracked(pb_.tracking == Tracking::Surface);
    }
    flush_solidify(true);
    if (st_.nsol) drain_rdf(true);
  }
  if (hist_on_) {
    sync_all();
    drain_hist(true);
  }
  read_counts();
  const auto t1 = std::chrono::steady_clock::now();
  stats.wall_s = std::chrono::duration<double>(t1 - t0).count();
  if (pb_.mode == Mode::Solidification) {  // Main.cpp:104-112: Stork writes its records instead
    if (pb_.tracking != Tracking::Stork) output("Solidification.Final");
    else if (opt.outputs) write_rrdf(pb_, rr_idx_, rr_ts_, rr_Ts_, opt.full_dir);
  }
  if ((pb_.out.T_hist || st_.nsol) && opt.outputs) {  // Main.cpp:113-118
    std::vector<uint8_t> flags;
    if (st_.oflag) {
      flags.resize(static_cast<size_t>(pb_.grid.pnum));
      THESIS_HIP(hipMemcpyAsync(flags.data(), st_.oflag, flags.size(), hipMemcpyDeviceToHost, stream_));
      sync();
    }
    if (pb_.out.T_hist && hist_on_) write_thist_tracked(pb_, hist_records(), flags, opt.full_dir);
    else if (pb_.out.T_hist) write_thist(pb_, thist_t_, thist_T_, flags, opt.full_dir);
    if (st_.nsol) write_rdf(pb_, rdf_, flags, opt.full_dir);
  }
}

// Run::Stork (Run.cpp:711-1079). The melt pool is found as in Surface
// tracking (recheck, beam-trace seeds, surface neighbour waves) but nothing
// solidifies; each liquid surface point's column is probed for its depth and
// marked liquid, and every grid cell whose number of liquid corners changed
// or is partial, now or at the previous step, is recorded with its corner
// temperatures at t - dt and t, in ascending cell order.
void Engine::stork() {
  const GridSpec& gs = pb_.grid;
  const long long ncell = std::max(0LL, static_cast<long long>(gs.xnum - 1) * (gs.ynum - 1) * (gs.znum - 1));
  const size_t pnum = static_cast<size_t>(gs.pnum);
  DevBuf<uint8_t> c_buf[2], is_liq;
  DevBuf<int> rel;
  DevBuf<double> recs;
  const size_t cells = static_cast<size_t>(std::max(1LL, ncell));
  c_buf[0].ensure(cells);
  c_buf[1].ensure(cells);
  is_liq.ensure(pnum);
  rel.ensure(cells);
  const int chunk = 1 << 16;
  recs.ensure(16 * static_cast<size_t>(chunk));
  launch_fill_u8(c_buf[0].get(), 0, cells, stream_);
  launch_fill_u8(c_buf[1].get(), 0, cells, stream_);
  const DList rel_list{rel.get(), counters_.get() + cRel, static_cast<int>(std::min<size_t>(cells, INT_MAX))};
  rr_idx_.clear();
  rr_ts_.clear();
  rr_Ts_.clear();
  const NodeSlot empty{};
  double t = 0.0;
  int itert = 0, n_liq = 0, n_reset = 0;
  std::vector<int> seeds, rel_h;
  std::vector<double> rec_h;
  std::vector<int> order;
  while (true) {
    uint8_t* c_prev = c_buf[(itert % 2) ? 0 : 1].get();
    uint8_t* c_cur = c_buf[(itert % 2) ? 1 : 0].get();
    launch_fill_u8(c_cur, 0, cells, stream_);
    launch_fill_u8(is_liq.get(), 0, pnum, stream_);
    cur_ ^= 1;
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[cur_];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    const NodeSetDev cur = view(slot);
    const NodeSetDev prev = itert ? view(slots_[cur_ ^ 1]) : view(empty);
    seeds.clear();
    if (itert && t < pb_.scan_end + pb_.dt) seeds_between(t - pb_.dt, t, seeds);
    const int n_seeds = upload_seeds(seeds);

    // the previous step's touched points: T_last = T, unclaimed; the previous
    // liquid points are rechecked and start the new touched list
    std::swap(L_[cLiq], L_[cLast]);
    std::swap(L_[cReset], L_[cReset2]);
    const int n_last = n_liq;
    PointState stp = st_;
    stp.ntm = nullptr;  // no melt times here: they are a Surface tracking record (Run.cpp:607)
    launch_step_prologue(L_[cReset2], std::max(n_reset, n_liq), L_[cLast], L_[cReset], DList{}, L_[cLiq].n,
                         L_[cTest].n, nullptr, nullptr, stp, dc_, stream_);
    EvalAction act;
    act.liq = L_[cLiq];
    act.next = L_[cTest];
    act.zero_depth = 1;
    timed(kCatSum, [&] {
      launch_eval_list(cur, geom_, L_[cLast], n_last, pb_.mat.T_init, pb_.mat.T_liq, st_, act, dc_, stream_);
      launch_seed_eval(seeds_.get(), n_seeds, cur, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, L_[cLiq], L_[cTest],
                       L_[cReset], dc_, stream_);
    });
    bfs(cur, true, n_last + n_seeds);
    const int n_liq_now = count(L_[cLiq]);
    timed(kCatCol, [&] {
      launch_stork_depth(L_[cLiq], n_liq_now, cur, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, is_liq.get(),
                         dc_cat_[kCatCol], stream_);
    });
    if (n_last || n_liq_now) {
      set_count(rel_list, 0);
      launch_stork_cells(geom_, is_liq.get(), c_prev, c_cur, rel_list, dc_, stream_);
      read_counts();
      const int nrel = std::min(count_of(cRel), rel_list.cap);
      if (nrel > 0) {
        rel_h.resize(static_cast<size_t>(nrel));
        THESIS_HIP(hipMemcpyAsync(rel_h.data(), rel.get(), rel_h.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
        sync();
        std::sort(rel_h.begin(), rel_h.end());  // the one-thread record order: ascending cell
        THESIS_HIP(hipMemcpyAsync(rel.get(), rel_h.data(), rel_h.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
        rec_h.resize(16 * static_cast<size_t>(nrel));
        for (int c0 = 0; c0 < nrel; c0 += chunk) {
          const int m = std::min(chunk, nrel - c0);
          timed(kCatCol, [&] {
            launch_stork_records(rel.get() + c0, m, geom_, prev, cur, pb_.mat.T_init, recs.get(), dc_cat_[kCatCol],
                                 stream_);
          });
          THESIS_HIP(hipMemcpyAsync(rec_h.data() + 16 * static_cast<size_t>(c0), recs.get(),
                                    16 * static_cast<size_t>(m) * sizeof(double), hipMemcpyDeviceToHost, stream_));
          sync();
        }
        const int cy = gs.ynum - 1, cz = gs.znum - 1;
        for (int r = 0; r < nrel; ++r) {
          const int c = rel_h[r];
          rr_idx_.push_back(static_cast<uint32_t>((c / cy) / cz));
          rr_idx_.push_back(static_cast<uint32_t>((c / cz) % cy));
          rr_idx_.push_back(static_cast<uint32_t>(c % cz));
          rr_ts_.push_back(t - pb_.dt);
          rr_ts_.push_back(t);
        }
        rr_Ts_.insert(rr_Ts_.end(), rec_h.begin(), rec_h.end());
      }
    }
    n_liq = n_liq_now;
    n_reset = count(L_[cReset]);
    ++stats.steps;
    if (opt.progress && itert % 1000 == 0) std::printf("Time step: %d\n", itert);
    if (t > pb_.scan_end && n_liq == 0) break;  // Util::sim_finish
    t += pb_.dt;
    ++itert;
  }
  read_counts();
}

// Move the device RDF logs to the host: melt times in log order (one per
// point per step, steps in order), solidification events sorted by step.
void Engine::drain_rdf(bool force) {
  read_counts();
  const int ntm = count_of(cTmLog), nsl = count_of(cSolLog);
  const int cap = st_.tm_log.cap;
  if (!force && ntm < cap / 2 && nsl < cap / 2) return;
  if (ntm > 0) {
    std::vector<int> p(static_cast<size_t>(ntm));
    std::vector<double> t(static_cast<size_t>(ntm));
    THESIS_HIP(hipMemcpyAsync(p.data(), st_.tm_log.p, p.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
    THESIS_HIP(hipMemcpyAsync(t.data(), st_.tm_log.a, t.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    for (int i = 0; i < ntm; ++i) rdf_.tm[p[i]].push_back(t[i]);
    launch_set_int(st_.tm_log.n, 0, stream_);
  }
  if (nsl > 0) {
    std::vector<int> p(static_cast<size_t>(nsl)), step(static_cast<size_t>(nsl));
    std::vector<double> a(static_cast<size_t>(nsl)), b(static_cast<size_t>(nsl));
    THESIS_HIP(hipMemcpyAsync(p.data(), st_.sol_log.p, p.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
    THESIS_HIP(hipMemcpyAsync(step.data(), st_.sol_log.step, step.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
    THESIS_HIP(hipMemcpyAsync(a.data(), st_.sol_log.a, a.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    THESIS_HIP(hipMemcpyAsync(b.data(), st_.sol_log.b, b.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    for (int i = 0; i < nsl; ++i) rdf_sol_[p[i]].push_back({step[i], {a[i], b[i]}});
    launch_set_int(st_.sol_log.n, 0, stream_);
  }
  if (force) {
    for (size_t q = 0; q < rdf_sol_.size(); ++q) {
      auto& v = rdf_sol_[q];
      std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      for (const auto& e : v) {
        rdf_.tl[q].push_back(e.second.first);
        rdf_.cr[q].push_back(e.second.second);
      }
      v.clear();
    }
  }
  read_counts();
}

// Grid::add_T_hist / add_t_hist (Grid.h:377-378) for a step in which every
// point was evaluated with set: T of all points at t.
void Engine::record_history(double t) {
  if (!pb_.out.T_hist) return;
  const size_t n = static_cast<size_t>(pb_.grid.pnum);
  const size_t at = thist_T_.size();
  thist_T_.resize(at + n);
  THESIS_HIP(hipMemcpyAsync(thist_T_.data() + at, T_.get(), n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  sync();
  thist_t_.push_back(t);
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
    // Run.cpp:86-89: every point is output (a custom point deck may still name a tracking mode)
    if (st_.oflag) launch_fill_u8(st_.oflag, 1, static_cast<size_t>(pnum), stream_);
    record_history(t);
    output("Snapshot." + zero_pad(static_cast<int>(s), 2));
  }
}

// Seed list upload through the alternating staging buffers.
int Engine::upload_seeds(const std::vector<int>& seeds) {
  if (seeds.empty()) return 0;
  Staging& sg = next_staging();
  sg.seeds.ensure(seeds.size() + 4);
  sg.ranges.ensure(UploadList::storage_for(1, (seeds.size() + 4) * sizeof(int), 0));
  std::copy(seeds.begin(), seeds.end(), sg.seeds.get());
  if (seeds_.capacity() < seeds.size() + 4) sync();
  seeds_.ensure(seeds.size() + 4);
  UploadList ul{reinterpret_cast<UploadRange*>(sg.ranges.get())};
  ul.add(sg.seeds.dev(), seeds_.get(), seeds.size() * sizeof(int));
  ul.launch(reinterpret_cast<const UploadRange*>(sg.ranges.dev()), stream_);
  THESIS_HIP(hipEventRecord(sg.done, stream_));
  return static_cast<int>(seeds.size());
}

// Melt::neighbor_check waves (Melt.cpp:185-269) until the frontier is empty.
// Waves are launched kBfsBatch at a time between host checks; a wave whose
// frontier is empty costs one near-empty launch.
void Engine::bfs(const NodeSetDev& ns, bool surface_only, int hint, bool next_zeroed, int nseeds, int done) {
  constexpr int kBfsBatch = 3;
  // three frontier lists in rotation: a level reads cTest, appends to cNext
  // and clears cCand's counter, which the level after appends to
  auto rotate = [this] {
    const DList t = L_[cTest];
    L_[cTest] = L_[cNext];
    L_[cNext] = L_[cCand];
    L_[cCand] = t;
  };
  hint = std::max(hint, 64);
  if (!next_zeroed) set_count(L_[cNext], 0);
  for (int b = 0; b < done; ++b) rotate();  // levels the caller launched
  int b0 = std::min(done, kBfsBatch - 1);
  while (true) {
    for (int b = b0; b < kBfsBatch; ++b) {
      timed(kCatSum, [&] {
        launch_bfs_level(ns, geom_, L_[cTest], hint, surface_only ? 1 : 0, pb_.mat.T_init, pb_.mat.T_liq, st_,
                         L_[cLiq], L_[cNext], L_[cReset], L_[cCand].n, dc_, stream_, nseeds ? seeds_.get() : nullptr,
                         nseeds);
      });
      nseeds = 0;  // the first level claims them
      rotate();
    }
    b0 = 0;
    read_counts_polled();
    const int n = count(L_[cTest]);
    if (n == 0) break;
    hint = std::max(n, 64);
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
    st_.step = static_cast<int>(s);
    note_hist_time(static_cast<int>(s), t);
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    NodeSlot& slot = slots_[0];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    const NodeSetDev ns = view(slot);
    // the frontier starts from the liquid points of earlier snapshots (the
    // reference never clears them; duplicates only repeat BFS sources, so
    // they are dropped here)
    read_counts();
    const int n_liq = count(L_[cLiq]);
    liq_host.resize(n_liq);
    if (n_liq)
      copy_sync(liq_host.data(), L_[cLiq].v, n_liq * sizeof(int), hipMemcpyDeviceToHost);
    std::sort(liq_host.begin(), liq_host.end());
    liq_host.erase(std::unique(liq_host.begin(), liq_host.end()), liq_host.end());
    const int n_u = static_cast<int>(liq_host.size());
    if (n_u) {
      copy_sync(L_[cLiq].v, liq_host.data(), n_u * sizeof(int), hipMemcpyHostToDevice);
      copy_sync(L_[cTest].v, liq_host.data(), n_u * sizeof(int), hipMemcpyHostToDevice);
    }
    set_count(L_[cLiq], n_u);
    set_count(L_[cTest], n_u);
    // seeds over the whole path so far (Run.cpp:166-178)
    seeds.clear();
    seeds_between(0.0, t, seeds);
    const int ns_count = upload_seeds(seeds);
    timed(kCatSum, [&] {
      launch_seed_eval(seeds_.get(), ns_count, ns, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, L_[cLiq], L_[cTest],
                       L_[cReset], dc_, stream_);
    });
    bfs(ns, false, n_u + ns_count);
    read_counts();
    n_reset = count(L_[cReset]);
    ++stats.steps;
    drain_hist(false);
    output("Snapshot." + zero_pad(static_cast<int>(s), 2));
  }
}

// Run::Solidify_NoTracking (Run.cpp:409-467)
void Engine::solidify_none() {
  const int pnum = static_cast<int>(pb_.grid.pnum);
  {
    std::vector<double> init(static_cast<size_t>(pnum), pb_.mat.T_init);
    copy_sync(T_.get(), init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  Tnew_.ensure(static_cast<size_t>(pnum));
  double t = 0.0;
  int itert = 0;
  while (true) {
    host_nodes_.clear();
    quad_.build(t, false, host_nodes_);
    note_step(itert, t, quad_.hints());
    NodeSlot& slot = slots_[cur_];
    stage_nodes(host_nodes_.data(), static_cast<int>(host_nodes_.size()), slot);
    timed(kCatSum, [&] {
      if (opt.precision == kFP64)
        launch_tsum_grid_f64(slot.f64.get(), slot.n64, geom_, 0, pnum, pb_.mat.T_init, Tnew_.get(), stream_);
      else
        launch_tsum_grid_f32c(slot.f32.get(), slot.n32, geom_, 0, pnum, pb_.mat.T_init, Tnew_.get(), opt.ppt,
                              stream_);
    });
    host_grid_evals_ += static_cast<double>(pnum) * slot.n64;
    set_count(DList{nullptr, counters_.get() + cLiquid, 0}, 0);
    launch_none_apply(Tnew_.get(), pnum, pb_.mat.T_liq, st_, SQ_, counters_.get() + cLiquid, dc_, stream_);
    read_counts();
    const int liq = count_of(cLiquid);
    record_history(t);
    ++stats.steps;
    flush_solidify(false);
    if (st_.nsol) drain_rdf(false);
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

// Run::Solidify_Volume (Run.cpp:469-573) and Run::Solidify_Surface (Run.cpp:575-709).
// All point work of a step runs on the device; the host builds the step's
// node set (overlapping the previous step's tail), uploads it with the seed
// candidates, and synchronises once per batch of BFS waves. Solidification of
// the points a step freezes is deferred to the end of the step: their T and
// T_last are final by then, so the result equals the reference's immediate
// Grid::Solidify calls.
void Engine::solidify_tracked(bool surface) {
  const std::string prefix = surface ? "Solidification." : "";
  double t = 0.0;
  int itert = 0;
  int n_liq = 0, n_reset = 0;
  // node sets and seeds of the coming steps, built in batches on the workers
  // (t accumulates in the stream as here, Run.cpp:566/:704)
  if (!pool_) {
    const unsigned hc = std::thread::hardware_concurrency();
    pool_.reset(new WorkerPool(static_cast<int>(std::min<unsigned>(hc ? hc : 1u, 8u)) - 1));
  }
  // queued solidifications are solved on their own stream in large ranges
  // while the steps go on, unless event logs that the solves append to (RDF,
  // T_hist) are drained as the run goes
  const bool async = opt.async_sol && !st_.nsol && !hist_on_;
  if (async) {
    init_async_sol();
    if (!ev_step_) THESIS_HIP(hipEventCreateWithFlags(&ev_step_, hipEventDisableTiming));
    if (!ev_sol_end_) THESIS_HIP(hipEventCreateWithFlags(&ev_sol_end_, hipEventDisableTiming));
  }
  constexpr int kAheadSteps = 64;
  StepStream steps(
      quad_, pb_.dt, opt.precision == kFP32c,
      [this](int s, double ts, std::vector<int>& out) {
        if (s && ts < pb_.scan_end + pb_.dt) seeds_between(ts - pb_.dt, ts, out);  // Run.cpp:539-541
      },
      *pool_, 3 * kAheadSteps, step_ring_);
  while (true) {
    // nodes of this step; the other slot keeps the previous step's
    cur_ ^= 1;
    if (!steps.has(itert)) {
      steps.drop_before(itert);
      steps.ensure(itert, kAheadSteps);
      steps.ahead(itert + 3 * kAheadSteps);  // the next batches while these steps run
    }
    const StepData& sd = steps.at(itert);
    t = sd.t;
    note_step(itert, t, sd.hints);
    NodeSlot& slot = slots_[cur_];
    DeferredUpload up;  // launched with the step prologue
    stage_step_nodes(sd, slot, &sd.seeds, &up);
    const NodeSetDev cur = view(slot);
    const NodeSetDev last = view(slots_[cur_ ^ 1]);
    const int n_seeds = static_cast<int>(sd.seeds.size());

    // previous step's touched points: T_last = T, unclaimed; last = liq, reset = liq
    std::swap(L_[cLiq], L_[cLast]);
    std::swap(L_[cReset], L_[cReset2]);
    PointState stp = st_;
    if (!surface) stp.ntm = nullptr;  // melt times are recorded by Surface tracking only (Run.cpp:607)
    // Volume steps whose node set fits the LDS recheck and run the first BFS
    // level in one launch; the prologue keeps the previous liquid points claimed
    const bool fused = !surface && recheck_bfs_fits(cur);
    launch_step_prologue(L_[cReset2], std::max(n_reset, n_liq), L_[cLast], L_[cReset],
                         surface ? L_[cTest] : DList{}, L_[cLiq].n, L_[cNext].n, L_[cCols].n,
                         surface ? nullptr : L_[cTest].n, stp, dc_, stream_, t - pb_.dt, pb_.mat.T_liq,
                         fused ? 1 : 0, up.ranges, up.map, up.blocks);
    THESIS_HIP(hipEventRecord(up.done, stream_));
    if (fused) {
      timed(kCatSum, [&] {
        launch_recheck_bfs(cur, geom_, L_[cLast], n_liq, pb_.mat.T_init, pb_.mat.T_liq, st_, L_[cLiq], L_[cNext],
                           L_[cReset], SQ_, L_[cCand].n, n_seeds ? seeds_.get() : nullptr, n_seeds, dc_, stream_);
      });
      bfs(cur, false, n_liq + n_seeds, true, 0, 1);
    } else {
      // recheck the previous liquid points
      EvalAction act;
      act.liq = L_[cLiq];
      act.solid = SQ_;
      if (surface) act.solid2 = L_[cCols];
      else act.next = L_[cTest];  // Volume: the still-liquid points seed the BFS
      timed(kCatSum, [&] {
        launch_eval_list(cur, geom_, L_[cLast], n_liq, pb_.mat.T_init, pb_.mat.T_liq, st_, act, dc_, stream_);
      });
      if (surface)
        timed(kCatCol, [&] {
          launch_column_expand(L_[cCols], n_liq, cur, last, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, SQ_, L_[cReset],
                               dc_cat_[kCatCol], stream_);
        });
      // the seeds from the beam's travel during this step are claimed by the
      // first BFS level; the prologue cleared that level's output list
      bfs(cur, surface, n_liq + n_seeds, true, n_seeds);
    }
    const int n_liq_now = count(L_[cLiq]);
    if (surface) {
      timed(kCatCol, [&] {
        launch_calc_depth(L_[cLiq], n_liq_now, cur, last, geom_, pb_.mat.T_init, pb_.mat.T_liq, st_, SQ_, L_[cReset],
                          dc_cat_[kCatCol], stream_);
      });
      if (pb_.out.depth && pb_.grid.znum > 1) launch_depth_max(L_[cLiq], n_liq_now, geom_, pb_.mat.T_liq, st_, stream_);
      if (mp_) {  // Melt::calc_mp_info after calc_depth_max (Run.cpp:681-684)
        std::vector<int> liq(static_cast<size_t>(std::min(n_liq_now, L_[cLiq].cap)));
        std::vector<int> depths(static_cast<size_t>(pb_.grid.xnum) * pb_.grid.ynum);
        if (!liq.empty())
          THESIS_HIP(hipMemcpyAsync(liq.data(), L_[cLiq].v, liq.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
        THESIS_HIP(hipMemcpyAsync(depths.data(), st_.depths, depths.size() * sizeof(int), hipMemcpyDeviceToHost,
                                  stream_));
        sync();
        mp_->step(t, liq, depths);
      }
    }
    // the BFS ended on a host check, so the liquid list is final; the reset
    // list grows in the tail kernels and is bounded here for the next grid
    n_liq = n_liq_now;
    n_reset = count(L_[cReset]) + n_liq * pb_.grid.znum;
    ++stats.steps;
    if (async) {
      const int queued = h_counters_.get()[cSolQ];  // as of the BFS's last count read
      if (queued >= SQ_.cap / 2 || step_t_.size() >= kStepTable) {
        flush_solidify(true);  // waits for the ranges in flight, solves the rest, empties the queue
      } else if (queued - sol_done_ >= opt.step_sol_batch) {
        THESIS_HIP(hipEventRecord(ev_step_, stream_));
        const int launched = sol_done_;
        solidify_async(queued, ev_step_, opt.step_sol_batch, opt.step_sol_cus);
        if (opt.step_sol_cus > 0 && sol_done_ != launched) {
          THESIS_HIP(hipEventRecord(ev_sol_end_, stream_sol_));
          set_reserved_cus(opt.step_sol_cus);
          cus_reserved_ = true;
        }
      } else if (cus_reserved_) {
        const hipError_t e = hipEventQuery(ev_sol_end_);
        if (e == hipSuccess) release_cus();
        else if (e != hipErrorNotReady) THESIS_HIP(e);
      }
    } else {
      flush_solidify(false);
    }
    if (st_.nsol) drain_rdf(false);
    drain_hist(false);
    if (opt.progress && itert % 1000 == 0) std::printf("Time step: %d\n", itert);
    const int n_liq_out = std::min(n_liq, L_[cLiq].cap);
    if (itert && (itert % pb_.out_freq == 0)) output(prefix + zero_pad(itert, 7), L_[cLiq].v, n_liq_out);
    if (t > pb_.scan_end && n_liq == 0) {
      if (pb_.out_freq != INT_MAX) {
        itert = ((itert / pb_.out_freq) + 1) * pb_.out_freq;
        output(prefix + zero_pad(itert, 7), L_[cLiq].v, n_liq_out);
      }
      break;
    }
    t += pb_.dt;
    ++itert;
  }
  read_counts();
}

// Run::Solidify_Volume / Run::Solidify_Surface (Run.cpp:469-709) in windows of
// up to opt.window steps. For each window the host plans boxes around the
// liquid points, the beams' positions and the seeds of its steps; the device
// evaluates every box point at every step of the window, replays the
// reference's per-step tracking on those values in one workgroup, and applies
// the resulting events to the point state. A point the tracking needs outside
// the boxes ends the window before that step, and the next window (with wider
// boxes) resumes there.
//
// The windows are pipelined. While window w is tracked on the main stream, the
// box sums of window w+1 -- planned on the assumption that w runs to its end --
// run on the second stream into the other buffer set, and w+1's tracker is
// queued behind w's: it checks w's result on the device and skips itself if w
// ended early. The host reads w's result (copied on a third stream) while w+1
// runs; if w ended early, w+1 is planned again from where w stopped. Windows
// are not pipelined across output steps or when queued solidifications are
// due: there the stream drains first. Node sets and seeds of later steps are
// built on worker threads meanwhile.
struct Engine::WinPlan {
  int s0 = 0, W = 0, row0 = 0, nrows = 0, k0 = 0, list_cap = 0, buf = 0;
  bool fits = false;
  bool two_pass = false;  // Surface box sums in two passes (launch_surface_two_pass)
  WinGeom wg{};
  std::vector<int> noff;
  size_t o_noff = 0, o_soff = 0, o_seed = 0, o_fin = 0, o_time = 0;
  double spec = 0.0;
};

bool Engine::solidify_windowed(bool surface) {
  const std::string prefix = surface ? "Solidification." : "";
  const GridSpec& gs = pb_.grid;
  const int nz = gs.znum;
  const int out_freq = pb_.out_freq;
  const bool f32 = opt.precision == kFP32c;
  const bool depth_on = surface && st_.depth != nullptr && nz > 1;
  if (!pool_) {
    const unsigned hc = std::thread::hardware_concurrency();
    pool_.reset(new WorkerPool(static_cast<int>(std::min<unsigned>(hc ? hc : 1u, 8u)) - 1));
  }
  init_async_sol();
  if (!stream_copy_) {
    THESIS_HIP(hipStreamCreateWithFlags(&stream_copy_, hipStreamNonBlocking));
    for (int b = 0; b < kPipe; ++b) {
      THESIS_HIP(hipEventCreateWithFlags(&w_evald_[b], hipEventDisableTiming));
      THESIS_HIP(hipEventCreateWithFlags(&w_tracked_[b], hipEventDisableTiming));
      THESIS_HIP(hipEventCreateWithFlags(&w_free_[b], hipEventDisableTiming));
      THESIS_HIP(hipEventCreateWithFlags(&w_copied_[b], hipEventDisableTiming));
      THESIS_HIP(hipEventRecord(w_evald_[b], stream_eval_));
      THESIS_HIP(hipEventRecord(w_free_[b], stream_));
    }
  }
  if (opt.sol_stream == 2) stream_sol_ = stream_copy_;
  if (opt.sol_stream == 3) stream_sol_ = stream_;
  StepStream steps(
      quad_, pb_.dt, f32,
      [this](int s, double t, std::vector<int>& out) {
        if (s && t < pb_.scan_end + pb_.dt) seeds_between(t - pb_.dt, t, out);  // Run.cpp:539-541
      },
      *pool_, 6 * std::max(1, opt.window) + 64, step_ring_);
  BoxPlanner planner(geom_);
  const size_t lds_limit = opt.tracker_lds ? track_lds_limit() : 0;
  const int wmax = std::max(1, opt.window);
  const PathData& pd = quad_.data();
  const int nbeams = static_cast<int>(pd.beams.size());

  auto cell_of = [&](double x, double y, int& i, int& j) {
    const double fi = gs.xnum > 1 ? std::floor((x - gs.xmin) / gs.xres) : 0.0;
    const double fj = gs.ynum > 1 ? std::floor((y - gs.ymin) / gs.yres) : 0.0;
    if (!(fi >= 0.0 && fi <= gs.xnum - 1 && fj >= 0.0 && fj <= gs.ynum - 1)) return false;
    i = static_cast<int>(fi);
    j = static_cast<int>(fj);
    return true;
  };
  // window length from s0: at most wmax, ending at the next output step
  auto window_len = [&](int s0) {
    int W = wmax;
    if (out_freq != INT_MAX) {
      const long long o = (static_cast<long long>(std::max(s0, 1)) + out_freq - 1) / out_freq * out_freq;
      W = static_cast<int>(std::min<long long>(W, o - s0 + 1));
    }
    return W;
  };

  int n_liq = 0, max_depth = 0, min_k = nz - 1;
  // lateral margin of the boxes around their key columns: widened by a miss
  // outside the boxes, narrowed again by one after kMarginDecay windows
  // without one (a miss only costs a re-planned window)
  constexpr int kMarginDecay = 32;
  const int kMarginBase = opt.box_margin;
  int margin = kMarginBase, zbonus = 0, miss_k = nz - 1, clean = 0;
  // Volume windows that did not fit the LDS tracker; the device-memory
  // tracker's neighbour waves are slower than the per-step kernels
  int outgrown = 0;
  const double switch_steps = pb_.scan_end / pb_.dt / 8.0;
  std::vector<int> liq_host, keys, extra;
  // liquid lists hold at most every grid point; result blocks per result slot
  const size_t lcap = static_cast<size_t>(pb_.grid.pnum);
  const size_t rints = static_cast<size_t>(kTrInts + 4 * wmax);
  for (int q = 0; q <= kPipe; ++q) {
    w_liq_[q].ensure(lcap);
    w_liq_h_[q].ensure(lcap);
  }
  for (int q = 0; q < kPipe; ++q) {
    w_res_[q].ensure(rints);
    w_res_h_[q].ensure(rints);
  }

  // Plan the boxes of the window [s0, s0+W) whose key points are the liquid
  // points known now and the beam path from step path0 on.
  // (i, j) of grid point p = (i ny + j) nz + k, by reciprocal multiplies
  // corrected to the exact quotients (two integer divides per key were a
  // third of the planning time)
  const double r_nyz = 1.0 / (static_cast<double>(gs.ynum) * nz), r_nz = 1.0 / nz;
  const long long nyz = static_cast<long long>(gs.ynum) * nz;
  auto col_of = [&](int p, int& i, int& j) {
    i = static_cast<int>(p * r_nyz);
    if (i * nyz > p) --i;
    else if ((i + 1) * nyz <= p) ++i;
    const int rem = static_cast<int>(p - i * nyz);
    j = static_cast<int>(rem * r_nz);
    if (j * nz > rem) --j;
    else if ((j + 1) * nz <= rem) ++j;
  };
  auto plan = [&](int s0, int W, int path0, int buf, bool shrink) {
    WinPlan P;
    P.s0 = s0;
    P.buf = buf;
    steps.ensure(std::max(path0, 0), s0 + W - std::max(path0, 0));
    while (true) {
      keys.clear();
      keys.reserve(2 * (liq_host.size() + 64));
      for (int p : liq_host) {
        int i, j;
        col_of(p, i, j);
        keys.push_back(i);
        keys.push_back(j);
      }
      keys.insert(keys.end(), extra.begin(), extra.end());
      for (int s = std::max(path0, 0); s < s0 + W; ++s) {
        const StepData& sd = steps.at(s);
        for (int b = 0; b < nbeams; ++b) {
          // the beam's position, formed by the step's builder
          const double xb = sd.beam_xy[2 * static_cast<size_t>(b)], yb = sd.beam_xy[2 * static_cast<size_t>(b) + 1];
          int i, j;
          if (cell_of(xb, yb, i, j)) {
            keys.push_back(i);
            keys.push_back(j);
          }
        }
        if (s >= s0)
          for (int p : sd.seeds) {
            int i, j;
            col_of(p, i, j);
            keys.push_back(i);
            keys.push_back(j);
          }
      }
      if (surface) P.k0 = nz - 1 - std::min(nz - 1, max_depth + opt.depth_margin + zbonus);
      else P.k0 = std::max(0, std::min(min_k, nz - 1) - opt.depth_margin - zbonus);
      P.k0 = std::max(0, std::min(P.k0, miss_k - 1));
      planner.plan(keys, margin, P.k0, P.wg);
      if (surface) {  // surface-layer bitmaps
        P.fits = track_surface_lds_bytes(P.wg.nsurfw) <= lds_limit && P.wg.nsurfw <= surface_max_words();
      } else {  // list capacity: what the LDS left over holds, at most 8192 entries
        const size_t base = track_lds_bytes(P.wg.nslots, P.wg.ncols, 0);
        P.list_cap = base < lds_limit
                         ? static_cast<int>(std::min<size_t>(8192, (lds_limit - base) / (kTrackLists * sizeof(int))))
                         : 0;
        P.fits = P.list_cap >= 256;
      }
      if (P.fits || W == 1 || !shrink) break;
      W = std::max(1, W / 2);
    }
    if (!P.fits && !surface) P.list_cap = std::max(256, std::min(P.wg.nslots, 1 << 16));
    P.W = W;
    P.nrows = W + 1;
    P.row0 = s0 == 0 ? 1 : 0;
    return P;
  };

  // Upload the window's node sets, seeds and finish flags and start its box
  // sums on the evaluation stream, into buffer set P.buf.
  auto stage = [&](WinPlan& P) {
    const int b = P.buf;
    const size_t nsz = f32 ? sizeof(NodeF32c) : sizeof(NodeF64);
    P.noff.assign(static_cast<size_t>(P.nrows) + 1, 0);
    std::vector<int> soff(static_cast<size_t>(P.W) + 1, 0);
    for (int r = 0; r < P.nrows; ++r) {
      const int s = P.s0 - 1 + r;
      const size_t n = s < 0 ? 0 : (f32 ? steps.at(s).n32.size() : steps.at(s).n64.size());
      P.noff[static_cast<size_t>(r) + 1] = P.noff[static_cast<size_t>(r)] + static_cast<int>(n);
    }
    for (int r = 0; r < P.W; ++r)
      soff[static_cast<size_t>(r) + 1] = soff[static_cast<size_t>(r)] + static_cast<int>(steps.at(P.s0 + r).seeds.size());
    auto a256 = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    P.o_noff = a256(static_cast<size_t>(P.noff.back()) * nsz);
    P.o_soff = P.o_noff + a256(P.noff.size() * sizeof(int));
    P.o_seed = P.o_soff + a256(soff.size() * sizeof(int));
    P.o_fin = P.o_seed + a256(static_cast<size_t>(soff.back()) * sizeof(int));
    P.o_time = P.o_fin + a256(static_cast<size_t>(P.W));
    const size_t total = P.o_time + a256((static_cast<size_t>(P.W) + 1) * sizeof(double));
    // Surface windows sum the surface layer, then the listed columns in depth
    // (their evaluations are added when the listing comes back)
    P.two_pass = surface && two_pass_fits(P.wg);
    P.spec = 0.0;
    for (int r = P.row0; r < P.nrows; ++r)
      P.spec += static_cast<double>(P.noff[static_cast<size_t>(r) + 1] - P.noff[static_cast<size_t>(r)]) *
                (P.two_pass ? P.wg.ncols : P.wg.nslots);
    const size_t cells = static_cast<size_t>(P.nrows) * static_cast<size_t>(std::max(P.wg.stride, 128));
    const size_t dcells = static_cast<size_t>(P.nrows) * static_cast<size_t>(std::max(P.wg.ncols, 1));
    const size_t swords = surface ? static_cast<size_t>(P.nrows) * static_cast<size_t>(std::max(surf_stride(P.wg), 2)) : 0;
    const size_t hbytes =
        total - P.o_noff +
        UploadList::storage_for(static_cast<size_t>(P.nrows) + 5, total,
                                static_cast<size_t>(P.nrows) * P.wg.stride + static_cast<size_t>(P.nrows) * P.wg.stride / 8 +
                                    (depth_on ? dcells * sizeof(short) + 8 : 0));
    THESIS_HIP(hipEventSynchronize(w_evald_[b]));  // the pinned staging's last upload is done
    if (w_up_h_[b].capacity() < hbytes || w_up_[b].capacity() < total || w_Tc_[b].capacity() < cells ||
        w_ev_[b].capacity() < cells || w_lq_[b].capacity() < cells / 64 ||
        (depth_on && w_dev_[b].capacity() < dcells + 8) || w_sl_[b].capacity() < swords ||
        (P.two_pass && (w_cols_[b].capacity() < dcells || w_ccnt_[b].capacity() < static_cast<size_t>(P.nrows)))) {
      sync_all();  // buffers about to be reallocated are idle (the set's tracker has finished)
      w_up_h_[b].ensure(hbytes);
      w_up_[b].ensure(total);
      w_Tc_[b].ensure(cells);
      w_lq_[b].ensure(cells / 64);
      w_ev_[b].ensure(cells);
      if (depth_on) w_dev_[b].ensure(dcells + 8);  // the fill rounds up to 16 bytes
      w_sl_[b].ensure(swords);
      if (P.two_pass) {
        w_cols_[b].ensure(dcells);
        w_ccnt_[b].ensure(static_cast<size_t>(P.nrows));
        w_ccnt_h_[b].ensure(static_cast<size_t>(P.nrows));
      }
    }
    // the per-window tables go through the pinned staging buffer; the node
    // sets are gathered by the upload kernel from the steps' own pinned vectors
    uint8_t* const hb = w_up_h_[b].get();
    auto tab = [&](size_t o) { return hb + (o - P.o_noff); };  // table at device offset o
    std::memcpy(tab(P.o_noff), P.noff.data(), P.noff.size() * sizeof(int));
    std::memcpy(tab(P.o_soff), soff.data(), soff.size() * sizeof(int));
    for (int r = 0; r < P.W; ++r) {
      const StepData& sd = steps.at(P.s0 + r);
      std::memcpy(tab(P.o_seed) + sizeof(int) * static_cast<size_t>(soff[static_cast<size_t>(r)]), sd.seeds.data(),
                  sd.seeds.size() * sizeof(int));
      tab(P.o_fin)[r] = sd.t > pb_.scan_end ? 1 : 0;
    }
    // step times s0 .. s0 + W (the last accumulated as the step stream does)
    double* times = reinterpret_cast<double*>(tab(P.o_time));
    for (int r = 0; r < P.W; ++r) times[r] = steps.at(P.s0 + r).t;
    times[P.W] = times[P.W - 1] + pb_.dt;
    hipStream_t s1 = stream_eval_;
    THESIS_HIP(hipStreamWaitEvent(s1, w_free_[b], 0));  // the tracker that last read this set is done
    uint8_t* dup = w_up_[b].get();
    UploadList ul{reinterpret_cast<UploadRange*>(tab(total))};
    for (int r = 0; r < P.nrows; ++r) {
      const int s = P.s0 - 1 + r;
      const int n = P.noff[static_cast<size_t>(r) + 1] - P.noff[static_cast<size_t>(r)];
      if (s < 0 || n == 0) continue;
      const StepData& sd = steps.at(s);
      ul.add(f32 ? static_cast<const void*>(sd.n32.data()) : static_cast<const void*>(sd.p64.data()),
             dup + static_cast<size_t>(P.noff[static_cast<size_t>(r)]) * nsz, static_cast<size_t>(n) * nsz);
    }
    ul.add(w_up_h_[b].dev(), dup + P.o_noff, total - P.o_noff);
    ul.add(reinterpret_cast<const void*>(kUploadZero), w_ev_[b].get(), static_cast<size_t>(P.nrows) * P.wg.stride);
    if (P.two_pass)  // pass B ORs each listed column's liquid bits into zeroed rows
      ul.add(reinterpret_cast<const void*>(kUploadZero), w_lq_[b].get() + static_cast<size_t>(P.row0) * (P.wg.stride / 64),
             static_cast<size_t>(P.nrows - P.row0) * (P.wg.stride / 8));
    if (depth_on) ul.add(reinterpret_cast<const void*>(kUploadOnes), w_dev_[b].get(), dcells * sizeof(short));
    ul.launch(reinterpret_cast<const UploadRange*>(w_up_h_[b].dev() + (total - P.o_noff)), s1);
    if (opt.timing) timer_.begin(kCatSum, s1, kKBox);
    if (P.two_pass)
      launch_surface_two_pass(dup, f32, reinterpret_cast<const int*>(dup + P.o_noff), P.nrows, P.row0, P.wg, geom_,
                              pb_.mat.T_init, pb_.mat.T_liq, w_Tc_[b].get(), w_lq_[b].get(), w_sl_[b].get(),
                              w_cols_[b].get(), w_ccnt_[b].get(), s1);
    else if (f32)
      launch_window_eval_f32c(reinterpret_cast<const NodeF32c*>(dup), reinterpret_cast<const int*>(dup + P.o_noff),
                              P.nrows - P.row0, P.row0, P.wg, geom_, pb_.mat.T_init, pb_.mat.T_liq, w_Tc_[b].get(),
                              w_lq_[b].get(), s1);
    else
      launch_window_eval_f64(reinterpret_cast<const NodeF64*>(dup), reinterpret_cast<const int*>(dup + P.o_noff),
                             P.nrows - P.row0, P.row0, P.wg, geom_, pb_.mat.T_init, pb_.mat.T_liq, w_Tc_[b].get(),
                             w_lq_[b].get(), s1);
    if (opt.timing) timer_.end(s1);
    if (surface && !P.two_pass) launch_surface_bits(P.wg, P.nrows, P.row0, w_lq_[b].get(), w_sl_[b].get(), s1);
    THESIS_HIP(hipEventRecord(w_evald_[b], s1));
  };

  // Track window P on the main stream once its box sums are in, into result
  // slot `slot`; with `prev`, skip unless the window before ran prev_n steps.
  // The replay follows on the copy stream, so the next window's tracker, which
  // reads only tracker outputs (liquid list, column depths, result block),
  // does not wait for it; then the result block and liquid list are copied to
  // the host on that stream.
  auto track = [&](const WinPlan& P, int slot, const int* prev, int prev_n, int liq_in_idx, int n_in, int lst) {
    const int b = P.buf;
    uint8_t* dup = w_up_[b].get();
    THESIS_HIP(hipStreamWaitEvent(stream_, w_evald_[b], 0));
    TrackArgs ta{};
    ta.w = P.wg;
    ta.g = geom_;
    ta.nsteps = P.W;
    ta.surface = surface ? 1 : 0;
    ta.depth_events = depth_on ? 1 : 0;
    ta.lds_state = P.fits ? 1 : 0;
    ta.lq = w_lq_[b].get();
    ta.sl = surface ? w_sl_[b].get() : nullptr;
    ta.surf_stride = surf_stride(P.wg);
    ta.seed_off = reinterpret_cast<const int*>(dup + P.o_soff);
    ta.seeds = reinterpret_cast<const int*>(dup + P.o_seed);
    ta.finish = dup + P.o_fin;
    ta.liq_in = w_liq_[liq_in_idx].get();
    ta.n_liq_in = n_in;
    ta.depths = st_.depths;
    ta.ev = w_ev_[b].get();
    ta.depth_ev = depth_on ? w_dev_[b].get() : nullptr;
    ta.liq_out = w_liq_[lst].get();
    ta.liq_cap = static_cast<int>(lcap);
    ta.prev_result = prev;
    ta.prev_nsteps = prev_n;
    ta.result = w_res_[slot].get();
    ta.step_counts = w_res_[slot].get() + kTrInts;
    ta.list_cap = P.list_cap;
    ta.phase_times = opt.phase_times ? 1 : 0;
    ta.surface_threads = opt.surface_threads;
    ta.surface_rows_max = opt.surface_rows;
    if (surface) {
      const size_t rows = 3 * static_cast<size_t>(P.W + 1) * static_cast<size_t>(std::max(ta.surf_stride, 2));
      const size_t drows = static_cast<size_t>(P.W + 1) * static_cast<size_t>(std::max(P.wg.ncols, 1));
      if (w_srows_[slot].capacity() < rows || w_dstep_[slot].capacity() < drows || w_colmiss_[slot].capacity() < 2) {
        sync_all();  // the last window of this slot may still be walking its columns
        w_srows_[slot].ensure(rows);
        w_dstep_[slot].ensure(drows);
        w_colmiss_[slot].ensure(2);
      }
      ta.srows = w_srows_[slot].get();
      ta.dstep = w_dstep_[slot].get();
      ta.colmiss = w_colmiss_[slot].get();
    }
    if (!P.fits) {
      if (w_gflags_.capacity() < static_cast<size_t>(std::max(P.wg.stride, 128)) ||
          w_gbits_.capacity() < 14 * static_cast<size_t>(std::max(P.wg.nsurfw, 1)) ||
          w_glists_.capacity() < kTrackLists * static_cast<size_t>(std::max(P.list_cap, 1)) ||
          w_gdepth_.capacity() < 2 * ((static_cast<size_t>(P.wg.ncols) + 7) & ~static_cast<size_t>(7)) + 8)
        sync_all();  // the device-memory tracker state of a window in flight is about to move
      w_gflags_.ensure(static_cast<size_t>(std::max(P.wg.stride, 128)));
      w_gdepth_.ensure(2 * ((static_cast<size_t>(P.wg.ncols) + 7) & ~static_cast<size_t>(7)) + 8);
      w_glists_.ensure(kTrackLists * static_cast<size_t>(std::max(P.list_cap, 1)));
      w_gbits_.ensure(14 * static_cast<size_t>(std::max(P.wg.nsurfw, 1)));  // 7 bitmaps + the 48-byte word table
      ta.gflags = w_gflags_.get();
      ta.gdepth = w_gdepth_.get();
      ta.glists = w_glists_.get();
      ta.gbits = w_gbits_.get();
    }
    timed(kCatCol, [&] { launch_track_window(ta, stream_); }, kKTrack);
    THESIS_HIP(hipEventRecord(w_tracked_[slot], stream_));
    THESIS_HIP(hipStreamWaitEvent(stream_copy_, w_tracked_[slot], 0));
    // Surface: the column walks and the window's final result, beside the next tracker
    if (surface) timed_on(kCatCol, stream_copy_, [&] { launch_surface_columns(ta, stream_copy_); }, kKColumns);
    // the tracker's results first: the host plans the window after next from
    // them (liquid lists longer than kLiqPrefix are fetched when seen)
    launch_window_export(w_res_[slot].get(), kTrInts + 4 * P.W, w_liq_[lst].get(), std::min(P.wg.nslots, kLiqPrefix),
                         P.two_pass ? w_ccnt_[b].get() : nullptr, P.two_pass ? P.nrows : 0, w_res_h_[slot].dev(),
                         w_liq_h_[lst].dev(), P.two_pass ? w_ccnt_h_[b].dev() : nullptr, stream_copy_);
    THESIS_HIP(hipEventRecord(w_copied_[slot], stream_copy_));
  };
  // Replay window P (result slot `slot`) on the stream of the solidification
  // ranges, once the host has seen its export complete (so no device-side wait
  // is needed: a wait on an event still pending costs the host ~13 us). Then
  // buffer set P.buf (= slot) is free again.
  auto replay = [&](const WinPlan& P, int slot) {
    const int b = P.buf;
    MeltTimes mt;
    if (surface && pb_.out.RDF) {  // melt times are a Surface tracking record (Run.cpp:607)
      mt.step_time = reinterpret_cast<const double*>(w_up_[b].get() + P.o_time);
      mt.dt = pb_.dt;
      mt.T_liq = pb_.mat.T_liq;
    }
    timed_on(kCatCol, stream_solo_, [&] {
      launch_window_apply(P.wg, geom_, P.W, P.s0, w_ev_[b].get(), w_Tc_[b].get(), pb_.mat.T_liq, w_res_[slot].get(),
                          st_, SQ_, dc_cat_[kCatCol], mt, stream_solo_);
      if (depth_on)
        launch_window_depth(P.wg, geom_, P.W, w_dev_[b].get(), w_Tc_[b].get(), w_res_[slot].get(), pb_.mat.T_liq, st_,
                            stream_solo_);
    }, kKApply);
    THESIS_HIP(hipEventRecord(w_free_[b], stream_solo_));
  };

  const long qthresh = std::min<long>(static_cast<long>(pb_.grid.pnum) / 2 + 1, 1 << 20);
  using clk = std::chrono::steady_clock;
  auto host_ms = [&](int field, clk::time_point& t) {
    const clk::time_point now = clk::now();
    stats.window[field] += std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
  };
  // Windows in flight, oldest first: window number k uses buffer set,
  // result slot and liquid list k % kPipe. Each is planned from the liquid
  // points of the last window processed and the beam path from the oldest
  // window in flight on, so up to `depth` windows are queued on the device
  // while the host waits for the oldest one's result.
  struct Flight {
    WinPlan P;
    int seq;
    int lst;  // its liquid list
  };
  const int depth = std::max(1, std::min(opt.pipeline, kPipe));
  std::deque<Flight> fl;
  int seq = 0;               // number of the next window
  int liq_dev = kPipe;   // device list of the liquid points after the last window processed
  int queued = 0;
  // a liquid list that neither a window in flight nor liq_dev holds: a
  // window's tracker writes its list while windows ahead of it may still end
  // early and send the host back to liq_dev
  auto free_list = [&] {
    for (int q = 0; q <= kPipe; ++q) {
      bool used = q == liq_dev;
      for (const Flight& x : fl) used = used || x.lst == q;
      if (!used) return q;
    }
    throw Error(kDevice, "no free liquid list for a window");
  };
  auto start_fresh = [&](int s0) {
    const int k = seq++;
    WinPlan P = plan(s0, window_len(s0), s0 - 1, k % kPipe, true);
    stage(P);
    const int lst = free_list();
    track(P, k % kPipe, nullptr, 0, liq_dev, n_liq, lst);
    fl.push_back(Flight{std::move(P), k, lst});
  };
  start_fresh(0);
  clk::time_point tk = clk::now();
  while (true) {
    // queue windows behind the newest, each assuming the ones before it run to their ends
    while (static_cast<int>(fl.size()) < depth) {
      const Flight& b = fl.back();
      const int s_next = b.P.s0 + b.P.W;
      const int s_end = s_next - 1;
      const bool output_after = out_freq != INT_MAX && s_end > 0 && s_end % out_freq == 0;
      const bool flush_due = queued >= qthresh / 2 ||
                             step_t_.size() + static_cast<size_t>((depth + 1) * wmax) >= kStepTable ||
                             (hist_on_ && hist_pending_ >= st_.hist_log.cap / 4);
      if (output_after || flush_due) break;  // the pipeline drains to that step first
      const int k = seq;
      WinPlan N = plan(s_next, window_len(s_next), fl.front().P.s0 - 1, k % kPipe, false);
      host_ms(22, tk);
      if (!N.fits) break;  // a window that would not fit its tracker in LDS waits for fresh boxes
      stage(N);
      host_ms(23, tk);
      const int prev_slot = b.seq % kPipe, prev_w = b.P.W;
      const int lst = free_list();
      track(N, k % kPipe, w_res_[prev_slot].get(), prev_w, b.lst, 0, lst);
      host_ms(24, tk);
      ++seq;
      fl.push_back(Flight{std::move(N), k, lst});
    }
    steps.ahead(fl.back().P.s0 + fl.back().P.W + 3 * wmax);  // later node sets, built while the device works
    host_ms(25, tk);
    const Flight f = std::move(fl.front());
    fl.pop_front();
    const WinPlan& cur = f.P;
    const int slot = f.seq % kPipe;
    THESIS_HIP(hipEventSynchronize(w_copied_[slot]));
    host_ms(26, tk);
    replay(cur, slot);

    const int* R = w_res_h_[slot].get();
    const int done = R[kTrDone], status = R[kTrStatus];
    if (status == 3) throw Error(kBadInput, "surface tracking met a liquid point below the surface");
    stats.evals_window += cur.spec;
    if (cur.two_pass) {  // the listed columns' layers below the surface
      const int* cnt = w_ccnt_h_[cur.buf].get();
      const int nd = cur.wg.b[0].nk - 1;
      for (int r = cur.row0; r < cur.nrows; ++r)
        stats.evals_window += static_cast<double>(cnt[r]) * nd *
                              static_cast<double>(cur.noff[static_cast<size_t>(r) + 1] - cur.noff[static_cast<size_t>(r)]);
    }
    for (int r = 1; r <= done; ++r) {
      const int s = cur.s0 + r - 1;
      const StepData& sd = steps.at(s);
      note_step(s, sd.t, sd.hints);
      const int* c = R + kTrInts + 4 * (r - 1);
      queued += c[2];  // the replay queues one solidification per solidified slot
      hist_pending_ += c[0];  // and logs one T_hist entry per evaluated slot
      const double ev_step = static_cast<double>(c[0]) * static_cast<double>(sd.n64.size()) +
                             (s > 0 ? static_cast<double>(c[1]) * static_cast<double>(steps.at(s - 1).n64.size()) : 0.0);
      host_grid_evals_ += ev_step;
      stats.window[41] += ev_step;
      if (opt.progress && s % 1000 == 0) std::printf("Time step: %d\n", s);
    }
    stats.steps += done;
    const WinGeom& wg = cur.wg;
    stats.window[0] += 1;
    stats.window[1] += status == 2 ? 1 : 0;
    stats.window[2] += wg.nslots;
    stats.window[3] += done;
    stats.window[4] += cur.fits ? 1 : 0;
    outgrown += cur.fits ? 0 : 1;
    stats.window[5] += R[kTrLevels];
    for (int q = 0; q < 8; ++q) stats.window[6 + q] += R[kTrTime + q] * 1e-5;  // 10 ns ticks -> ms
    if (surface)
      for (int q = 0; q < 3; ++q) stats.window[42 + q] += R[kTrColTime + q] * 1e-5;
    stats.window[14] += R[kTrScans];
    stats.window[15] = std::max(stats.window[15], static_cast<double>(R[kTrMaxFront]));
    stats.window[16] = std::max(stats.window[16], static_cast<double>(R[kTrMaxLiq]));
    stats.window[17] += wg.nbox;
    stats.window[18] += wg.ncols;
    stats.window[19] += wg.nbox ? wg.b[0].nk : 0;
    stats.window[20] += margin;
    stats.window[21] += zbonus;
    if (status == 2) {  // widen the boxes where the miss was
      const int mi = R[kTrMissI], mj = R[kTrMissJ], mk = R[kTrMissK];
      bool in_xy = false;
      for (int bx = 0; bx < wg.nbox; ++bx) {
        const WinBox& x = wg.b[bx];
        in_xy = in_xy || (mi >= x.i0 && mi < x.i0 + x.ni && mj >= x.j0 && mj < x.j0 + x.nj);
      }
      if (in_xy) {
        zbonus += 2;
        miss_k = std::min(miss_k, mk);
      } else {
        margin = std::min(margin + 2, std::max(gs.xnum, gs.ynum));
      }
      extra.assign({mi, mj});
      clean = 0;
    } else {
      extra.clear();
      if (++clean >= kMarginDecay && margin > kMarginBase) {
        --margin;
        clean = 0;
      }
    }
    if (done > 0) {
      liq_dev = f.lst;
      n_liq = R[kTrNliq];
      if (n_liq > wg.nslots) throw Error(kCapacity, "liquid list larger than its window");
      if (n_liq > kLiqPrefix)  // the rest of a long list
        THESIS_HIP(hipMemcpy(w_liq_h_[f.lst].get() + kLiqPrefix, w_liq_[f.lst].get() + kLiqPrefix,
                             static_cast<size_t>(n_liq - kLiqPrefix) * sizeof(int), hipMemcpyDeviceToHost));
      liq_host.assign(w_liq_h_[f.lst].get(), w_liq_h_[f.lst].get() + n_liq);
      max_depth = R[kTrMaxDepth];
      min_k = R[kTrMinK];
      miss_k = nz - 1;
    }
    const bool completed = status == 0 && done == cur.W;
    if (!completed && !fl.empty()) {  // the windows queued behind it skip themselves on the device
      sync_all();
      fl.clear();
      seq = f.seq + 1;
    }
    if (fl.empty()) {  // drained: counters, due solidifications, event logs, outputs
      if (!surface && opt.volume_switch > 0 && outgrown >= opt.volume_switch && cur.s0 + done < switch_steps &&
          status != 1) {
        sync_all();
        stats.nodes_built += steps.nodes_built();
        return false;  // the caller starts the run again with the per-step driver
      }
      THESIS_HIP(hipStreamSynchronize(stream_solo_));  // the replay
      read_counts();
      flush_solidify(false);
      if (st_.nsol) {
        sync_all();  // solidification ranges append to the RDF log too
        drain_rdf(false);
      }
      drain_hist(false);
      read_counts();  // the queue as the flush left it
      queued = h_counters_.get()[cSolQ];
      const int s_last = cur.s0 + done - 1;
      if (done > 0 && s_last > 0 && out_freq != INT_MAX && s_last % out_freq == 0)
        output(prefix + zero_pad(s_last, 7), w_liq_[liq_dev].get(), n_liq);
      if (status == 1) {
        if (out_freq != INT_MAX)
          output(prefix + zero_pad(((s_last / out_freq) + 1) * out_freq, 7), w_liq_[liq_dev].get(), n_liq);
        break;
      }
    } else if (opt.async_sol) {
      // ranges of at least window_sol_batch entries (a launch costs the host ~25 us)
      solidify_async(queued, w_free_[slot], opt.window_sol_batch);  // runs after the replay
    }
    const int s0 = cur.s0 + done;
    steps.drop_before(std::max(s0 - 1, 0));
    host_ms(27, tk);
    if (fl.empty()) start_fresh(s0);
  }
  sync_all();
  stats.nodes_built += steps.nodes_built();
  return true;
}

void Engine::solidify_points(double t, const int* pts, int n, const double* T, const double* T_last, double* out7) {
  if (pb_.mode != Mode::Solidification) throw Error(kBadArgument, "solidification quantities need Solidification mode");
  if (!st_.T) alloc_state();
  std::vector<SolEntry> q(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    if (pts[i] < 0 || pts[i] >= pb_.grid.pnum) throw Error(kBadArgument, "point index outside the grid");
    q[i] = SolEntry{pts[i], 0, T[i], T_last[i]};
  }
  // hints of a main build at t, as the reference's static hint would hold
  host_nodes_.clear();
  quad_.build(t, false, host_nodes_);
  const std::vector<int>& h = quad_.hints();
  DevBuf<SolEntry> dq;
  DevBuf<double> dt;
  DevBuf<int> dh;
  dq.ensure(static_cast<size_t>(std::max(n, 1)));
  dt.ensure(1);
  dh.ensure(h.size());
  if (n) copy_sync(dq.get(), q.data(), n * sizeof(SolEntry), hipMemcpyHostToDevice);
  copy_sync(dt.get(), &t, sizeof(double), hipMemcpyHostToDevice);
  copy_sync(dh.get(), h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
  DevBuf<int> cnt;
  cnt.ensure(1);
  copy_sync(cnt.get(), &n, sizeof(int), hipMemcpyHostToDevice);
  const SolQueue sq{dq.get(), cnt.get(), n};
  PointState st = st_;  // every column, no deferral stamps
  st.stamp_liq = st.stamp_sol = nullptr;
  st.numMelt = nullptr;
  DevBuf<double> cols[7];
  double** dst[7] = {&st.tSol, &st.G, &st.Gx, &st.Gy, &st.Gz, &st.dTdt, &st.V};
  for (int c = 0; c < 7; ++c) {
    cols[c].ensure(static_cast<size_t>(pb_.grid.pnum));
    *dst[c] = cols[c].get();
  }
  st.eqFrac = nullptr;
  launch_solidify(sq, 0, n, 0, 1, dt.get(), dh.get(), geom_, sol_params(), 1, st, dc_cat_[kCatSol], stream_);
  sync();
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 7; ++c)
      copy_sync(out7 + 7 * i + c, *dst[c] + pts[i], sizeof(double), hipMemcpyDeviceToHost);
}

}  // namespace thesis
