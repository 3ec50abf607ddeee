// Engine: one parsed problem, its device state and the mode drivers
// (the reference's Run::Simulate and the Grid it drives, Run.cpp:27-709).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "model.hpp"
#include "nodes.hpp"
#include "output.hpp"
#include "track.hpp"
#include "melt_stats.hpp"
#include "window_host.hpp"

namespace thesis {

template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  // Grows by half again at least, so that buffers sized per window settle
  // after a few reallocations (each waits for the device).
  void ensure(size_t n) {
    if (n <= cap_) return;
    const size_t c = cap_ ? std::max(n, cap_ + cap_ / 2) : n;
    release();
    THESIS_HIP(hipMalloc(reinterpret_cast<void**>(&p_), std::max<size_t>(c, 1) * sizeof(T)));
    cap_ = c;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    cap_ = 0;
  }
  T* get() const { return p_; }
  size_t capacity() const { return cap_; }

 private:
  T* p_ = nullptr;
  size_t cap_ = 0;
};

template <class T>
class PinnedBuf {
 public:
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void ensure(size_t n) {
    if (n <= cap_) return;
    const size_t c = cap_ ? std::max(n, cap_ + cap_ / 2) : n;
    release();
    THESIS_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), std::max<size_t>(c, 1) * sizeof(T), hipHostMallocDefault));
    THESIS_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_), p_, 0));
    cap_ = c;
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = d_ = nullptr;
    cap_ = 0;
  }
  T* get() const { return p_; }
  T* dev() const { return d_; }  // the same memory as kernels address it
  size_t capacity() const { return cap_; }

 private:
  T* p_ = nullptr;
  T* d_ = nullptr;
  size_t cap_ = 0;
};

struct Options {
  int precision = kFP64;  // Precision enum
  int ppt = 2;            // points per thread of the FP32c grid kernel
  std::string full_dir;   // non-empty: also write full-precision dumps there
  bool progress = false;  // print the reference's progress lines
  bool outputs = true;    // write the CSV outputs (off: timing runs)
  bool timing = false;    // time every kernel launch per category with HIP events
  bool phase_times = false;  // the window tracker reads its clock at every phase (window stats)
  int window = 40;        // steps per tracked Solidification window (0: one step at a time, per-step kernels)
  bool tracker_lds = true;  // tracker state in LDS when it fits (false: always in device memory)
  // solve queued solidifications on the box-sum stream while tracking
  // continues (1), or only where the driver drains the queue (0)
  int async_sol = 1;
  // a Volume run whose windows do not fit the LDS tracker this many times
  // within the first eighth of its scan steps starts again with the per-step
  // driver (0: never)
  int volume_switch = 24;
  // lateral margin (cells) of the windowed box planner around its key columns
  // before any point outside the boxes widens it
  int box_margin = 5;
  // layers of margin below the deepest liquid column (Surface) or the lowest
  // liquid layer (Volume) before a column walk out of the boxes deepens them
  int depth_margin = 4;
  // Surface tracker workgroup size (64, 256 or 1024; 0: by the number of
  // surface words of the window)
  int surface_threads = 0;
  // Surface tracker: the window's surface rows and seed masks in LDS where
  // they fit (2), the rows only (1), neither (0)
  int surface_rows = 2;
  // tracked windows queued on the device at once (1 .. kPipe): the host plans
  // each from the last window it has processed
  int pipeline = 4;
  // queued solidification ranges on a stream of their own (1), on the
  // box-sum stream (0), the replay stream (2) or the trackers' stream (3)
  int sol_stream = 1;
  // CUs a solidification range takes while the per-step driver runs on (0:
  // the whole device, ranges then wait for the steps' kernels and back)
  int step_sol_cus = 32;
  int step_sol_batch = 1 << 18;  // queue entries that start such a range
  int window_sol_batch = 8192;   // the windowed driver's: entries that start a range
};

constexpr int kPipe = 6;  // most windows in flight
// steps whose times and segment hints the solidification queue may refer to
// between two drains of the queue
constexpr size_t kStepTable = 1 << 16;

// Kernel categories for the per-category timing and evaluation counts.
enum KernelCat { kCatSum = 0, kCatSol = 1, kCatCol = 2, kNumCats = 3 };
// Kernels timed on their own within the categories (opt.timing): the windowed
// box sums, the window tracker, the window replay (with its depth pass), the
// solidification builds, and everything else.
enum KernelId { kKBox = 0, kKTrack = 1, kKApply = 2, kKSolidify = 3, kKColumns = 4, kKOther = 5, kNumKern = 6 };
constexpr int kWindowFields = 45;

struct RunStats {
  double evals = 0;        // point-node evaluations performed
  double wall_s = 0;       // wall time of the mode driver (Run::Simulate scope)
  long steps = 0;          // time steps (Solidification) or snapshots
  double nodes_built = 0;  // host-built nodes
  double cat_evals[kNumCats] = {0, 0, 0};  // evaluations per kernel category
  double cat_ms[kNumCats] = {0, 0, 0};     // summed launch durations (opt.timing)
  long cat_launches[kNumCats] = {0, 0, 0};
  double evals_window = 0;  // evaluations of the windowed box sums (speculative superset)
  // windowed tracking: windows, windows ended by a point outside the boxes,
  // box slots summed over windows, steps, windows tracked in LDS, neighbour
  // waves, the tracker's 8 phase times (ms, device clock), waves that fell
  // back to flag scans, largest frontier, largest liquid set, then summed
  // over windows: boxes, box columns, box layers (first box), margin, z margin,
  // then host time (ms) planning, staging, launching trackers, building
  // node sets ahead, waiting for results, processing results, then 1 if a
  // Volume run left the windows for the per-step driver; then per kernel
  // (KernelId) the summed launch durations (ms, opt.timing) and the launch
  // counts; then the evaluations the windowed steps stand for (each step's
  // evaluated points times its node count, as the reference counts them)
  double window[kWindowFields] = {0};
};

// HIP-event timing of launches by category, collected at synchronisation points.
class LaunchTimer {
 public:
  ~LaunchTimer();
  void begin(int cat, hipStream_t s, int kern = kKOther);
  void end(hipStream_t s);
  void collect(RunStats& st);  // call after the stream has synchronised
 private:
  hipEvent_t get();
  std::vector<hipEvent_t> pool_;
  size_t used_ = 0;
  std::vector<int> cats_, kerns_;
};

class Engine {
 public:
  explicit Engine(Problem pb);
  ~Engine();

  const Problem& problem() const { return pb_; }
  Quadrature& quadrature() { return quad_; }
  const GridGeom& geom() const { return geom_; }
  hipStream_t stream() const { return stream_; }

  Options opt;
  RunStats stats;

  // Run::Simulate for the configured mode; writes every output file.
  void run();

  // T_init + node sum at grid points [p_begin, p_begin+count) into T_dev
  // (device array of count doubles), from an FP64 node set held on the host.
  void eval_grid(const NodeF64* nodes, int n, int precision, int p_begin, int count, double* T_dev);
  // Same for an explicit device list of point indices.
  void eval_list(const NodeF64* nodes, int n, int precision, const int* list_dev, int count, double* T_dev);
  // Device-resident node set (uploaded once) for repeated timed evaluation.
  void upload_nodes(const NodeF64* nodes, int n);
  void eval_grid_resident(int precision, int p_begin, int count, double* T_dev, hipStream_t s);
  int resident_count(int precision) const;
  const NodeF64* resident_f64() const { return slots_[0].f64.get(); }

  // Solidification quantities at the given points for a liquidus crossing
  // between (t - dt, T_last) and (t, T) -- Grid::Solidify without the grid.
  void solidify_points(double t, const int* pts, int n, const double* T, const double* T_last, double* out7);

 private:
  struct NodeSlot {
    DevBuf<NodeF64> f64;
    DevBuf<NodeF32c> f32;
    int n64 = 0, n32 = 0;
  };
  void stage_nodes(const NodeF64* nodes, int n, NodeSlot& slot);
  // An upload built but not yet launched: the caller launches it inside the
  // step prologue (launch_step_prologue) and then records `done`.
  struct DeferredUpload {
    const UploadRange* ranges = nullptr;
    const unsigned short* map = nullptr;
    unsigned blocks = 0;
    hipEvent_t done = nullptr;
  };
  // defer: build the upload into *defer instead of launching it
  void stage_step_nodes(const StepData& sd, NodeSlot& slot, const std::vector<int>* seeds = nullptr,
                        DeferredUpload* defer = nullptr);
  NodeSetDev view(const NodeSlot& slot) const;

  void alloc_state();
  void download_state(HostState& hs);
  // liq[0, n_liq): the device list of the points liquid at this step, for
  // tracked runs whose solidification ranges ran beside the tracking
  void output(const std::string& tag, const int* liq = nullptr, int n_liq = 0);
  void sync();
  void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);  // on stream_, then wait
  void sync_all();  // both streams
  int count_of(int idx);
  void read_counts();
  void read_counts_polled();
  void set_count(const DList& L, int v);
  int count(const DList& L);
  SolParams sol_params();
  void upload_hints();
  void seeds_between(double t0, double t1, std::vector<int>& out) const;

  void snapshots_none();
  void snapshots_tracked();
  void solidify_none();
  void solidify_tracked(bool surface);
  bool solidify_windowed(bool surface);
  void stork();
  // Stork records (Grid.h:111-113): cell index triples, (t - dt, t) and 16 corner temperatures
  std::vector<uint32_t> rr_idx_;
  std::vector<double> rr_ts_, rr_Ts_;
  // Neighbour waves until the frontier is empty; with nseeds the first level
  // also claims the uploaded seeds.
  // done: levels the caller has launched already (their lists rotate in)
  void bfs(const NodeSetDev& ns, bool surface_only, int hint, bool next_zeroed = false, int nseeds = 0, int done = 0);
  int upload_seeds(const std::vector<int>& seeds);

  Problem pb_;
  Quadrature quad_;
  GridGeom geom_{};
  hipStream_t stream_ = nullptr;

  // nodes: two slots (current, previous step)
  NodeSlot slots_[2];
  int cur_ = 0;
  // pinned staging for uploads, two sets used alternately; an event after each
  // set's copies says when the host may overwrite it again
  struct Staging {
    PinnedBuf<NodeF64> h64;
    PinnedBuf<NodeF32c> h32;
    PinnedBuf<int> hints, seeds;
    PinnedBuf<uint8_t> ranges;  // an UploadList's ranges and block map
    hipEvent_t done = nullptr;
  };
  Staging stage_[2];
  int stage_cur_ = 0;
  Staging& next_staging();
  void upload_staged(Staging& sg, const NodeF64* n64, int n, const NodeF32c* n32p, size_t n32, DeferredUpload* defer,
                     const std::vector<int>& hints, const std::vector<int>* seeds, NodeSlot& slot);
  std::vector<NodeF64> host_nodes_;

  // path data on the device for per-point builds
  DevBuf<PathSeg> d_segs_;
  DevBuf<int> d_seg_off_;
  DevBuf<BeamQ> d_beams_;
  DevBuf<WallImage> d_images_;
  DevBuf<double> d_pts_;  // custom points, x y z each
  // T_hist output (Grid.h:100-101): the untracked drivers evaluate every point
  // at every step, so one time list serves all points; thist_T_ is step-major
  std::vector<double> thist_t_, thist_T_;
  void record_history(double t);
  // T_hist with melt-pool tracking on a grid: the device log of (point, step,
  // T) per evaluation with set, drained into hist_p_/hist_s_/hist_T_;
  // hist_step_t_[s] is the time of step (or snapshot) s, hist_pending_ the
  // entries the windowed driver has let the log collect since its last drain
  bool hist_on_ = false;
  DevBuf<int> hist_ip_;
  DevBuf<double> hist_dp_;
  std::vector<int> hist_p_, hist_s_;
  std::vector<double> hist_T_, hist_step_t_;
  long hist_pending_ = 0;
  void drain_hist(bool force);
  void note_hist_time(int step, double t);
  HistRecords hist_records() const;
  // RDF: per-point counters and event logs on the device, events on the host
  DevBuf<int> nsol_, ntm_, rdf_ip_;
  DevBuf<double> rdf_dp_;
  RdfEvents rdf_;
  std::vector<std::vector<std::pair<int, std::pair<double, double>>>> rdf_sol_;  // (step, (tl, cr)) per point
  void drain_rdf(bool force);
  // MP_Stats on the host, fed after every Surface step
  std::unique_ptr<MeltPoolStats> mp_;
  DevBuf<int> d_hints_;

  // per-point state
  DevBuf<double> T_, T_last_, tSol_, G_, V_, Gx_, Gy_, Gz_, dTdt_, eqFrac_, depth_, depth_max_, Tnew_;
  DevBuf<uint32_t> flag_;
  DevBuf<uint8_t> oflag_;
  DevBuf<uint16_t> numMelt_;
  DevBuf<int> depths_, stamp_liq_, stamp_sol_;
  PointState st_{};

  // deferred solidification: queue, and each queued step's time and hints
  DevBuf<SolEntry> solq_;
  SolQueue SQ_{};
  std::vector<double> step_t_;
  std::vector<int> step_hints_;
  int step0_ = 0;  // step of step_t_[0]
  DevBuf<double> d_step_t_;
  DevBuf<int> d_step_hints_;
  void note_step(int step, double t, const std::vector<int>& hints);
  void flush_solidify(bool force);

  // lists and counters
  // list counters occupy ints [0, kNumLists); flags follow; three u64
  // evaluation counters start at int 16
  enum { cLiq, cLast, cReset, cReset2, cTest, cNext, cCand, cSolid, cCols, kNumLists, cSolQ = 10, cOverflow = 12,
         cLiquid = 13, cTmLog = 14, cSolLog = 15, cRel = 11, cHistLog = 22 };
  DevBuf<int> counters_;
  PinnedBuf<int> h_counters_;
  PinnedBuf<unsigned> h_seq_;  // the counter export's sequence number (read_counts_polled)
  unsigned count_seq_ = 0;
  DevBuf<int> lst_[kNumLists];
  DList L_[kNumLists];
  DevBuf<int> seeds_;
  DevCounters dc_{};            // summation kernels
  DevCounters dc_cat_[kNumCats];
  LaunchTimer timer_;
  double host_grid_evals_ = 0;  // grid sums, counted on the host

  template <class F>
  void timed(int cat, F&& launch, int kern = kKOther) {
    timed_on(cat, stream_, launch, kern);
  }
  template <class F>
  void timed_on(int cat, hipStream_t s, F&& launch, int kern = kKOther) {
    if (opt.timing) timer_.begin(cat, s, kern);
    launch();
    if (opt.timing) timer_.end(s);
  }
  // Solidification queue entries solved while tracking continues: ranges of
  // the append-only queue go to their own stream as windows finish replaying
  // (outputs and the queue reset wait for it). Their step times and hints
  // reach the device through an append-only pinned mirror.
  hipStream_t stream_sol_ = nullptr;  // the box-sum stream, or stream_solo_
  hipStream_t stream_solo_ = nullptr;  // owned; used with opt.sol_stream
  void make_window_streams();
  int sol_done_ = 0;         // queue entries launched so far
  bool sol_async_ran_ = false;      // a range was launched beside the tracking since the last output
  hipEvent_t ev_step_ = nullptr;    // end of a step of the per-step driver, for its ranges
  hipEvent_t ev_sol_end_ = nullptr; // end of the per-step driver's latest range
  bool cus_reserved_ = false;       // the LDS-staged kernels leave that range's CUs free
  size_t sol_steps_up_ = 0;  // step-table entries copied to the device so far
  size_t sol_steps_cap_ = 0;
  PinnedBuf<double> sol_t_h_;
  PinnedBuf<int> sol_hints_h_;
  void solidify_async(int queued, hipEvent_t ready, int min_batch, int max_blocks = 0);
  void init_async_sol();
  void release_cus();
  void gather_evals();
  void after_counts();

  // windowed tracking: worker threads, a second stream for the box sums, two
  // buffer sets (one window tracked while the next is evaluated)
  struct WinPlan;
  std::unique_ptr<WorkerPool> pool_;
  std::vector<StepData> step_ring_;  // the step streams' ring, kept across runs
  hipStream_t stream_eval_ = nullptr;
  hipStream_t stream_copy_ = nullptr;            // replays, then result copies to the host
  hipEvent_t w_evald_[kPipe] = {};    // box sums of the set done
  hipEvent_t w_tracked_[kPipe] = {};  // per result slot: tracker done
  hipEvent_t w_free_[kPipe] = {};     // tracker and replay of the set done
  hipEvent_t w_copied_[kPipe] = {};   // per result slot: result block and liquid list on the host
  static constexpr int kLiqPrefix = 8192;         // liquid-list entries copied with the result block
  DevBuf<uint8_t> w_up_[kPipe];                     // uploaded nodes, offsets, seeds, finish flags
  PinnedBuf<uint8_t> w_up_h_[kPipe];
  DevBuf<double> w_Tc_[kPipe];
  DevBuf<unsigned long long> w_lq_[kPipe];
  DevBuf<unsigned long long> w_sl_[kPipe];          // Surface: the surface layer's liquid words per cache row
  DevBuf<int> w_cols_[kPipe], w_ccnt_[kPipe];           // Surface two-pass sums: listed columns per cache row, their counts
  PinnedBuf<int> w_ccnt_h_[kPipe];
  DevBuf<uint8_t> w_ev_[kPipe], w_gflags_;
  DevBuf<short> w_dev_[kPipe], w_gdepth_;
  // liquid lists: one per window in flight and the list of the last window
  // processed, which a window started afresh reads
  DevBuf<int> w_liq_[kPipe + 1], w_res_[kPipe], w_glists_;
  DevBuf<unsigned long long> w_gbits_;
  // Surface, per result slot: the tracker's surface words per step, column
  // depths per step and the column walks' miss and progress words
  DevBuf<unsigned long long> w_srows_[kPipe], w_colmiss_[kPipe];
  DevBuf<short> w_dstep_[kPipe];
  PinnedBuf<int> w_res_h_[kPipe], w_liq_h_[kPipe + 1];
};

}  // namespace thesis
