// Host side of the windowed Solidification driver: node sets of upcoming
// steps built ahead on worker threads, and the boxes a window evaluates.
#pragma once

#include <atomic>
#include <condition_variable>
#include <exception>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <algorithm>
#include <vector>

#include "nodes.hpp"
#include "window.hpp"

namespace thesis {

// A fixed set of worker threads; run(n, fn) calls fn(index, worker) for every
// index in [0, n) on the workers and the caller and returns when all are done.
class WorkerPool {
 public:
  explicit WorkerPool(int workers);
  ~WorkerPool();
  WorkerPool(const WorkerPool&) = delete;
  WorkerPool& operator=(const WorkerPool&) = delete;
  int size() const { return static_cast<int>(threads_.size()) + 1; }
  void run(int n, const std::function<void(int, int)>& fn);

 private:
  void loop(int id);
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable wake_, idle_;
  const std::function<void(int, int)>* fn_ = nullptr;
  int n_ = 0, next_ = 0, busy_ = 0;
  long gen_ = 0;
  bool stop_ = false;
};

// Allocator of pinned host memory that kernels address with the host
// pointer itself (checked at each allocation), so that a kernel can gather
// data held in such vectors without a staging copy.
template <class T>
struct PinnedAlloc {
  using value_type = T;
  PinnedAlloc() = default;
  template <class U>
  PinnedAlloc(const PinnedAlloc<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    THESIS_HIP(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    void* d = nullptr;
    THESIS_HIP(hipHostGetDevicePointer(&d, p, 0));
    if (d != p) {
      (void)hipHostFree(p);
      throw Error(kDevice, "pinned host memory is not addressed by kernels at its host address");
    }
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  template <class U>
  bool operator==(const PinnedAlloc<U>&) const { return true; }
  template <class U>
  bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
template <class T>
using PinnedVec = std::vector<T, PinnedAlloc<T>>;

// Everything one timestep of a Solidification run needs from the host.
struct StepData {
  double t = 0.0;
  std::vector<NodeF64> n64;   // the reference's node list at t
  PinnedVec<NodeF32c> n32;    // its FP32-compensated form (FP32c runs)
  PinnedVec<NodeF64> p64;     // a pinned copy of n64, which uploads gather
  std::vector<int> hints;     // per-beam segment after the build (Calc.cpp:104)
  std::vector<double> beam_xy;  // per beam: its (x, y) at t (the window planner's keys)
  std::vector<int> seeds;     // beam-trace candidates of the step (Melt.cpp:20-183)
};

// Steps 0, 1, 2, ... of a run, with t accumulated as the reference does
// (t += dt, Run.cpp:460/:566/:704). A builder thread of its own builds the
// steps in order, ahead of the caller, into a ring of `capacity` steps; each
// batch is split over the worker pool, and a worker's segment search starts
// from the hints of the last step built before the batch, which is at or
// below every step of the batch, so it finds the segment the reference's
// sequential search finds. The caller asks for steps with ensure (which
// waits for them) or ahead (which does not), and releases them with
// drop_before; at() is valid for steps made available by ensure or has.
class StepStream {
 public:
  using SeedFn = std::function<void(int step, double t, std::vector<int>& out)>;
  // `ring` holds the steps; the caller keeps it across runs, so that its
  // pinned node vectors keep their capacity and are not allocated again
  StepStream(const Quadrature& q, double dt, bool f32c, SeedFn seeds, WorkerPool& pool, int capacity,
             std::vector<StepData>& ring);
  ~StepStream();
  StepStream(const StepStream&) = delete;
  StepStream& operator=(const StepStream&) = delete;
  void ensure(int s, int n);  // steps [s, s+n) available
  void ahead(int s);          // build on up to step s (exclusive) without waiting
  const StepData& at(int s) const { return ring_[static_cast<size_t>(s % cap_)]; }
  bool has(int s) const { return s >= first_ && s < built_.load(std::memory_order_acquire); }
  void drop_before(int s);
  double nodes_built() const;

 private:
  void builder();
  std::vector<Quadrature> quads_;  // one per worker
  double dt_;
  bool f32c_;
  SeedFn seeds_;
  WorkerPool& pool_;
  std::vector<StepData>& ring_;
  int cap_;
  std::atomic<int> built_{0};  // steps [0, built_) are built
  int first_ = 0;              // the oldest step the caller keeps
  int target_ = 0;             // the builder builds up to here
  double t_next_ = 0.0;        // time of step built_ (builder only)
  std::vector<int> last_hints_;
  double nodes_built_ = 0.0;
  bool stop_ = false;
  std::exception_ptr err_;     // a failed build, rethrown to the caller
  mutable std::mutex m_;
  std::condition_variable cv_;
  std::thread th_;
};

// Boxes covering every grid column within `margin` cells of a key column,
// from layer k0 to the surface. Nearby key columns share a box (a coarse
// occupancy grid groups them); at most kMaxBoxes boxes result.
class BoxPlanner {
 public:
  explicit BoxPlanner(const GridGeom& g);
  void plan(const std::vector<int>& key_ij, int margin, int k0, WinGeom& out);

 private:
  GridGeom g_;
  struct Rect {
    int i0, i1, j0, j1;
  };
  static constexpr int kShift = 2;  // coarse cells of 4 x 4 columns
  int cx_ = 0, cy_ = 0;
  std::vector<Rect> ext_;  // per coarse cell: its keys' extent (i0 == INT_MAX: none)
  std::vector<int> cells_;  // cells with keys
  std::vector<int> label_;
  std::vector<int> marked_;
  std::vector<int> cover_;
};

}  // namespace thesis
