#include "window_host.hpp"

#include <algorithm>
#include <climits>

namespace thesis {

// ------------------------------------------------------------------ workers

WorkerPool::WorkerPool(int workers) {
  for (int i = 0; i < workers; ++i) threads_.emplace_back([this, i] { loop(i + 1); });
}

WorkerPool::~WorkerPool() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  wake_.notify_all();
  for (std::thread& t : threads_) t.join();
}

void WorkerPool::loop(int id) {
  long seen = 0;
  while (true) {
    std::unique_lock<std::mutex> lk(m_);
    wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
    if (stop_) return;
    seen = gen_;
    ++busy_;
    while (next_ < n_) {
      const int i = next_++;
      const std::function<void(int, int)>* fn = fn_;
      lk.unlock();
      (*fn)(i, id);
      lk.lock();
    }
    if (--busy_ == 0) idle_.notify_all();
  }
}

void WorkerPool::run(int n, const std::function<void(int, int)>& fn) {
  if (n <= 0) return;
  std::unique_lock<std::mutex> lk(m_);
  fn_ = &fn;
  n_ = n;
  next_ = 0;
  ++gen_;
  wake_.notify_all();
  ++busy_;
  while (next_ < n_) {
    const int i = next_++;
    lk.unlock();
    fn(i, 0);
    lk.lock();
  }
  --busy_;
  idle_.wait(lk, [&] { return busy_ == 0; });
  fn_ = nullptr;
  n_ = 0;
}

// ------------------------------------------------------------------ steps

StepStream::StepStream(const Quadrature& q, double dt, bool f32c, SeedFn seeds, WorkerPool& pool, int capacity,
                       std::vector<StepData>& ring)
    : dt_(dt), f32c_(f32c), seeds_(std::move(seeds)), pool_(pool), ring_(ring), cap_(std::max(capacity, 1)) {
  quads_.assign(static_cast<size_t>(pool.size()), q);
  if (ring_.size() < static_cast<size_t>(cap_)) ring_.resize(static_cast<size_t>(cap_));
  last_hints_.assign(quads_[0].data().beams.size(), 1);
  th_ = std::thread([this] { builder(); });
}

StepStream::~StepStream() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  cv_.notify_all();
  th_.join();
}

void StepStream::drop_before(int s) {
  {
    std::lock_guard<std::mutex> lk(m_);
    first_ = std::max(first_, s);
  }
  cv_.notify_all();
}

void StepStream::ahead(int s) {
  {
    std::lock_guard<std::mutex> lk(m_);
    target_ = std::max(target_, std::min(s, first_ + cap_));
  }
  cv_.notify_all();
}

void StepStream::ensure(int s, int n) {
  std::unique_lock<std::mutex> lk(m_);
  if (s < first_) throw Error(kBadArgument, "step stream asked for a dropped step");
  if (s + n > first_ + cap_) throw Error(kCapacity, "step stream asked for more steps than its ring holds");
  target_ = std::max(target_, s + n);
  cv_.notify_all();
  cv_.wait(lk, [&] { return err_ || built_.load(std::memory_order_relaxed) >= s + n; });
  if (err_) std::rethrow_exception(err_);
}

double StepStream::nodes_built() const {
  std::lock_guard<std::mutex> lk(m_);
  return nodes_built_;
}

void StepStream::builder() {
  constexpr int kBatch = 64;  // steps per batch at most, so that a waiting caller sees steps soon
  while (true) {
    int base, cnt;
    {
      std::unique_lock<std::mutex> lk(m_);
      cv_.wait(lk, [&] {
        const int b = built_.load(std::memory_order_relaxed);
        return stop_ || (!err_ && b < target_ && b < first_ + cap_);
      });
      if (stop_) return;
      base = built_.load(std::memory_order_relaxed);
      cnt = std::min({target_, first_ + cap_, base + kBatch}) - base;
    }
    try {
      for (int i = 0; i < cnt; ++i) {
        StepData& sd = ring_[static_cast<size_t>((base + i) % cap_)];
        sd.t = t_next_;
        t_next_ += dt_;
      }
      // contiguous chunks per worker, each starting from the hints of the last step built
      const int chunks = std::min(pool_.size(), cnt);
      const std::vector<int> h0 = last_hints_;
      std::exception_ptr werr;
      std::mutex werr_m;
      pool_.run(chunks, [&](int c, int wid) {
        try {
        Quadrature& qd = quads_[static_cast<size_t>(wid)];
        qd.hints() = h0;
        const int a = base + static_cast<int>(static_cast<long long>(cnt) * c / chunks);
        const int b = base + static_cast<int>(static_cast<long long>(cnt) * (c + 1) / chunks);
        for (int st = a; st < b; ++st) {
          StepData& sd = ring_[static_cast<size_t>(st % cap_)];
          sd.n64.clear();
          qd.build(sd.t, false, sd.n64);
          sd.hints = qd.hints();
          {
            const PathData& pd = qd.data();
            const size_t nb = pd.beams.size();
            sd.beam_xy.resize(2 * nb);
            for (size_t b = 0; b < nb; ++b) {
              double xb, yb, zb, q;
              beam_at(pd.segs.data() + pd.off[b], sd.hints[b], sd.t, pd.rbox, xb, yb, zb, q);
              sd.beam_xy[2 * b] = xb;
              sd.beam_xy[2 * b + 1] = yb;
            }
          }
          sd.p64.assign(sd.n64.begin(), sd.n64.end());
          if (f32c_) {
            sd.n32.resize(sd.n64.size());
            sd.n32.resize(pack_f32c(sd.n64.data(), sd.n64.size(), sd.n32.data()));
          } else {
            sd.n32.clear();
          }
          sd.seeds.clear();
          seeds_(st, sd.t, sd.seeds);
        }
        } catch (...) {  // a worker must not throw: the batch fails on this thread instead
          std::lock_guard<std::mutex> lk(werr_m);
          if (!werr) werr = std::current_exception();
        }
      });
      if (werr) std::rethrow_exception(werr);
      last_hints_ = ring_[static_cast<size_t>((base + cnt - 1) % cap_)].hints;
      double nb = 0.0;
      for (int i = 0; i < cnt; ++i) nb += static_cast<double>(ring_[static_cast<size_t>((base + i) % cap_)].n64.size());
      {
        std::lock_guard<std::mutex> lk(m_);
        nodes_built_ += nb;
        built_.store(base + cnt, std::memory_order_release);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(m_);
      err_ = std::current_exception();
    }
    cv_.notify_all();
  }
}

// ------------------------------------------------------------------ boxes

BoxPlanner::BoxPlanner(const GridGeom& g) : g_(g) {
  cx_ = ((g.nx - 1) >> kShift) + 1;
  cy_ = ((g.ny - 1) >> kShift) + 1;
  label_.assign(static_cast<size_t>(cx_) * cy_, 0);
  ext_.assign(static_cast<size_t>(cx_) * cy_, Rect{INT_MAX, INT_MIN, INT_MAX, INT_MIN});
}

void BoxPlanner::plan(const std::vector<int>& key_ij, int margin, int k0, WinGeom& out) {
  out = WinGeom{};
  const size_t nkey = key_ij.size() / 2;
  if (nkey == 0) return;
  // the keys' extents per coarse cell: the squares below are formed per cell
  // of keys (a pool's points share few cells), not per key
  cells_.clear();
  for (size_t q = 0; q < nkey; ++q) {
    const int i = key_ij[2 * q], j = key_ij[2 * q + 1];
    const int c = (std::min(std::max(i, 0), g_.nx - 1) >> kShift) * cy_ + (std::min(std::max(j, 0), g_.ny - 1) >> kShift);
    Rect& r = ext_[static_cast<size_t>(c)];
    if (r.i0 == INT_MAX) cells_.push_back(c);
    r.i0 = std::min(r.i0, i);
    r.i1 = std::max(r.i1, i);
    r.j0 = std::min(r.j0, j);
    r.j1 = std::max(r.j1, j);
  }
  // mark coarse cells touched by a cell's square (its keys' extent widened by
  // the margin): each square adds its corners to a 2-D difference array over
  // the squares' bounding box, whose prefix sums are then the number of
  // squares covering each cell
  marked_.clear();
  int bi0 = INT_MAX, bi1 = -1, bj0 = INT_MAX, bj1 = -1;
  auto square = [&](int c, int& ci0, int& ci1, int& cj0, int& cj1) {
    const Rect& r = ext_[static_cast<size_t>(c)];
    ci0 = std::max(0, r.i0 - margin) >> kShift;
    ci1 = std::max(0, std::min(g_.nx - 1, r.i1 + margin)) >> kShift;
    cj0 = std::max(0, r.j0 - margin) >> kShift;
    cj1 = std::max(0, std::min(g_.ny - 1, r.j1 + margin)) >> kShift;
  };
  for (int c : cells_) {
    int ci0, ci1, cj0, cj1;
    square(c, ci0, ci1, cj0, cj1);
    bi0 = std::min(bi0, ci0);
    bi1 = std::max(bi1, ci1);
    bj0 = std::min(bj0, cj0);
    bj1 = std::max(bj1, cj1);
  }
  if (bi1 >= 0) {
    const int bh = bi1 - bi0 + 2, bw = bj1 - bj0 + 2;
    cover_.assign(static_cast<size_t>(bh) * bw, 0);
    auto at = [&](int ci, int cj) -> int& { return cover_[static_cast<size_t>(ci - bi0) * bw + (cj - bj0)]; };
    for (int c : cells_) {
      int ci0, ci1, cj0, cj1;
      square(c, ci0, ci1, cj0, cj1);
      ++at(ci0, cj0);
      --at(ci0, cj1 + 1);
      --at(ci1 + 1, cj0);
      ++at(ci1 + 1, cj1 + 1);
    }
    for (int ci = bi0; ci <= bi1; ++ci)
      for (int cj = bj0; cj <= bj1; ++cj) {
        int v = at(ci, cj);
        if (ci > bi0) v += at(ci - 1, cj);
        if (cj > bj0) v += at(ci, cj - 1);
        if (ci > bi0 && cj > bj0) v -= at(ci - 1, cj - 1);
        at(ci, cj) = v;
        if (v > 0) {
          label_[static_cast<size_t>(ci) * cy_ + cj] = -1;
          marked_.push_back(ci * cy_ + cj);
        }
      }
  }
  // connected components (8-neighbour) of the marked cells
  std::vector<Rect> comp;
  std::vector<int> stack;
  for (int m : marked_) {
    if (label_[static_cast<size_t>(m)] != -1) continue;
    comp.push_back(Rect{INT_MAX, INT_MIN, INT_MAX, INT_MIN});
    const int id = static_cast<int>(comp.size());
    label_[static_cast<size_t>(m)] = id;
    stack.assign(1, m);
    while (!stack.empty()) {
      const int c = stack.back();
      stack.pop_back();
      const int ci = c / cy_, cj = c % cy_;
      for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj) {
          const int ni = ci + di, nj = cj + dj;
          if (ni < 0 || ni >= cx_ || nj < 0 || nj >= cy_) continue;
          int& l = label_[static_cast<size_t>(ni) * cy_ + nj];
          if (l == -1) {
            l = id;
            stack.push_back(ni * cy_ + nj);
          }
        }
    }
  }
  // exact extent of each component's key squares
  for (int c : cells_) {
    Rect& e = ext_[static_cast<size_t>(c)];
    Rect& r = comp[static_cast<size_t>(label_[static_cast<size_t>(c)] - 1)];
    r.i0 = std::min(r.i0, std::max(0, e.i0 - margin));
    r.i1 = std::max(r.i1, std::min(g_.nx - 1, e.i1 + margin));
    r.j0 = std::min(r.j0, std::max(0, e.j0 - margin));
    r.j1 = std::max(r.j1, std::min(g_.ny - 1, e.j1 + margin));
    e = Rect{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  }
  for (int m : marked_) label_[static_cast<size_t>(m)] = 0;
  // merge rectangles that overlap or touch (a neighbour of a box point is then
  // either in the same box or in none), then the cheapest pairs down to kMaxBoxes
  auto area = [](const Rect& r) { return static_cast<long long>(r.i1 - r.i0 + 1) * (r.j1 - r.j0 + 1); };
  auto uni = [](const Rect& a, const Rect& b) {
    return Rect{std::min(a.i0, b.i0), std::max(a.i1, b.i1), std::min(a.j0, b.j0), std::max(a.j1, b.j1)};
  };
  bool changed = true;
  while (changed) {
    changed = false;
    for (size_t a = 0; a < comp.size() && !changed; ++a)
      for (size_t b = a + 1; b < comp.size() && !changed; ++b) {
        const Rect& x = comp[a];
        const Rect& y = comp[b];
        if (x.i0 <= y.i1 + 1 && y.i0 <= x.i1 + 1 && x.j0 <= y.j1 + 1 && y.j0 <= x.j1 + 1) {
          comp[a] = uni(x, y);
          comp.erase(comp.begin() + static_cast<long>(b));
          changed = true;
        }
      }
    if (!changed && comp.size() > static_cast<size_t>(kMaxBoxes)) {
      size_t ba = 0, bb = 1;
      long long best = LLONG_MAX;
      for (size_t a = 0; a < comp.size(); ++a)
        for (size_t b = a + 1; b < comp.size(); ++b) {
          const long long cost = area(uni(comp[a], comp[b])) - area(comp[a]) - area(comp[b]);
          if (cost < best) {
            best = cost;
            ba = a;
            bb = b;
          }
        }
      comp[ba] = uni(comp[ba], comp[bb]);
      comp.erase(comp.begin() + static_cast<long>(bb));
      changed = true;
    }
  }
  const int nk = g_.nz - k0;
  int off = 0, coff = 0, woff = 0;
  for (const Rect& r : comp) {
    WinBox& b = out.b[out.nbox++];
    b.i0 = r.i0;
    b.j0 = r.j0;
    b.k0 = k0;
    b.ni = r.i1 - r.i0 + 1;
    b.nj = r.j1 - r.j0 + 1;
    b.nk = nk;
    b.off = off;
    b.coff = coff;
    b.rnj = 1.0f / static_cast<float>(b.nj);
    b.rnk = 1.0f / static_cast<float>(b.nk);
    b.nw = (b.nj + 63) / 64;
    b.woff = woff;
    off += b.ni * b.nj * nk;
    coff += b.ni * b.nj;
    woff += b.ni * b.nw;
  }
  out.nslots = off;
  out.ncols = coff;
  out.nsurfw = woff;
  out.stride = (off + 127) & ~127;
}

}  // namespace thesis
