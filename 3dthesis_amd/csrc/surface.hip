// Surface melt-pool tracking of a window (Melt.cpp:185-317 as Run.cpp:600-709
// drives it), in two kernels:
//
//   k_track_surface   -- one workgroup replays the per-step set logic on the
//                        surface layer alone: the recheck of the previous
//                        liquid points, the beam-trace seeds and the
//                        neighbour waves, on 64-bit row bitmaps (registers
//                        and LDS). It writes each step's evaluated,
//                        solidified and liquid surface words.
//   k_surface_columns -- one wavefront per surface word, lane = column, walks
//                        every column step by step: the solidified column
//                        under a surface point that solidifies (Run.cpp:628-643)
//                        and Melt::calc_depth under a liquid one
//                        (Melt.cpp:271-317). It writes the event bytes, the
//                        step counts and each step's column depths; its last
//                        workgroup ends the window before the first step whose
//                        column walk left the boxes and hands on that step's
//                        liquid list, column depths and extents. It runs on
//                        another stream, beside the next window's tracker.
//
// Surface tracking grows the pool only within the surface layer
// (Melt.cpp:202-227), and a column's walks read only that column's surface
// words, its liquid bits and its own depth, so they are independent across
// columns and leave the one-workgroup loop.
#include <algorithm>

#include "window.hpp"
#include "window_dev.hpp"

namespace thesis {
namespace {

typedef unsigned long long u64;

constexpr int kSeedCache = 2048;  // seeds of a window held in LDS (16-bit codes)
constexpr int kStepCache = 256;   // steps of a window whose seed offsets and finish flags are held in LDS
constexpr int kSurfPf = 4;        // surface words per thread prefetched into registers (device-memory tracker)
constexpr int kSurfWordsPerThread = 4;  // most surface words a thread of the LDS tracker owns

enum { kMapLQ, kMapLP, kMapLN, kMapEV, kMapFA, kMapFB, kMapSO, kNumMaps };

// Where a surface bitmap word sits, computed once per window.
struct SurfWord {
  int ssl0;             // surface slot of bit 0 (bit b: ssl0 + b * nk)
  int col0;             // column of bit 0
  int gi, gj0;          // grid i of the row, grid j of bit 0
  int up, dn, lf, rt;   // neighbouring words in the box (-1: none)
  int nbits;            // valid bits
  int face;             // 1: i-1 row exists in the grid but not the box, 2: i+1, 4: j-1 of bit 0, 8: j+1 of the last bit
  int pad[2];
};
static_assert(sizeof(SurfWord) == 48, "the global word table is sized for 48-byte words");

// Word q -> box, box-local row, word within the row.
__device__ __forceinline__ void word_pos(const WinGeom& w, int q, int& b, int& li, int& ww) {
  b = 0;
  while (b + 1 < w.nbox && q >= w.b[b + 1].woff) ++b;
  const int r = q - w.b[b].woff;
  li = r / w.b[b].nw;
  ww = r - li * w.b[b].nw;
}

// Wave-wide maximum.
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// The hand-off fields of the result block before k_surface_columns adds to
// them: no liquid point, no depth, empty extents.
__device__ __forceinline__ void reset_handoff(int* R, const GridGeom& g) {
  R[kTrColTime + 0] = R[kTrColTime + 1] = R[kTrColTime + 2] = 0;
  R[kTrNliq] = 0;
  R[kTrMaxDepth] = 0;
  R[kTrMinK] = g.nz;
  R[kTrLiqBox + 0] = g.nx;
  R[kTrLiqBox + 1] = -1;
  R[kTrLiqBox + 2] = g.ny;
  R[kTrLiqBox + 3] = -1;
}

// The liquid points of surface word q after the tracker's last step into
// the hand-off list, one thread per word: a slot in the list by an LDS
// counter, the extents (min k, i_lo, i_hi, j_lo, j_hi) by LDS atomics.
__device__ __forceinline__ void handoff_thread(const TrackArgs& a, int q, u64 lpw, int* s_nl, int* s_ext) {
  if (!lpw) return;
  const WinGeom& w = a.w;
  const int nz = a.g.nz;
  int b, li, ww;
  word_pos(w, q, b, li, ww);
  const int i = w.b[b].i0 + li, j0 = w.b[b].j0 + ww * 64;
  int n = atomicAdd(s_nl, __popcll(lpw));
  for (u64 m = lpw; m; m &= m - 1ull, ++n)
    if (n < a.liq_cap) a.liq_out[n] = (i * a.g.ny + j0 + __builtin_ctzll(m)) * nz + (nz - 1);
  atomicMin(&s_ext[0], nz - 1);
  atomicMin(&s_ext[1], i);
  atomicMax(&s_ext[2], i);
  atomicMin(&s_ext[3], j0 + __builtin_ctzll(lpw));
  atomicMax(&s_ext[4], j0 + 63 - __builtin_clzll(lpw));
}

__device__ __forceinline__ void handoff_init(int* s_nl, int* s_ext, const GridGeom& g) {
  *s_nl = 0;
  s_ext[0] = g.nz;
  s_ext[1] = g.nx;
  s_ext[2] = -1;
  s_ext[3] = g.ny;
  s_ext[4] = -1;
}

__device__ __forceinline__ void handoff_write(int* R, int nl, const int* s_ext) {
  R[kTrNliq] = nl;
  R[kTrMinK] = s_ext[0];
  R[kTrLiqBox + 0] = s_ext[1];
  R[kTrLiqBox + 1] = s_ext[2];
  R[kTrLiqBox + 2] = s_ext[3];
  R[kTrLiqBox + 3] = s_ext[4];
}

// Barrier of the tracker's step loop: a full workgroup barrier, which also
// waits for the wave's global stores. An LDS-only wait before s_barrier
// (the waves share nothing else within a step) measured the same run time
// on cfg2 (19.0 against 19.1 ms), so the plain barrier is kept.
template <bool LDS>
__device__ __forceinline__ void step_barrier() {
  __syncthreads();
}

// ------------------------------------------------- tracker, device memory

// The tracker for windows whose surface words outgrow LDS: the same steps as
// k_track_surface, with every bitmap and the word table in device memory
// (TrackArgs::gbits) and one thread per word in loops over the words.
__global__ __launch_bounds__(1024) void k_track_surface_global(TrackArgs a) {
  constexpr int NT = 1024;
  const int tid = threadIdx.x;
  if (a.prev_result && (a.prev_result[kTrDone] != a.prev_nsteps || a.prev_result[kTrStatus] != 0)) {
    if (tid == 0) {  // the window before ended early: this one was planned for a start that never came
      for (int q = 0; q < kTrInts; ++q) a.result[q] = 0;
      a.result[kTrStatus] = 4;
    }
    return;
  }
  const WinGeom& w = a.w;
  const GridGeom& g = a.g;
  const int lane = tid & 63;
  const int nz = g.nz;
  const int NW = w.nsurfw;
  const int SS = a.surf_stride;
  const size_t RS = static_cast<size_t>(a.nsteps + 1) * SS;  // rows of one output map
  u64* out_ev = a.srows;
  u64* out_so = a.srows + RS;
  u64* out_ln = a.srows + 2 * RS;
  __shared__ int s_miss, s_mp[3], s_cnt[2], s_done, s_stop, s_bad, s_any[3], s_nl, s_ext[5];
  // the window's seeds (as surface word * 64 + bit, 0xffff outside the
  // boxes), seed offsets and finish flags, resolved once so that no step
  // waits on a global load for them
  __shared__ uint16_t s_seed[kSeedCache];
  __shared__ int s_soff[kStepCache + 1];
  __shared__ uint8_t s_fin[kStepCache];
  // phase clock: wall_clock64 ticks at 100 MHz
  long long tick = a.phase_times ? wall_clock64() : 0;
  long long ph0 = 0, ph1 = 0, ph2 = 0, ph3 = 0, ph5 = 0, ph6 = 0, ph7 = 0;
  int levels = 0, max_liq = 0;
#define LAP(v)                               \
  do {                                       \
    if (a.phase_times) {                     \
      const long long now_ = wall_clock64(); \
      v += now_ - tick;                      \
      tick = now_;                           \
    }                                        \
  } while (0)

  u64* M = a.gbits;
  SurfWord* SW = reinterpret_cast<SurfWord*>(M + kNumMaps * NW);
  u64* LQS = M + kMapLQ * NW;
  u64* LPS = M + kMapLP * NW;
  u64* LNS = M + kMapLN * NW;
  u64* EVS = M + kMapEV * NW;
  u64* SOS = M + kMapSO * NW;
  // the next step's surface words, held in registers while a step runs
  const bool sprefetch = NW <= kSurfPf * NT;
  u64 ps[kSurfPf];
#pragma unroll
  for (int u = 0; u < kSurfPf; ++u) ps[u] = 0ull;
  auto surf_load = [&](int r) __attribute__((always_inline)) {
    const u64* src = a.sl + static_cast<size_t>(r) * SS;
#pragma unroll
    for (int u = 0; u < kSurfPf; ++u)
      if (tid + u * NT < NW) ps[u] = src[tid + u * NT];
  };
  auto surf_stage = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kSurfPf; ++u)
      if (tid + u * NT < NW) LQS[tid + u * NT] = ps[u];
  };

  if (tid == 0) {
    s_miss = 0;
    s_stop = 0;
    s_done = 0;
    s_bad = 0;
    s_cnt[0] = s_cnt[1] = 0;
    s_any[0] = s_any[1] = s_any[2] = 0;
    handoff_init(&s_nl, s_ext, g);
    a.colmiss[0] = ~0ull;  // no column walk out of the boxes yet
    a.colmiss[1] = 0ull;   // k_surface_columns' workgroups done
    reset_handoff(a.result, g);
  }
  for (int q = tid; q < kNumMaps * NW; q += NT) M[q] = 0ull;
  for (int q = tid; q < NW; q += NT) {
    int b, li, ww;
    word_pos(w, q, b, li, ww);
    const WinBox& x = w.b[b];
    SurfWord d;
    d.ssl0 = x.off + (li * x.nj + ww * 64) * x.nk + (x.nk - 1);
    d.col0 = x.coff + li * x.nj + ww * 64;
    d.gi = x.i0 + li;
    d.gj0 = x.j0 + ww * 64;
    d.up = li > 0 ? q - x.nw : -1;
    d.dn = li + 1 < x.ni ? q + x.nw : -1;
    d.lf = ww > 0 ? q - 1 : -1;
    d.rt = ww + 1 < x.nw ? q + 1 : -1;
    d.nbits = min(64, x.nj - ww * 64);
    d.face = ((li == 0 && x.i0 > 0) ? 1 : 0) | ((li == x.ni - 1 && x.i0 + x.ni < g.nx) ? 2 : 0) |
             ((ww == 0 && x.j0 > 0) ? 4 : 0) | ((ww == x.nw - 1 && x.j0 + x.nj < g.ny) ? 8 : 0);
    d.pad[0] = d.pad[1] = 0;
    SW[q] = d;
  }
  if (sprefetch && a.nsteps >= 1) surf_load(1);
  __syncthreads();
  const int n_in = a.prev_result ? a.prev_result[kTrNliq] : a.n_liq_in;  // a queued window takes the last one's list
  for (int q = tid; q < n_in; q += NT) {
    int i, j, k;
    grid_ijk(g, a.liq_in[q], i, j, k);
    if (k != nz - 1) {  // Surface tracking never holds a point below the surface
      s_bad = 1;
      continue;
    }
    const int sl = win_slot(w, i, j, k);
    if (sl < 0) {
      note_miss(&s_miss, s_mp, i, j, k);
      continue;
    }
    int b, li, lj, lk;
    win_local(w, sl, b, li, lj, lk);
    const WinBox& x = w.b[b];
    atomicOr(&LPS[x.woff + li * x.nw + (lj >> 6)], 1ull << (lj & 63));
  }
  __syncthreads();
  if (s_bad) s_miss = 1;
  const int seed0 = a.seed_off[0];
  const bool cached = a.nsteps <= kStepCache && a.seed_off[a.nsteps] - seed0 <= kSeedCache && NW < 1024;
  if (cached) {
    for (int q = tid; q <= a.nsteps; q += NT) s_soff[q] = a.seed_off[q];
    for (int q = tid; q < a.nsteps; q += NT) s_fin[q] = a.finish[q];
    for (int q = tid; q < a.seed_off[a.nsteps] - seed0; q += NT) {
      int i, j, k;
      grid_ijk(g, a.seeds[seed0 + q], i, j, k);
      const int sl = win_slot(w, i, j, k);
      int code = -1;
      if (sl >= 0) {
        int b, li, lj, lk;
        win_local(w, sl, b, li, lj, lk);
        const WinBox& x = w.b[b];
        code = (x.woff + li * x.nw + (lj >> 6)) * 64 + (lj & 63);
      }
      s_seed[q] = code < 0 ? static_cast<uint16_t>(0xffffu) : static_cast<uint16_t>(code);
    }
  }
  // row 0 of the liquid words: the liquid points the window starts from
  for (int q = tid; q < NW; q += NT) out_ln[q] = LPS[q];
  if (sprefetch && a.nsteps >= 1) surf_stage();
  __syncthreads();
  LAP(ph0);

  for (int r = 1; r <= a.nsteps && !s_miss; ++r) {
    if (!sprefetch) {  // the surface words of this step (held in registers otherwise, staged by the last commit)
      for (int q = tid; q < NW; q += NT) LQS[q] = a.sl[static_cast<size_t>(r) * SS + q];
      __syncthreads();
    }
    if (sprefetch && r < a.nsteps) surf_load(r + 1);  // lands while this step runs
    LAP(ph1);

    // Recheck of the previous liquid points (Run.cpp:617-654), one thread per
    // word: each is evaluated, stays liquid or solidifies; the search restarts
    // from every one of them. Their columns are walked by k_surface_columns.
    for (int q = tid; q < NW; q += NT) {
      const u64 lqw = LQS[q], prev = LPS[q];
      EVS[q] = prev;
      LNS[q] = prev & lqw;
      M[kMapFA * NW + q] = prev;
      SOS[q] = prev & ~lqw;
    }
    __syncthreads();
    if (tid == 0) s_any[0] = s_any[1] = s_any[2] = 0;  // for the neighbour waves below
    // Beam-trace seeds (Melt.cpp:98-183): surface points, claimed if unclaimed.
    const int so0 = cached ? s_soff[r - 1] : a.seed_off[r - 1], so1 = cached ? s_soff[r] : a.seed_off[r];
    for (int q = so0 + tid; q < so1; q += NT) {
      int code = -1;
      if (cached) {
        const int c16 = s_seed[q - seed0];
        code = c16 == 0xffff ? -1 : c16;
      } else {
        int i, j, k;
        grid_ijk(g, a.seeds[q], i, j, k);
        const int sl = win_slot(w, i, j, k);
        if (sl >= 0) {
          int b, li, lj, lk;
          win_local(w, sl, b, li, lj, lk);
          const WinBox& x = w.b[b];
          code = (x.woff + li * x.nw + (lj >> 6)) * 64 + (lj & 63);
        }
      }
      if (code < 0) {  // outside the boxes
        int i, j, k;
        grid_ijk(g, a.seeds[q], i, j, k);
        note_miss(&s_miss, s_mp, i, j, k);
        continue;
      }
      const int q2 = code >> 6;
      const u64 bit = 1ull << (code & 63);
      const u64 old = atomicOr(&EVS[q2], bit);
      if (!(old & bit) && (LQS[q2] & bit)) {
        atomicOr(&LNS[q2], bit);
        atomicOr(&M[kMapFA * NW + q2], bit);
      }
    }
    __syncthreads();
    LAP(ph2);

    // Neighbour waves (Melt::neighbor_check, Melt.cpp:185-269): frontier rows
    // dilated by one in i and j, minus the claimed bits; the liquid ones form
    // the next frontier.
    int cur = kMapFA, nxt = kMapFB, lvl = 0;
    while (true) {
      const u64* FR = M + cur * NW;
      u64* FN = M + nxt * NW;
      bool any = false;
      for (int q = tid; q < NW; q += NT) {
        const SurfWord& sw = SW[q];
        // vertical dilation of this word and its left and right neighbours
        auto vert = [&](int wq) __attribute__((always_inline)) -> u64 {
          if (wq < 0) return 0ull;
          const SurfWord& n = SW[wq];
          u64 v = FR[wq];
          if (n.up >= 0) v |= FR[n.up];
          if (n.dn >= 0) v |= FR[n.dn];
          return v;
        };
        const u64 vm = vert(q), vl = vert(sw.lf), vr = vert(sw.rt);
        const u64 dil = vm | (vm << 1) | (vm >> 1) | (vl >> 63) | (vr << 63);
        const int nbits = sw.nbits;
        const u64 valid = nbits == 64 ? ~0ull : ((1ull << nbits) - 1ull);
        const u64 f = FR[q];
        // a frontier point on a box face with grid points beyond it
        if (f && sw.face) {
          const bool lo_j = (sw.face & 4) && (f & 1ull), hi_j = (sw.face & 8) && ((f >> (nbits - 1)) & 1ull);
          if ((sw.face & 3) || lo_j || hi_j) {  // report a grid point just outside the face
            int mi = sw.gi, mj = sw.gj0 + (__ffsll(f) - 1);
            if (sw.face & 1) mi = sw.gi - 1;
            else if (sw.face & 2) mi = sw.gi + 1;
            else if (lo_j) mj = sw.gj0 - 1;
            else mj = sw.gj0 + nbits;
            note_miss(&s_miss, s_mp, mi, mj, nz - 1);
          }
        }
        const u64 ev = EVS[q];
        const u64 claim = dil & valid & ~ev;
        const u64 newf = claim & LQS[q];
        EVS[q] = ev | claim;
        LNS[q] |= newf;
        FN[q] = newf;
        any = any || newf != 0ull;
      }
      ++levels;
      // any frontier left: three rotating flags, so the reset of the one for
      // two levels ahead cannot meet a late reader or an early writer
      if (any) s_any[lvl % 3] = 1;
      __syncthreads();
      const bool more = s_any[lvl % 3] != 0;
      if (tid == 0) s_any[(lvl + 2) % 3] = 0;
      ++lvl;
      if (!more || s_miss) break;
      const int t = cur;
      cur = nxt;
      nxt = t;
    }
    LAP(ph3);
    if (s_miss) break;  // the step is not committed

    // Commit: liquid-new becomes liquid-prev, the step's words go out for the
    // column walks and the hand-off, the next step's words come in.
    const size_t ro = static_cast<size_t>(r) * SS;
    int c_lq = 0;
    for (int q = tid; q < NW; q += NT) {
      const u64 ln = LNS[q];
      LPS[q] = ln;
      out_ev[ro + q] = EVS[q];
      out_so[ro + q] = SOS[q];
      out_ln[ro + q] = ln;
      c_lq += __popcll(ln);
    }
    if (sprefetch && r < a.nsteps) surf_stage();
    c_lq = wave_sum_i(c_lq);
    if (lane == 0 && c_lq) atomicAdd(&s_cnt[r & 1], c_lq);
    LAP(ph5);  // the commit's own work, before its barrier
    __syncthreads();
    const int nl = s_cnt[r & 1];
    const bool stop = (cached ? s_fin[r - 1] : a.finish[r - 1]) && nl == 0;  // Util::sim_finish (Util.cpp:239-243)
    max_liq = max(max_liq, nl);
    if (tid == 0) {
      int* sc = a.step_counts + 4 * (r - 1);
      sc[0] = sc[1] = sc[2] = 0;  // k_surface_columns adds these
      sc[3] = nl;
      s_done = r;
      s_stop = stop ? 1 : 0;
      s_cnt[(r + 1) & 1] = 0;
    }
    LAP(ph6);
    if (stop) break;
  }
  __syncthreads();
  // hand-off: the liquid points after the last step committed
  for (int q = tid; q < NW; q += NT) handoff_thread(a, q, LPS[q], &s_nl, s_ext);
  __syncthreads();
  if (tid == 0) {
    handoff_write(a.result, s_nl, s_ext);
    a.result[kTrDone] = s_done;
    a.result[kTrStatus] = s_bad ? 3 : (s_miss ? 2 : (s_stop ? 1 : 0));
    a.result[kTrMissI] = s_miss ? s_mp[0] : -1;
    a.result[kTrMissJ] = s_miss ? s_mp[1] : -1;
    a.result[kTrMissK] = s_miss ? s_mp[2] : -1;
    a.result[kTrLevels] = levels;
    LAP(ph7);
    a.result[kTrTime + 0] = static_cast<int>(ph0);
    a.result[kTrTime + 1] = static_cast<int>(ph1);
    a.result[kTrTime + 2] = static_cast<int>(ph2);
    a.result[kTrTime + 3] = static_cast<int>(ph5);
    a.result[kTrTime + 4] = static_cast<int>(ph3);
    a.result[kTrTime + 5] = 0;
    a.result[kTrTime + 6] = static_cast<int>(ph6);
    a.result[kTrTime + 7] = static_cast<int>(ph7);
    a.result[kTrScans] = 0;
    a.result[kTrMaxFront] = 0;
    a.result[kTrMaxLiq] = max_liq;
  }
#undef LAP
}

// ------------------------------------------------- tracker, registers + LDS

// 64-bit readlane with a wave-uniform lane index.
__device__ __forceinline__ u64 readlane64(u64 v, int l) {
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), l));
  return (static_cast<u64>(hi) << 32) | lo;
}

// The Surface tracker. Thread t owns surface words t, t + NT, ... (KW of
// them) and keeps their state in registers for the whole window: this
// step's liquid words, the previous step's liquid points, the claimed,
// liquid-new and solidified bits, and where its neighbouring words are. LDS
// holds only what other threads read -- the frontier rows of the neighbour
// waves, each with a zero word at index NW standing for a missing neighbour
// -- and, where they fit, the window's surface rows, loaded once so that no
// step waits on device memory, and each step's beam-trace seeds as bit masks
// of the surface words. Without the masks a step's seeds are claimed with
// LDS atomics on its claimed, liquid-new and liquid words.
template <int NT, int KW>
__global__ __launch_bounds__(NT) void k_track_surface(TrackArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  // the window's critical path: issue ahead of the box-sum waves that share its SIMDs
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x;
  if (a.prev_result && (a.prev_result[kTrDone] != a.prev_nsteps || a.prev_result[kTrStatus] != 0)) {
    if (tid == 0) {  // the window before ended early: this one was planned for a start that never came
      for (int q = 0; q < kTrInts; ++q) a.result[q] = 0;
      a.result[kTrStatus] = 4;
    }
    return;
  }
  const WinGeom& w = a.w;
  const GridGeom& g = a.g;
  const int lane = tid & 63;
  const int nz = g.nz;
  const int NW = w.nsurfw;
  const int SS = a.surf_stride;
  const size_t RS = static_cast<size_t>(a.nsteps + 1) * SS;  // rows of one output map
  u64* out_ev = a.srows;
  u64* out_so = a.srows + RS;
  u64* out_ln = a.srows + 2 * RS;
  __shared__ int s_miss, s_mp[3], s_cnt[2], s_bad, s_any[3], s_nl, s_ext[5];
  __shared__ u64 s_seedmiss;  // the first seed outside the boxes, step << 32 | grid point (seed masks)
  __shared__ int s_soff[kStepCache + 1];
  __shared__ uint8_t s_fin[kStepCache];
  long long tick = a.phase_times ? wall_clock64() : 0;
  long long ph0 = 0, ph1 = 0, ph2 = 0, ph3 = 0, ph5 = 0, ph6 = 0, ph7 = 0;
  int levels = 0, max_liq = 0;
#define LAP(v)                               \
  do {                                       \
    if (a.phase_times) {                     \
      const long long now_ = wall_clock64(); \
      v += now_ - tick;                      \
      tick = now_;                           \
    }                                        \
  } while (0)

  u64* LQS = reinterpret_cast<u64*>(lds);
  u64* EVS = LQS + NW;
  u64* LNS = EVS + NW;
  u64* F0 = LNS + NW;
  u64* F1 = F0 + NW + 1;
  u64* ROWS = F1 + NW + 1;  // surface rows 1 .. nsteps, NW words each (a.surf_rows_lds)
  u64* SEEDS = ROWS + static_cast<size_t>(a.nsteps) * NW;  // seed masks of steps 1 .. nsteps (a.surf_rows_lds == 2)
  const bool rows_lds = a.surf_rows_lds != 0;
  const bool masks = a.surf_rows_lds == 2;

  u64 lq[KW], lqn[KW], lp[KW], ln[KW], ev[KW], so[KW], sm[KW];
  int up[KW], dn[KW], meta[KW];  // meta: 1 a word to the left, 2 to the right, valid bits << 2, box faces << 9
#pragma unroll
  for (int u = 0; u < KW; ++u) {
    lq[u] = lqn[u] = lp[u] = ln[u] = ev[u] = so[u] = sm[u] = 0ull;
    up[u] = dn[u] = NW;
    meta[u] = 0;
  }
#define OWN(u) (tid + (u) * NT < NW)

  if (tid == 0) {
    s_miss = 0;
    s_bad = 0;
    s_cnt[0] = s_cnt[1] = 0;
    s_any[0] = s_any[1] = s_any[2] = 0;
    s_seedmiss = ~0ull;
    handoff_init(&s_nl, s_ext, g);
    a.colmiss[0] = ~0ull;  // no column walk out of the boxes yet
    a.colmiss[1] = 0ull;   // k_surface_columns' workgroups done
    reset_handoff(a.result, g);
  }
  for (int q = tid; q < 5 * NW + 2; q += NT) LQS[q] = 0ull;
  if (masks)
    for (int i = tid; i < a.nsteps * NW; i += NT) SEEDS[i] = 0ull;
  const int seed0 = a.seed_off[0];
  const bool cached = a.nsteps <= kStepCache;  // seed offsets and finish flags in LDS
  if (cached) {
    for (int q = tid; q <= a.nsteps; q += NT) s_soff[q] = a.seed_off[q];
    for (int q = tid; q < a.nsteps; q += NT) s_fin[q] = a.finish[q];
  }
#pragma unroll
  for (int u = 0; u < KW; ++u) {
    if (!OWN(u)) continue;
    const int q = tid + u * NT;
    int b, li, ww;
    word_pos(w, q, b, li, ww);
    const WinBox& x = w.b[b];
    up[u] = li > 0 ? q - x.nw : NW;
    dn[u] = li + 1 < x.ni ? q + x.nw : NW;
    const int nbits = min(64, x.nj - ww * 64);
    const int face = ((li == 0 && x.i0 > 0) ? 1 : 0) | ((li == x.ni - 1 && x.i0 + x.ni < g.nx) ? 2 : 0) |
                     ((ww == 0 && x.j0 > 0) ? 4 : 0) | ((ww == x.nw - 1 && x.j0 + x.nj < g.ny) ? 8 : 0);
    meta[u] = (ww > 0 ? 1 : 0) | (ww + 1 < x.nw ? 2 : 0) | (nbits << 2) | (face << 9);
  }
  if (rows_lds) {  // every surface row of the window, 16 loads in flight per thread
    const int n = a.nsteps * NW;
    for (int i0 = tid; i0 < n; i0 += 16 * NT) {
      u64 v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int i = i0 + k * NT;
        v[k] = 0ull;
        if (i < n) {
          const int r = i / NW;
          v[k] = a.sl[static_cast<size_t>(r + 1) * SS + (i - r * NW)];
        }
      }
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (i0 + k * NT < n) ROWS[i0 + k * NT] = v[k];
    }
  }
  __syncthreads();
  // the liquid points the window starts from, into LNS first
  const int n_in = a.prev_result ? a.prev_result[kTrNliq] : a.n_liq_in;  // a queued window takes the last one's list
  for (int q = tid; q < n_in; q += NT) {
    int i, j, k;
    grid_ijk(g, a.liq_in[q], i, j, k);
    if (k != nz - 1) {  // Surface tracking never holds a point below the surface
      s_bad = 1;
      continue;
    }
    const int sl = win_slot(w, i, j, k);
    if (sl < 0) {
      note_miss(&s_miss, s_mp, i, j, k);
      continue;
    }
    int b, li, lj, lk;
    win_local(w, sl, b, li, lj, lk);
    const WinBox& x = w.b[b];
    atomicOr(&LNS[x.woff + li * x.nw + (lj >> 6)], 1ull << (lj & 63));
  }
  // Seed masks: each step's beam-trace seeds as bits of its surface words.
  // A step then claims them all at once -- the same as claiming them one by
  // one, since a claim only sets bits that are not yet set. The first step
  // with a seed outside the boxes ends the window before it.
  if (masks) {
    const int ns = s_soff[a.nsteps] - seed0;
    for (int e = tid; e < ns; e += NT) {
      int lo = 0, hi = a.nsteps;  // the step: s_soff[lo] <= seed0 + e < s_soff[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_soff[mid] <= seed0 + e) lo = mid;
        else hi = mid;
      }
      const int p = a.seeds[seed0 + e];
      int i, j, k;
      grid_ijk(g, p, i, j, k);
      const int sl = win_slot(w, i, j, k);
      if (sl < 0) {
        atomicMin(&s_seedmiss, (static_cast<u64>(lo + 1) << 32) | static_cast<unsigned>(p));
        continue;
      }
      int b, li, lj, lk;
      win_local(w, sl, b, li, lj, lk);
      const WinBox& x = w.b[b];
      atomicOr(&SEEDS[static_cast<size_t>(lo) * NW + x.woff + li * x.nw + (lj >> 6)], 1ull << (lj & 63));
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < KW; ++u) {
    if (!OWN(u)) continue;
    const int q = tid + u * NT;
    lp[u] = LNS[q];
    out_ln[q] = lp[u];  // liquid row 0
    if (!rows_lds && a.nsteps >= 1) lqn[u] = a.sl[SS + q];
  }
  if (s_bad) s_miss = 1;
  const int seedmiss_r = s_seedmiss == ~0ull ? a.nsteps + 1 : static_cast<int>(s_seedmiss >> 32);
  LAP(ph0);

  int done = 0, pend = 0;  // the last step committed; a step whose counts thread 0 has yet to write
  bool stop = false;
  for (int r = 1; r <= a.nsteps && !s_miss; ++r) {
#pragma unroll
    for (int u = 0; u < KW; ++u) {
      if (!OWN(u)) continue;
      const int q = tid + u * NT;
      lq[u] = rows_lds ? ROWS[static_cast<size_t>(r - 1) * NW + q] : lqn[u];
      if (masks) sm[u] = SEEDS[static_cast<size_t>(r - 1) * NW + q];
      if (!rows_lds && r < a.nsteps) lqn[u] = a.sl[static_cast<size_t>(r + 1) * SS + q];
    }
    LAP(ph1);
    if (r == seedmiss_r) {  // a seed of this step lies outside the boxes
      if (tid == 0) {
        int i, j, k;
        grid_ijk(g, static_cast<int>(s_seedmiss & 0xffffffffull), i, j, k);
        note_miss(&s_miss, s_mp, i, j, k);
      }
      break;
    }
    const int so0 = cached ? s_soff[r - 1] : a.seed_off[r - 1], so1 = cached ? s_soff[r] : a.seed_off[r];
    const bool seeded = so1 > so0 && !masks;  // seeds claimed with LDS atomics below
    // Recheck of the previous liquid points (Run.cpp:617-654): each is
    // evaluated, stays liquid or solidifies, and the search restarts from
    // every one of them (their columns are k_surface_columns' work). Then the
    // beam-trace seeds (Melt.cpp:98-183): surface points, claimed if
    // unclaimed, from the step's seed mask where there is one.
    u64 f0[KW];  // the first frontier: the previous liquid points and the seeds that claim liquid points
#pragma unroll
    for (int u = 0; u < KW; ++u) {
      ev[u] = lp[u];
      ln[u] = lp[u] & lq[u];
      so[u] = lp[u] & ~lq[u];
      f0[u] = lp[u];
      const u64 claim = sm[u] & ~ev[u];
      const u64 nl = claim & lq[u];
      ev[u] |= claim;
      ln[u] |= nl;
      f0[u] |= nl;
    }
#pragma unroll
    for (int u = 0; u < KW; ++u) {
      if (!OWN(u)) continue;
      const int q = tid + u * NT;
      F0[q] = f0[u];
      if (seeded) {
        LQS[q] = lq[u];
        EVS[q] = ev[u];
        LNS[q] = ln[u];
      }
    }
    step_barrier<true>();
    if (tid == 0 && pend) {  // the counts of the step before, complete at this barrier
      const int nl = s_cnt[pend & 1];
      int* sc = a.step_counts + 4 * (pend - 1);
      sc[0] = sc[1] = sc[2] = 0;  // k_surface_columns adds these
      sc[3] = nl;
      max_liq = max(max_liq, nl);
      s_cnt[pend & 1] = 0;
    }
    pend = 0;
    if (seeded) {
      for (int q = so0 + tid; q < so1; q += NT) {
        int code = -1;
        {
          int i, j, k;
          grid_ijk(g, a.seeds[q], i, j, k);
          const int sl = win_slot(w, i, j, k);
          if (sl >= 0) {
            int b, li, lj, lk;
            win_local(w, sl, b, li, lj, lk);
            const WinBox& x = w.b[b];
            code = (x.woff + li * x.nw + (lj >> 6)) * 64 + (lj & 63);
          }
        }
        if (code < 0) {  // outside the boxes
          int i, j, k;
          grid_ijk(g, a.seeds[q], i, j, k);
          note_miss(&s_miss, s_mp, i, j, k);
          continue;
        }
        const int q2 = code >> 6;
        const u64 bit = 1ull << (code & 63);
        const u64 old = atomicOr(&EVS[q2], bit);
        if (!(old & bit) && (LQS[q2] & bit)) {
          atomicOr(&LNS[q2], bit);
          atomicOr(&F0[q2], bit);
        }
      }
      step_barrier<true>();
#pragma unroll
      for (int u = 0; u < KW; ++u) {
        if (!OWN(u)) continue;
        const int q = tid + u * NT;
        ev[u] = EVS[q];
        ln[u] = LNS[q];
      }
    }
    LAP(ph2);

    // Neighbour waves (Melt::neighbor_check, Melt.cpp:185-269): frontier rows
    // dilated by one in i and j, minus the claimed bits; the liquid ones form
    // the next frontier. All nine frontier words a word needs are read at once.
    u64* FR = F0;
    u64* FN = F1;
    int lvl = 0;
    while (true) {
      bool any = false;
#pragma unroll
      for (int u = 0; u < KW; ++u) {
        if (!OWN(u)) continue;
        const int q = tid + u * NT;
        const int m = meta[u], U = up[u], D = dn[u];
        const bool hl = m & 1, hr = m & 2;
        const u64 fc = FR[q];
        const u64 vm = fc | FR[U] | FR[D];
        const u64 vl = FR[hl ? q - 1 : NW] | FR[hl && U < NW ? U - 1 : NW] | FR[hl && D < NW ? D - 1 : NW];
        const u64 vr = FR[hr ? q + 1 : NW] | FR[hr && U < NW ? U + 1 : NW] | FR[hr && D < NW ? D + 1 : NW];
        const u64 dil = vm | (vm << 1) | (vm >> 1) | (vl >> 63) | (vr << 63);
        const int nbits = (m >> 2) & 127, face = m >> 9;
        const u64 valid = nbits == 64 ? ~0ull : ((1ull << nbits) - 1ull);
        if (fc && face) {  // a frontier point on a box face with grid points beyond it
          const bool lo_j = (face & 4) && (fc & 1ull), hi_j = (face & 8) && ((fc >> (nbits - 1)) & 1ull);
          if ((face & 3) || lo_j || hi_j) {  // report a grid point just outside the face
            int b, li, ww;
            word_pos(w, q, b, li, ww);
            const int gi = w.b[b].i0 + li, gj0 = w.b[b].j0 + ww * 64;
            int mi = gi, mj = gj0 + (__ffsll(fc) - 1);
            if (face & 1) mi = gi - 1;
            else if (face & 2) mi = gi + 1;
            else if (lo_j) mj = gj0 - 1;
            else mj = gj0 + nbits;
            note_miss(&s_miss, s_mp, mi, mj, nz - 1);
          }
        }
        const u64 claim = dil & valid & ~ev[u];
        const u64 newf = claim & lq[u];
        ev[u] |= claim;
        ln[u] |= newf;
        FN[q] = newf;
        any = any || newf != 0ull;
      }
      ++levels;
      // any frontier left: three rotating flags, so the reset of the one for
      // two levels ahead cannot meet a late reader or an early writer (all
      // three are clear again when a step's waves end)
      if (any) s_any[lvl % 3] = 1;
      step_barrier<true>();
      const bool more = s_any[lvl % 3] != 0;
      if (tid == 0) s_any[(lvl + 2) % 3] = 0;
      ++lvl;
      if (!more || s_miss) break;
      u64* t = FR;
      FR = FN;
      FN = t;
    }
    LAP(ph3);
    if (s_miss) break;  // the step is not committed

    // Commit: liquid-new becomes liquid-prev, and the step's words go out for
    // the column walks and the hand-off. The counts are written by thread 0
    // at the next step's first barrier, unless the scan may end here.
    const size_t ro = static_cast<size_t>(r) * SS;
    int c_lq = 0;
#pragma unroll
    for (int u = 0; u < KW; ++u) {
      if (!OWN(u)) continue;
      const int q = tid + u * NT;
      lp[u] = ln[u];
      out_ev[ro + q] = ev[u];
      out_so[ro + q] = so[u];
      out_ln[ro + q] = ln[u];
      c_lq += __popcll(ln[u]);
    }
    c_lq = wave_sum_i(c_lq);
    if (lane == 0 && c_lq) atomicAdd(&s_cnt[r & 1], c_lq);
    done = r;
    pend = r;
    LAP(ph5);
    if (cached ? s_fin[r - 1] : a.finish[r - 1]) {  // Util::sim_finish (Util.cpp:239-243): no liquid point left
      step_barrier<true>();
      stop = s_cnt[r & 1] == 0;
      if (stop) break;
    }
    LAP(ph6);
  }
  // Hand-off: the liquid points after the last step committed; the column
  // walks may end the window earlier and hand on an earlier step instead
#pragma unroll
  for (int u = 0; u < KW; ++u)
    if (OWN(u)) handoff_thread(a, tid + u * NT, lp[u], &s_nl, s_ext);
  __syncthreads();
  if (tid == 0) {
    handoff_write(a.result, s_nl, s_ext);
    if (pend) {
      const int nl = s_cnt[pend & 1];
      int* sc = a.step_counts + 4 * (pend - 1);
      sc[0] = sc[1] = sc[2] = 0;
      sc[3] = nl;
      max_liq = max(max_liq, nl);
    }
    a.result[kTrDone] = done;
    a.result[kTrStatus] = s_bad ? 3 : (s_miss ? 2 : (stop ? 1 : 0));
    a.result[kTrMissI] = s_miss ? s_mp[0] : -1;
    a.result[kTrMissJ] = s_miss ? s_mp[1] : -1;
    a.result[kTrMissK] = s_miss ? s_mp[2] : -1;
    a.result[kTrLevels] = levels;
    LAP(ph7);
    a.result[kTrTime + 0] = static_cast<int>(ph0);
    a.result[kTrTime + 1] = static_cast<int>(ph1);
    a.result[kTrTime + 2] = static_cast<int>(ph2);
    a.result[kTrTime + 3] = static_cast<int>(ph5);
    a.result[kTrTime + 4] = static_cast<int>(ph3);
    a.result[kTrTime + 5] = 0;
    a.result[kTrTime + 6] = static_cast<int>(ph6);
    a.result[kTrTime + 7] = static_cast<int>(ph7);
    a.result[kTrScans] = 0;
    a.result[kTrMaxFront] = 0;
    a.result[kTrMaxLiq] = max_liq;
  }
#undef OWN
#undef LAP
}

// ---------------------------------------------------------------- columns

// One wavefront (one workgroup) per surface word, lane = bit = column; the
// wave walks its columns through the tracker's steps. Every load comes first,
// independent of the tracker's step count -- lane l holds the words of step
// l + 1, and each lane moves its column's liquid bits of every step into LDS
// -- so the walks wait on no load behind their own event stores. A step
// whose words show no evaluated point is skipped by the whole wave. Event
// bytes of a column are its own slots: every byte has one writer.
//
// The wave then hands on its columns' depths as of the tracker's last step
// (the tracker hands on the liquid points). The last workgroup to finish
// checks for a column walk that left the boxes: the window then ends at the
// step before the first such walk, and that workgroup rewrites the hand-off
// from the words and depths of that step. The kernel runs on the replay's
// stream, while the next window's tracker runs on its own: it reads the
// previous window's final result, so a window queued behind one that its
// column walks ended early skips itself here.
constexpr int kColSteps = 32;  // steps of a window whose loads are issued at once
__global__ __launch_bounds__(64) void k_surface_columns(TrackArgs a) {
  __builtin_amdgcn_s_setprio(3);  // on the critical path, like the tracker
  const long long c0 = a.phase_times ? wall_clock64() : 0;
  long long c1 = c0;
  const WinGeom& w = a.w;
  const GridGeom& g = a.g;
  const int q = blockIdx.x;
  const int bit = threadIdx.x;
  __shared__ u64 s_cb[kColSteps][64];  // the column's liquid bits per step (nk <= 64)
  __shared__ int s_sc[kColSteps][3];    // the wave's counts per step: evaluated, T_last refreshed, solidified
  int b, li, ww;
  word_pos(w, q, b, li, ww);
  const WinBox& x = w.b[b];
  const int nz = g.nz, nk = x.nk, lk = nk - 1;
  const int lj = ww * 64 + bit;
  const bool live = lj < x.nj;
  const int col = x.coff + li * x.nj + lj;
  const int ssl = x.off + (li * x.nj + lj) * nk + lk;  // the column's slots are ssl - lk (bottom) .. ssl (surface)
  const int dnum = (x.i0 + li) * g.ny + x.j0 + lj;
  const int p0 = ssl - lk, w0 = p0 >> 6, sh = p0 & 63;
  const int SS = a.surf_stride;
  const size_t RS = static_cast<size_t>(a.nsteps + 1) * SS;
  const size_t lwords = static_cast<size_t>(w.stride) / 64;
  int* R = a.result;
  // A window queued behind one whose column walks ended it early started from
  // a state that never came: it is skipped (its replay then applies nothing).
  if (a.prev_result && (a.prev_result[kTrDone] != a.prev_nsteps || a.prev_result[kTrStatus] != 0)) {
    if (q == 0 && bit == 0) {
      R[kTrDone] = 0;
      R[kTrStatus] = 4;
    }
    return;
  }
  // loads: result, the column's depth, the step words, the column bits
  const int status = R[kTrStatus];
  const int done = R[kTrDone];
  int D = live ? a.depths[dnum] : 0;
  const int D0 = D;
  for (int r0 = 1; r0 <= a.nsteps; r0 += kColSteps) {
    const int n = min(kColSteps, a.nsteps - r0 + 1);
    u64 evl = 0ull, sol = 0ull, lnl = 0ull;  // lane l: step r0 + l
    if (bit < n) {
      const size_t o = static_cast<size_t>(r0 + bit) * SS + q;
      evl = a.srows[o];
      sol = a.srows[RS + o];
      lnl = a.srows[2 * RS + o];
    }
    if (nk <= 64 && nz > 1 && live) {
      u64 lo[kColSteps], hi[kColSteps];
#pragma unroll
      for (int u = 0; u < kColSteps; ++u) {
        lo[u] = hi[u] = 0ull;
        if (u < n) {
          const u64* LQ = a.lq + static_cast<size_t>(r0 + u) * lwords;
          lo[u] = LQ[w0];
          if (sh + nk > 64) hi[u] = LQ[w0 + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < kColSteps; ++u)
        if (u < n) s_cb[u][bit] = (lo[u] >> sh) | (sh ? hi[u] << (64 - sh) : 0ull);
    }
    if (status >= 3) return;  // skipped or refused: nothing was tracked
    if (a.phase_times && r0 == 1) {
      __builtin_amdgcn_s_waitcnt(0);  // the loads have landed
      c1 = wall_clock64();
    }
    if (r0 == 1 && live) a.dstep[col] = static_cast<short>(D0);  // depth row 0: the window's start
    if (bit < kColSteps) s_sc[bit][0] = s_sc[bit][1] = s_sc[bit][2] = 0;  // one wave: in order with the adds below
    bool dead = !live;  // past the column's first walk out of the boxes
    const int nd = min(n, done - r0 + 1);
    for (int u = 0; u < nd; ++u) {
      const int r = r0 + u;
      const u64 evw = readlane64(evl, u);
      if (evw) {
        int c_ev = 0, c_tl = 0, c_so = 0;
        if (!dead && ((evw >> bit) & 1ull)) {
          const bool so = (readlane64(sol, u) >> bit) & 1ull;
          const bool ln = (readlane64(lnl, u) >> bit) & 1ull;
          uint8_t* ev_row = a.ev + static_cast<size_t>(r) * w.stride;
          auto put = [&](int lo, int hi, uint32_t v) __attribute__((always_inline)) {
            if (lo > hi) return;
            for (int s = lo; s <= hi; ++s) ev_row[s] = static_cast<uint8_t>(v);
            const int cnt = hi - lo + 1;
            c_ev += cnt;  // every event byte carries kEvEval
            if (v & kEvTlast) c_tl += cnt;
            if (v & kEvSolid) c_so += cnt;
          };
          auto miss = [&](int depth) __attribute__((always_inline)) {
            atomicMin(&a.colmiss[0], (static_cast<u64>(r) << 32) | static_cast<unsigned>(dnum * nz + (nz - 1 - depth)));
            dead = true;
          };
          if (so) {  // the column solidifies with its surface point: levels d0 .. 1, T_last refreshed below d0
            const int d0 = D;
            if (d0 > lk) {
              miss(d0);
            } else {
              if (d0 > 0) {
                put(ssl - d0, ssl - d0, kEvEval | kEvSolid);
                put(ssl - (d0 - 1), ssl - 1, kEvEval | kEvSolid | kEvTlast);
              }
              D = 0;
            }
          }
          if (!dead) put(ssl, ssl, kEvEval | (so ? kEvSolid : 0u));
          if (!dead && ln && nz > 1) {
            // Melt::calc_depth: a walk down from the last depth while liquid;
            // if it found nothing deeper, a walk up while solid
            const int d0 = D;
            int depth = d0;
            bool ok = true;
            if (nk <= 64) {
              // depth d is bit lk - d of the column's bits and bit d of dm
              const u64 dm = __builtin_bitreverse64(s_cb[u][bit]) >> (63 - lk);  // bit d: liquid at depth d
              const int first = d0 + 1;
              bool deeper = false;
              if (first == nz) {
                depth = nz;  // nothing below: the downward walk stops at once
              } else if (first > lk) {
                miss(first);
                ok = false;
              } else {
                // downward walk: levels first .. dstop - 1 liquid, dstop the first
                // solid one within the box (levels below the box count as solid here)
                const u64 run = ~dm >> first;
                const int dstop = first + (run ? __builtin_ctzll(run) : 64 - first);
                deeper = dstop > first;
                if (dstop <= lk) {
                  put(ssl - dstop, ssl - first, kEvEval);
                  depth = dstop;
                } else if (dstop == nz) {  // liquid to the grid bottom
                  put(ssl - lk, ssl - first, kEvEval);
                  depth = nz;
                } else {  // liquid to the box bottom with grid below it
                  miss(dstop);
                  ok = false;
                }
              }
              if (ok) {
                if (deeper) {
                  depth -= 1;  // the deepest liquid level
                } else if (d0 > 0 && d0 > lk) {
                  miss(d0);
                  ok = false;
                } else {
                  // upward walk from d0 while solid, to the first liquid level (or the surface)
                  const u64 above = d0 > 0 ? (dm & ((2ull << d0) - 2ull)) : 0ull;  // liquid levels 1 .. d0
                  const int hl = above ? 63 - __builtin_clzll(above) : 0;
                  if (d0 > hl) {
                    put(ssl - d0, ssl - d0, kEvEval | kEvSolid);
                    put(ssl - (d0 - 1), ssl - (hl + 1), kEvEval | kEvSolid | kEvTlast);
                  }
                  if (hl > 0) put(ssl - hl, ssl - hl, kEvEval);
                  depth = d0 > 0 ? hl : 0;
                }
              }
            } else {
              const u64* LQ = a.lq + static_cast<size_t>(r) * lwords;
              auto lqbit = [&](int s) __attribute__((always_inline)) -> bool { return (LQ[s >> 6] >> (s & 63)) & 1ull; };
              bool deeper = false;
              while (true) {
                ++depth;
                if (depth == nz) break;
                if (lk - depth < 0) {
                  miss(depth);
                  ok = false;
                  break;
                }
                put(ssl - depth, ssl - depth, kEvEval);
                if (!lqbit(ssl - depth)) break;
                deeper = true;
              }
              while (ok) {
                --depth;
                if (depth == 0 || deeper) break;
                if (lk - depth < 0) {
                  miss(depth);
                  ok = false;
                  break;
                }
                if (lqbit(ssl - depth)) {
                  put(ssl - depth, ssl - depth, kEvEval);
                  break;
                }
                put(ssl - depth, ssl - depth, kEvEval | kEvSolid | (depth != d0 ? kEvTlast : 0u));
              }
            }
            if (ok) {
              if (depth == nz) --depth;
              D = depth;
              if (a.depth_events) a.depth_ev[static_cast<size_t>(r) * w.ncols + col] = static_cast<short>(depth);
            }
          }
        }
        // the step's counts, gathered in LDS
        if (c_ev) atomicAdd(&s_sc[u][0], c_ev);
        if (c_tl) atomicAdd(&s_sc[u][1], c_tl);
        if (c_so) atomicAdd(&s_sc[u][2], c_so);
      }
      if (live) a.dstep[static_cast<size_t>(r) * w.ncols + col] = static_cast<short>(D);
    }
    if (bit < nd) {  // lane u adds step r0 + u's counts
      int* sc = a.step_counts + 4 * (r0 + bit - 1);
      for (int k = 0; k < 3; ++k)
        if (s_sc[bit][k]) atomicAdd(&sc[k], s_sc[bit][k]);
    }
  }
  if (status >= 3) return;
  const long long c2 = a.phase_times ? wall_clock64() : 0;
  // hand-off as of the tracker's last step
  if (live) a.depths[dnum] = D;
  const int dmax = wave_max_i(live ? D : 0);
  if (bit == 0 && dmax > 0) atomicMax(&R[kTrMaxDepth], dmax);

  if (a.phase_times && bit == 0) {
    const long long c3 = wall_clock64();
    atomicMax(&R[kTrColTime + 0], static_cast<int>(c1 - c0));
    atomicMax(&R[kTrColTime + 1], static_cast<int>(c2 - c1));
    atomicMax(&R[kTrColTime + 2], static_cast<int>(c3 - c2));
  }
  // the last workgroup: a column walk out of the boxes ends the window earlier
  __threadfence();
  unsigned long long ticket = 0;
  if (bit == 0) ticket = atomicAdd(&a.colmiss[1], 1ull);
  ticket = __shfl(ticket, 0);
  if (ticket != static_cast<unsigned long long>(gridDim.x) - 1) return;
  __threadfence();
  const u64 key = atomicOr(&a.colmiss[0], 0ull);  // read at the device's coherence point
  if (key == ~0ull || static_cast<int>(key >> 32) > done) return;
  // the hand-off of step `kept`, rebuilt by this wave alone and written once
  const int kept = static_cast<int>(key >> 32) - 1;
  int dm = 0;
  for (int c = bit; c < w.ncols; c += 64) {
    int bb = 0;
    while (bb + 1 < w.nbox && c >= w.b[bb + 1].coff) ++bb;
    const WinBox& y = w.b[bb];
    const int cc = c - y.coff;
    const int ci = div_small(cc, y.nj, y.rnj);
    const int d = a.dstep[static_cast<size_t>(kept) * w.ncols + c];
    a.depths[(y.i0 + ci) * g.ny + y.j0 + (cc - ci * y.nj)] = d;
    dm = max(dm, d);
  }
  dm = wave_max_i(dm);
  const u64* LN = a.srows + 2 * RS + static_cast<size_t>(kept) * SS;
  int nl = 0, i_lo = g.nx, i_hi = -1, j_lo = g.ny, j_hi = -1;
  for (int q2 = 0; q2 < w.nsurfw; ++q2) {
    const u64 lnw = LN[q2];
    if (!lnw) continue;
    int b2, li2, ww2;
    word_pos(w, q2, b2, li2, ww2);
    const int i = w.b[b2].i0 + li2, j0 = w.b[b2].j0 + ww2 * 64;
    if ((lnw >> bit) & 1ull) {
      const int m = nl + __popcll(lnw & ((1ull << bit) - 1ull));
      if (m < a.liq_cap) a.liq_out[m] = (i * g.ny + j0 + bit) * nz + (nz - 1);
    }
    nl += __popcll(lnw);
    i_lo = min(i_lo, i);
    i_hi = max(i_hi, i);
    j_lo = min(j_lo, j0 + __builtin_ctzll(lnw));
    j_hi = max(j_hi, j0 + 63 - __builtin_clzll(lnw));
  }
  if (bit == 0) {
    int i, j, k;
    grid_ijk(g, static_cast<int>(key & 0xffffffffull), i, j, k);
    R[kTrDone] = kept;
    R[kTrStatus] = 2;
    R[kTrMissI] = i;
    R[kTrMissJ] = j;
    R[kTrMissK] = k;
    R[kTrNliq] = nl;
    R[kTrMaxDepth] = dm;
    R[kTrMinK] = nl ? nz - 1 : nz;
    R[kTrLiqBox + 0] = i_lo;
    R[kTrLiqBox + 1] = i_hi;
    R[kTrLiqBox + 2] = j_lo;
    R[kTrLiqBox + 3] = j_hi;
  }
}

template <int NT, int KW>
void launch_surface_lds(const TrackArgs& a, size_t bytes, hipStream_t s) {
  // the dynamic LDS attribute is raised once per kernel, to the device's limit
  static size_t raised = 0;
  if (bytes > 65536 && bytes > raised) {
    const size_t lim = track_lds_limit();
    THESIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_track_surface<NT, KW>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lim)));
    raised = lim;
  }
  hipLaunchKernelGGL((k_track_surface<NT, KW>), dim3(1), dim3(NT), bytes, s, a);
}

// Words per thread: 1, 2 or 4 (1024 threads take at most two each, within
// their 128 registers).
template <int NT>
void launch_surface_nt(const TrackArgs& a, size_t bytes, hipStream_t s) {
  const int kw = (a.w.nsurfw + NT - 1) / NT;
  if (kw <= 1) launch_surface_lds<NT, 1>(a, bytes, s);
  else if (kw <= 2 || NT == 1024) launch_surface_lds<NT, 2>(a, bytes, s);
  else launch_surface_lds<NT, NT == 1024 ? 2 : 4>(a, bytes, s);
}

}  // namespace

// The LDS tracker's maps: liquid, claimed and liquid-new words, and two
// frontier rows with their zero word.
size_t track_surface_lds_bytes(int nsurfw) { return (5 * static_cast<size_t>(nsurfw) + 2) * sizeof(u64); }

int surface_track_threads(int nsurfw, int forced) {
  const int need = (nsurfw + kSurfWordsPerThread - 1) / kSurfWordsPerThread;  // threads for four words each
  if (forced > 0 && forced >= need) return forced;
  // one word per thread where a workgroup of 1024 holds them: four waves
  // with a barrier per phase beat one wave with several words per lane
  if (nsurfw <= 64) return 64;
  if (nsurfw <= 256) return 256;
  return 1024;
}

int surface_max_words() { return 2 * 1024; }

void launch_track_surface(const TrackArgs& a0, hipStream_t s) {
  if (a0.nsteps > 0 && (!a0.sl || a0.surf_stride < a0.w.nsurfw))
    throw Error(kBadArgument, "Surface tracking needs the surface words of every cache row");
  if (!a0.srows || !a0.dstep || !a0.colmiss || !a0.lq)
    throw Error(kBadArgument, "Surface tracking needs its output rows, depth rows and miss word");
  TrackArgs a = a0;
  const int NW = a.w.nsurfw;
  if (a.lds_state) {
    if (NW > surface_max_words()) throw Error(kBadArgument, "too many surface words for the LDS tracker");
    size_t bytes = track_surface_lds_bytes(NW);
    const size_t limit = track_lds_limit();
    if (bytes > limit) throw Error(kBadArgument, "Surface tracker state larger than the LDS");
    // the window's surface rows, and with them its seed masks, in LDS where they fit
    const size_t rows = static_cast<size_t>(a.nsteps) * NW * sizeof(u64);
    a.surf_rows_lds = bytes + 2 * rows <= limit && a.nsteps <= kStepCache ? 2 : (bytes + rows <= limit ? 1 : 0);
    a.surf_rows_lds = std::min(a.surf_rows_lds, a.surface_rows_max);
    bytes += a.surf_rows_lds * rows;
    switch (surface_track_threads(NW, a.surface_threads)) {
      case 64: launch_surface_nt<64>(a, bytes, s); break;
      case 256: launch_surface_nt<256>(a, bytes, s); break;
      case 1024: launch_surface_nt<1024>(a, bytes, s); break;
      default: throw Error(kBadArgument, "Surface tracker threads must be 64, 256 or 1024");
    }
  } else {
    if (!a.gbits) throw Error(kBadArgument, "device-memory Surface tracking needs its bitmap storage");
    hipLaunchKernelGGL(k_track_surface_global, dim3(1), dim3(1024), 0, s, a);
  }
  THESIS_HIP(hipGetLastError());
}

void launch_surface_columns(const TrackArgs& a, hipStream_t s) {
  if (!a.srows || !a.dstep || !a.colmiss || !a.lq)
    throw Error(kBadArgument, "Surface column walks need the tracker's rows, depth rows and miss words");
  if (a.w.nsurfw <= 0) return;
  hipLaunchKernelGGL(k_surface_columns, dim3(a.w.nsurfw), dim3(64), 0, s, a);
  THESIS_HIP(hipGetLastError());
}

}  // namespace thesis
