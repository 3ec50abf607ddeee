// Windowed melt-pool tracking for Solidification runs (Surface and Volume).
//
// The reference advances one timestep at a time: recheck the liquid points,
// claim the beam-trace seeds, grow the pool by neighbour waves, walk the
// liquid columns down (Run.cpp:469-709, Melt.cpp:185-317). Every decision it
// takes depends only on which points are liquid at the step's own time, and
// every temperature it evaluates is T(p, t_s) (or T(p, t_{s-1}) for the
// column T_last refresh). So a window of W steps is split into:
//
//   1. k_window_eval   -- T(p, t_s) for every point of a few boxes around the
//                         melt pools, every step of the window, one launch
//                         (the bulk of the arithmetic, all CUs busy);
//   2. k_track_window  -- one workgroup replays the reference's per-step set
//                         logic on LDS flag bytes, reading only the liquid
//                         bytes of step 1, and writes per-step event bytes;
//                         a point outside the boxes stops the window there
//                         (the host widens the boxes and resumes at that step).
//                         Surface windows split this into a surface-layer
//                         tracker and a column-walk kernel (surface.hip);
//   3. k_window_apply  -- one thread per box point replays its events in step
//                         order onto the persistent per-point state (T, T_last,
//                         melt flags, numMelt) and queues its solidifications.
//
// The values and decisions equal the per-step order's: the events are the
// same sets, and each point's updates are applied in the same order.
#pragma once

#include "device.hpp"
#include "track.hpp"

namespace thesis {

constexpr int kMaxBoxes = 8;
constexpr int kTrackLists = 6;  // tracker lists: liquid (2), frontier (3), touched

// An axis-aligned block of grid points [i0, i0+ni) x [j0, j0+nj) x [k0, k0+nk).
// Slots of box b are off + ((i - i0) * nj + (j - j0)) * nk + (k - k0), the
// grid's own (i, j, k) order, so a box column is nk consecutive slots.
struct WinBox {
  int i0, j0, k0, ni, nj, nk;
  int off;         // first slot
  int coff;        // first column (columns numbered over all boxes)
  float rnj, rnk;  // 1/nj, 1/nk for the slot -> (i, j, k) divisions
  int woff;        // first 64-bit word of the box's surface-layer bitmaps
  int nw;          // words per surface row (bits are j)
};

struct WinGeom {
  int nbox = 0;
  int nslots = 0;
  int ncols = 0;
  int stride = 0;  // row length of the per-step slot arrays (nslots rounded up to 128)
  int nsurfw = 0;  // surface-layer bitmap words over all boxes
  int pad[3];
  WinBox b[kMaxBoxes];
};

THD int win_slot(const WinGeom& w, int i, int j, int k) {
  for (int b = 0; b < w.nbox; ++b) {
    const WinBox& x = w.b[b];
    const int di = i - x.i0, dj = j - x.j0, dk = k - x.k0;
    if (di >= 0 && di < x.ni && dj >= 0 && dj < x.nj && dk >= 0 && dk < x.nk)
      return x.off + (di * x.nj + dj) * x.nk + dk;
  }
  return -1;
}

// a / d for 0 <= a < 2^22 from a float reciprocal: the estimate is within
// one of the quotient and one correction step fixes it.
THD int div_small(int a, int d, float rd) {
  if (a >= (1 << 22)) return a / d;
  int q = static_cast<int>(static_cast<float>(a) * rd);
  const int r = a - q * d;
  if (r < 0) --q;
  else if (r >= d) ++q;
  return q;
}

// Box and box-local (li, lj, lk) of a slot.
THD void win_local(const WinGeom& w, int slot, int& b, int& li, int& lj, int& lk) {
  b = 0;
  while (b + 1 < w.nbox && slot >= w.b[b + 1].off) ++b;
  const WinBox& x = w.b[b];
  const int r = slot - x.off;
  const int c = div_small(r, x.nk, x.rnk);
  lk = r - c * x.nk;
  li = div_small(c, x.nj, x.rnj);
  lj = c - li * x.nj;
}

THD void win_ijk(const WinGeom& w, int slot, int& i, int& j, int& k, int& col) {
  int b, li, lj, lk;
  win_local(w, slot, b, li, lj, lk);
  const WinBox& x = w.b[b];
  i = x.i0 + li;
  j = x.j0 + lj;
  k = x.k0 + lk;
  col = x.coff + li * x.nj + lj;
}

// Event bits per (step, slot), written by the tracker.
enum : uint8_t {
  kEvEval = 0x04,  // evaluated with "set" this step (Grid::Calc_T set branch)
  kEvTlast = 0x20, // T_last refreshed from the previous step's nodes (Run.cpp:637, Melt.cpp:299-302)
  kEvSolid = 0x40, // Grid::Solidify this step
};

// Tracker result block (ints).
enum {
  kTrDone = 0,     // steps completed
  kTrStatus = 1,   // 0 window end, 1 finished (t > scan end, no liquid), 2 point outside the boxes
  kTrNliq = 2,     // liquid points after the last completed step
  kTrMissI = 3, kTrMissJ = 4, kTrMissK = 5,
  kTrMaxDepth = 6, // deepest column depth after the last completed step (Surface)
  kTrMinK = 7,     // lowest layer of a liquid point (Volume)
  kTrLiqBox = 8,   // i_lo, i_hi, j_lo, j_hi of the liquid points (4 ints)
  kTrLevels = 12,  // neighbour waves run in the window
  kTrScans = 13,   // waves that overflowed their list and scanned the flags
  kTrMaxFront = 14,
  kTrMaxLiq = 15,
  kTrTime = 16,    // 8 phase times in 10 ns ticks (with TrackArgs::phase_times, else 0): setup, stage, recheck and
                   // seeds, (Surface) the commit's work before its barrier, waves, depth, commit, hand-off
  kTrColTime = 24, // Surface column walks, the slowest workgroup's times in 10 ns ticks (with phase_times):
                   // loads, walks, hand-off
  kTrInts = 28
};
// kTrStatus values: 0 window end, 1 finished, 2 point outside the boxes,
// 3 liquid point below the surface in Surface tracking, 4 skipped (the
// window before it did not run to its end).

struct TrackArgs {
  WinGeom w;
  GridGeom g;
  int nsteps;              // steps s0 .. s0+nsteps-1 occupy cache rows 1..nsteps; row 0 is step s0-1
  int surface;             // Surface tracking (else Volume)
  int depth_events;        // record per-step column depths for the depth output
  int lds_state;           // flag/liquid bytes in LDS (else in `gflags` / read from `lq`)
  const unsigned long long* lq;  // [(nsteps+1) * stride/64] T >= T_liq, bit (slot & 63) of word slot/64
  const unsigned long long* sl;  // Surface: [(nsteps+1) * surf_stride] the surface layer's liquid bits per surface word
  int surf_stride;               // ... words per row (nsurfw rounded up to even)
  const int* seed_off;     // [nsteps+1] CSR offsets into seeds (per step)
  const int* seeds;        // grid point indices
  const uint8_t* finish;   // [nsteps] t_s > scan end
  const int* liq_in;       // liquid points after step s0-1
  int n_liq_in;            // their count (with prev_result: prev_result[kTrNliq] instead)
  int* depths;             // per grid column (nx*ny), read at start and written at end
  uint8_t* ev;             // [(nsteps+1) * stride] event bytes (row r for step s0+r-1)
  short* depth_ev;         // [(nsteps+1) * ncols] column depth after calc_depth; the caller fills it with -1
  int* liq_out;            // liquid points after the last completed step
  int liq_cap;
  const int* prev_result;  // the previous window's result block (nullptr: none in flight)
  int prev_nsteps;         // ... which must show this many steps done and status 0, else this window skips
  int* result;             // kTrInts ints
  int* step_counts;        // [nsteps][4]: evaluated, T_last refreshed, solidified, liquid
  int list_cap;            // entries per tracker list (liquid, frontier and touched lists; overflow falls back to flag scans)
  int phase_times;         // record the phase times of the result block (each reading of the clock costs a wait)
  uint8_t* gflags;         // stride bytes of global flag storage (lds_state == 0)
  short* gdepth;           // 2 * ncols rounded up to 8 (lds_state == 0)
  int* glists;             // kTrackLists * list_cap (lds_state == 0, Volume)
  unsigned long long* gbits;  // 14 * nsurfw (lds_state == 0, Surface: 7 bitmaps and a 48-byte word table)
  // Surface: the tracker's evaluated, solidified and liquid surface words per
  // step, [3][(nsteps+1) * surf_stride] (liquid row 0: the starting points),
  // each column's depth at the start and after every step, [(nsteps+1) *
  // ncols], and two words: the first column walk out of the boxes as
  // (step << 32 | grid point), and the column workgroups done
  unsigned long long* srows;
  short* dstep;
  unsigned long long* colmiss;
  int surface_threads;     // Surface tracker workgroup size: 64, 256 or 1024 (0: by the number of surface words)
  int surface_rows_max;    // at most this LDS mode (below): 2 lets the launcher choose
  int surf_rows_lds;       // set by the launcher: 1 the window's surface rows are loaded into LDS at the start, 2 with its seed masks
};

// T(p, t_s) for every slot of every cache row: Tc[r * stride + slot] and the
// liquid bit. nodes of row r are [node_off[r], node_off[r+1]).
void launch_window_eval_f64(const NodeF64* nodes, const int* node_off, int nrows, int row0, const WinGeom& w,
                            const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                            hipStream_t s);
void launch_window_eval_f32c(const NodeF32c* nodes, const int* node_off, int nrows, int row0, const WinGeom& w,
                             const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                             hipStream_t s);
// The surface layer's liquid bits in the tracker's surface-word layout (bit
// b of word q: the surface slot of column b of that word's row segment), for
// cache rows row0 .. nrows-1, from the liquid bits of launch_window_eval_*:
// sl[r * surf_stride + q]; words nsurfw .. surf_stride-1 of a row are zero.
int surf_stride(const WinGeom& w);
void launch_surface_bits(const WinGeom& w, int nrows, int row0, const unsigned long long* lq, unsigned long long* sl,
                         hipStream_t s);
// Surface windows in two passes (instead of launch_window_eval_* and
// launch_surface_bits): every surface slot of every cache row row0 ..
// nrows-1, with the surface words sl; then per row the listed columns
// (surface point liquid at the row's step or the step before; cols[r * ncols
// ..], ccount[r]) in the layers below the surface, with their liquid bits in
// lq (rows row0 .. nrows-1 must be zero on entry: the window's upload
// clears them). Tc and lq hold only those
// slots; the Surface tracker and the replay read no others. nodes are
// NodeF32c (f32) or NodeF64. Needs two_pass_fits(w): at most 64 layers below
// the surface.
bool two_pass_fits(const WinGeom& w);
void launch_surface_two_pass(const void* nodes, bool f32, const int* node_off, int nrows, int row0, const WinGeom& w,
                             const GridGeom& g, double T0, double T_liq, double* Tc, unsigned long long* lq,
                             unsigned long long* sl, int* cols, int* ccount, hipStream_t s);
// Largest dynamic LDS the tracker may use on the current device, and what a
// window of `nslots` slots, `ncols` columns and lists of `list_cap` entries needs.
size_t track_lds_limit();
size_t track_lds_bytes(int nslots, int ncols, int list_cap);
// Surface tracking keeps the surface layer as row bitmaps instead of lists:
// its LDS size, and the workgroup size it runs with (`forced`, else by the
// number of surface words).
size_t track_surface_lds_bytes(int nsurfw);
int surface_track_threads(int nsurfw, int forced);
int surface_max_words();  // most surface words of a window the LDS tracker takes
// Surface windows run the tracker (launch_track_window) and then, on
// another stream once it has finished, the column walks with the window's
// final result (launch_surface_columns, surface.hip); Volume windows one
// tracker kernel.
void launch_track_surface(const TrackArgs& a, hipStream_t s);
void launch_surface_columns(const TrackArgs& a, hipStream_t s);
void launch_track_window(const TrackArgs& a, hipStream_t s);
// Melt times of a window's replay (Surface tracking with RDF output):
// step_time[r] is the time of step s0 + r for r = 0 .. nsteps; nullptr records none.
struct MeltTimes {
  const double* step_time = nullptr;
  double dt = 0.0, T_liq = 0.0;
};
// The event rows must be zero on entry: the tracker writes the bytes of the
// points it touched. Replay rows 1..result[kTrDone] onto the point state; solidifications go to q
// with step numbers step0 + r - 1 (and count in st.nsol when that is kept).
void launch_window_apply(const WinGeom& w, const GridGeom& g, int nsteps, int step0, const uint8_t* ev,
                         const double* Tc, double T_liq, const int* result, const PointState& st,
                         const SolQueue& q, const DevCounters& dc, const MeltTimes& mt, hipStream_t s);
// A window's staging on the compute queue, as ranges of 8-byte units (node
// records are 40 or 64 bytes, so a node row is whole units and rows that
// follow each other are never written by two ranges): to dst from src (a
// device-usable pointer, pinned host memory included), or zero bytes (src
// kUploadZero) or 0xff bytes (src kUploadOnes). Copies take one unit per
// thread (every load of a block in flight at once), fills kUploadFillUnits
// per block. A block reads its range's index from the block map.
struct UploadRange {
  unsigned long long src, dst;
  unsigned n8, blk0;
};
constexpr unsigned long long kUploadZero = 0, kUploadOnes = 1;
constexpr unsigned kUploadCopyUnits = 256, kUploadFillUnits = 4096;
THD unsigned upload_units(unsigned long long src) {
  return src == kUploadZero || src == kUploadOnes ? kUploadFillUnits : kUploadCopyUnits;
}
// Block `bid` of an upload with NT threads per block (the block counts of
// UploadList assume NT = 256): its range from the block map, then its units.
#if defined(__HIPCC__)
template <int NT>
__device__ __forceinline__ void upload_block(const UploadRange* __restrict__ r, const unsigned short* __restrict__ map,
                                             unsigned bid) {
  const UploadRange d = r[map[bid]];
  const unsigned per = upload_units(d.src);
  const unsigned u0 = (bid - d.blk0) * per;
  const unsigned u1 = u0 + per < d.n8 ? u0 + per : d.n8;
  unsigned long long* dst = reinterpret_cast<unsigned long long*>(d.dst);
  if (d.src == kUploadZero || d.src == kUploadOnes) {
    const unsigned long long f = d.src == kUploadZero ? 0ull : ~0ull;
    for (unsigned u = u0 + threadIdx.x; u < u1; u += NT) dst[u] = f;
  } else {
    const unsigned long long* src = reinterpret_cast<const unsigned long long*>(d.src);
    for (unsigned u = u0 + threadIdx.x; u < u1; u += NT) dst[u] = src[u];
  }
}
#endif
// ranges[0 .. n) and the block map (readable by kernels) over `blocks` blocks.
void launch_window_upload(const UploadRange* ranges, const unsigned short* map, int n, unsigned blocks,
                          hipStream_t s);
// Builds the ranges of one upload and its block map in caller-provided
// (pinned) storage: ranges for at most kUploadMaxRanges, then the map. A
// range of `bytes` moves them rounded up to 8: source and destination of a
// length that is not a multiple of 8 must hold the rounded length.
constexpr int kUploadMaxRanges = 65535;
struct UploadList {
  UploadRange* r = nullptr;
  int n = 0;
  unsigned blocks = 0;
  void add(const void* src, void* dst, size_t bytes) {
    if (!bytes) return;
    const unsigned n8 = static_cast<unsigned>((bytes + 7) / 8);
    const unsigned long long s = reinterpret_cast<unsigned long long>(src);
    r[n++] = UploadRange{s, reinterpret_cast<unsigned long long>(dst), n8, blocks};
    blocks += (n8 + upload_units(s) - 1) / upload_units(s);
  }
  // bytes of storage for `nr` ranges and `nb` blocks
  static size_t bytes_for(size_t nr, size_t nb) {
    return ((nr * sizeof(UploadRange) + 15) & ~static_cast<size_t>(15)) + nb * sizeof(unsigned short);
  }
  // storage for at most `nr` ranges moving copy_bytes and fill_bytes in all
  static size_t storage_for(size_t nr, size_t copy_bytes, size_t fill_bytes) {
    const size_t nb = copy_bytes / 8 / kUploadCopyUnits + fill_bytes / 8 / kUploadFillUnits + 2 * nr + 2;
    return bytes_for(nr, nb);
  }
  // writes the block map after the ranges (storage for bytes_for(n, blocks));
  // returns its device address, given the ranges' device address r_dev
  const unsigned short* write_map(const UploadRange* r_dev) const;
  // write_map, then launch the upload
  void launch(const UploadRange* r_dev, hipStream_t s) const;
};

// A window's results to pinned host memory in one launch: nres ints of the
// result block, the first min(liq_max, result[kTrNliq]) liquid-list entries
// and nccnt column counts (ccnt may be null when nccnt is 0).
void launch_window_export(const int* result, int nres, const int* liq, int liq_max, const int* ccnt, int nccnt,
                          int* res_h, int* liq_h, int* ccnt_h, hipStream_t s);
// depth_max / depth output from the column depth events (Melt.cpp:319-357).
void launch_window_depth(const WinGeom& w, const GridGeom& g, int nsteps, const short* depth_ev, const double* Tc,
                         const int* result, double T_liq, const PointState& st, hipStream_t s);

}  // namespace thesis
