// Device state and kernels for Solidification runs and melt-pool tracking.
#pragma once

#include "device.hpp"
#include "quadrature.hpp"

namespace thesis {

// Append-only device log of RDF events (Grid.h:103-105): point, step and up
// to two values per entry; n counts appends (entries past cap are dropped and
// flag the overflow counter).
struct RdfLog {
  int* n = nullptr;
  int cap = 0;
  int* p = nullptr;
  int* step = nullptr;
  double* a = nullptr;
  double* b = nullptr;
};

// Device arrays of the per-point state (nullptr when the output selection does
// not need them, mirroring which arrays the reference allocates, Grid.h:120-280).
struct PointState {
  double* T = nullptr;
  double* T_last = nullptr;
  uint32_t* flag = nullptr;     // "temperature computed this step" claim flag
  uint8_t* oflag = nullptr;     // output flag
  double *tSol = nullptr, *G = nullptr, *V = nullptr, *Gx = nullptr, *Gy = nullptr, *Gz = nullptr,
         *dTdt = nullptr, *eqFrac = nullptr, *depth = nullptr;
  uint16_t* numMelt = nullptr;
  int* depths = nullptr;        // per (i,j) column: deepest liquid layer index below the surface
  double* depth_max = nullptr;  // per column
  // Deferred solidification: the step of the latest write that a queued
  // solidification must not overwrite -- liquid marking or solidification
  // (stamp_liq, guards G/V/eqFrac) and solidification (stamp_sol, guards the
  // columns only Solidify writes). nullptr: no deferral.
  int* stamp_liq = nullptr;
  int* stamp_sol = nullptr;
  int step = 0;  // current time step
  // RDF (nullptr: off): solidifications so far and melt times opened so far
  // per point, melt-time events (point, time) and solidification events
  // (point, step, liquidus time, cooling rate)
  int* nsol = nullptr;
  int* ntm = nullptr;
  RdfLog tm_log{}, sol_log{};
  // T_hist with melt-pool tracking (nullptr n: off): every evaluation with
  // set appends (point, step, T) (Grid.cpp:300-304); past cap the appends are
  // dropped and the host, which reads n, reports the overflow
  RdfLog hist_log{};
};

// Append (p, step, T) to the T_hist log.
__device__ __forceinline__ void log_history(const PointState& st, int p, int step, double T) {
  const int i = atomicAdd(st.hist_log.n, 1);
  if (i < st.hist_log.cap) {
    st.hist_log.p[i] = p;
    st.hist_log.step[i] = step;
    st.hist_log.a[i] = T;
  }
}

// A point that crossed the liquidus at `step`, with T and T_last as they were then.
struct SolEntry {
  int p, step;
  double T, T_last;
};
struct SolQueue {
  SolEntry* v = nullptr;
  int* n = nullptr;
  int cap = 0;
};

// A device list: entries and a counter in device memory. cap guards appends.
struct DList {
  int* v = nullptr;
  int* n = nullptr;  // points into the shared counter block
  int cap = 0;
};

// Node set of the current step in the precision in use.
struct NodeSetDev {
  const NodeF64* f64 = nullptr;
  const NodeF32c* f32 = nullptr;
  int n64 = 0, n32 = 0;
  int prec = kFP64;
  // Node count of the reference's list (the unit of the eval counter).
  THD int count() const { return n64; }
};

// Everything a per-point solidification build needs.
struct SolParams {
  const PathSeg* segs;
  const int* seg_off;
  const BeamQ* beams;
  const int* hints;  // per-beam segment hint at the current step
  int nbeams;
  RBox rb;
  double alpha;
  double T_init, T_liq, dt, sol_tol;
  int sol_iter, abs_mode;
  double cet_N0, cet_n, cet_a;
  const WallImage* images;  // wall images (Calc::AddBCs), nimages == 0 without walls
  int nimages;
  int compress;             // compressed quadrature with the minimum-size rule (Calc.cpp:49-58)
  GLTable gl;
};

// Device-side error flag (a list outgrew its capacity) and evaluation counter
// (point-node evaluations, the unit of the throughput metric).
struct DevCounters {
  int* overflow = nullptr;
  unsigned long long* evals = nullptr;
};

// What an evaluated point does with its temperature.
struct EvalAction {
  int set = 1;              // write T and the melt flags (Grid::Calc_T with set)
  int zero_depth = 0;       // if T < T_L: the point's column depth becomes 0 (Stork, Run.cpp:796-804)
  DList liq{};              // append if T >= T_L
  DList next{};             // append if T >= T_L (BFS frontier)
  SolQueue solid{};         // queue for solidification if T <  T_L
  DList solid2{};           // append if T < T_L (column expansion queue)
};

// Launchers. `hint` bounds the list length for grid sizing (-1: unknown); the
// kernels read the true length from device memory and loop grid-stride.
void launch_reset(const DList& reset, int hint, const PointState& st, int reset_oflag, hipStream_t s);
void launch_eval_list(const NodeSetDev& ns, const GridGeom& g, const DList& pts, int hint, double T_init,
                      double T_liq, const PointState& st, const EvalAction& act, const DevCounters& dc, hipStream_t s);
void launch_claim_neighbors(const DList& test, int hint, const GridGeom& g, int surface_only, const PointState& st,
                            const DList& cand, const DList& reset, const DevCounters& dc, hipStream_t s);
void launch_claim_seeds(const int* seeds, int count, const PointState& st, const DList& cand, const DList& reset,
                        const DevCounters& dc, hipStream_t s);
void launch_column_expand(const DList& cols, int hint, const NodeSetDev& cur, const NodeSetDev& last,
                          const GridGeom& g, double T_init, double T_liq, const PointState& st, const SolQueue& solid,
                          const DList& reset, const DevCounters& dc, hipStream_t s);
void launch_calc_depth(const DList& liq, int hint, const NodeSetDev& cur, const NodeSetDev& last, const GridGeom& g,
                       double T_init, double T_liq, const PointState& st, const SolQueue& solid, const DList& reset,
                       const DevCounters& dc, hipStream_t s);
void launch_depth_max(const DList& liq, int hint, const GridGeom& g, double T_liq, const PointState& st,
                      hipStream_t s);
// Solve the queued solidifications q.v[q0, q1) (Grid::Solidify after the
// fact). step_t[s - s0] and step_hints[(s - s0) * nbeams + b] hold each step's
// time and segment hints for steps s0 .. s0 + nsteps - 1; an entry of another
// step is not solved and raises the overflow flag. max_blocks > 0: at most
// that many blocks, each taking a whole CU (a range run beside other work).
void launch_solidify(const SolQueue& q, int q0, int q1, int s0, int nsteps, const double* step_t, const int* step_hints,
                     const GridGeom& g, const SolParams& sp, int do_sol, const PointState& st, const DevCounters& dc,
                     hipStream_t s, int max_blocks = 0);
void launch_copy_list(const DList& src, int hint, const DList& dst, const DevCounters& dc, hipStream_t s);
// Write the liquid marks (G = V = DBL_MIN, eqFrac = -1) of pts[0, n) again:
// a solidification range solved while later steps ran may have stored its
// results over the mark of a point that melted again meanwhile, so the points
// liquid at an output are marked once more after every range has finished.
void launch_remark_liquid(const int* pts, int n, const PointState& st, hipStream_t s);
// Solidification/None step: T_last = T, T = Tnew (set semantics), solidify list, liquid count.
void launch_none_apply(const double* Tnew, int pnum, double T_liq, const PointState& st, const SolQueue& solid,
                       int* liquid_count, const DevCounters& dc, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
// One fused BFS wave (claim + evaluate); frontier `test`, new frontier `next`.
// `zero` (may be null), a counter that no list of this level uses, is set to 0
// for the level after.
// With `seeds`, the level also claims and evaluates those points (beam-trace
// seeds, Melt.cpp:98-183): liquid ones join `next`. Claims are first-come, so
// the step's final sets are those of claiming the seeds before the waves.
void launch_bfs_level(const NodeSetDev& ns, const GridGeom& g, const DList& test, int hint, int surface_only,
                      double T_init, double T_liq, const PointState& st, const DList& liq, const DList& next,
                      const DList& reset, int* zero, const DevCounters& dc, hipStream_t s,
                      const int* seeds = nullptr, int nseeds = 0);
// A Volume step's recheck of the previous liquid points `last` (set; liquid
// ones to liq, the others queued on `solid`) fused with its first BFS level
// (the neighbours of the still-liquid points and the beam-trace seeds claimed
// and evaluated; won points to reset, liquid ones to liq and next). Needs the
// step prologue's preclaim and a node set that fits the LDS (recheck_bfs_fits).
bool recheck_bfs_fits(const NodeSetDev& ns);
void launch_recheck_bfs(const NodeSetDev& ns, const GridGeom& g, const DList& last, int hint, double T_init,
                        double T_liq, const PointState& st, const DList& liq, const DList& next, const DList& reset,
                        const SolQueue& solid, int* zero, const int* seeds, int nseeds, const DevCounters& dc,
                        hipStream_t s);
// Claim and evaluate beam-trace seeds; liquid ones go to liq and next.
void launch_seed_eval(const int* seeds, int count, const NodeSetDev& ns, const GridGeom& g, double T_init,
                      double T_liq, const PointState& st, const DList& liq, const DList& next, const DList& reset,
                      const DevCounters& dc, hipStream_t s);
// Tracked-step prologue: reset old_reset's points, copy last into new_reset
// (and frontier when given), zero the given counters. preclaim: the points of
// old_reset with T >= T_liq (the last list) keep their claim flags. With
// up_blocks > 0, the same launch also runs the step's upload (ranges and
// block map of an UploadList over up_blocks blocks).
struct UploadRange;
void launch_step_prologue(const DList& old_reset, int hint, const DList& last, const DList& new_reset,
                          const DList& frontier, int* z0, int* z1, int* z2, int* z3, const PointState& st,
                          const DevCounters& dc, hipStream_t s, double tm_time = 0.0, double T_liq = 0.0,
                          int preclaim = 0, const UploadRange* up_ranges = nullptr,
                          const unsigned short* up_map = nullptr, unsigned up_blocks = 0);
// Stork tracking (Run.cpp:908-1073). Column depth search below each liquid
// surface point (evaluations without set), marking the liquid column in is_liq.
void launch_stork_depth(const DList& liq, int hint, const NodeSetDev& ns, const GridGeom& g, double T_init,
                        double T_liq, const PointState& st, uint8_t* is_liq, const DevCounters& dc, hipStream_t s);
// Liquid corners per grid cell into c_cur; cells whose count changed or is
// partial (now or before) are appended to rel (unordered).
void launch_stork_cells(const GridGeom& g, const uint8_t* is_liq, const uint8_t* c_prev, uint8_t* c_cur,
                        const DList& rel, const DevCounters& dc, hipStream_t s);
// Corner temperatures of cells rel[0..n): out[16 r + n] = T at corner n with
// the previous step's nodes (n < 8) and this step's (n >= 8).
void launch_stork_records(const int* rel, int n, const GridGeom& g, const NodeSetDev& prev, const NodeSetDev& cur,
                          double T_init, double* out, const DevCounters& dc, hipStream_t s);

// Copy n device counters to pinned host memory dst_h (device-usable address),
// then store seq to *seq_h (pinned) with a system-scope release.
void launch_export_counts(const int* src, int n, int* dst_h, unsigned* seq_h, unsigned seq, hipStream_t s);
// Store one int on the device, stream-ordered (counter resets).
void launch_set_int(int* p, int v, hipStream_t s);
// Blocks per CU of the point-list kernels that stage the node set in LDS.
void set_lds_blocks_per_cu(int v);
// CUs the point-list kernels that stage the node set in LDS leave to other work.
void set_reserved_cus(int v);

}  // namespace thesis
