/*
 * C-ABI of lib3dthesis_amd.so: the MI355X (gfx950) implementation of
 * 3DThesis's Calc/Grid/Melt/Run path. Plain pointers and sizes only; every
 * entry point returns a status code (0 = success) and leaves a message for
 * thesis_last_error() on failure. No exceptions cross this boundary and there
 * is no CPU fallback: calls that need the GPU fail with THESIS_E_NODEVICE when
 * none is visible.
 *
 * A context is not re-entrant: use one per host thread. Host buffers belong to
 * the caller; device memory allocated by a context is freed by
 * thesis_ctx_destroy. Functions ending in _dev take device pointers and an
 * optional hipStream_t (passed as void*, NULL = the context's own stream) and
 * return without synchronising.
 */
#ifndef THESIS_AMD_H
#define THESIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct thesis_ctx thesis_ctx;

enum {
  THESIS_OK = 0,
  THESIS_E_INPUT = 1,     /* malformed or out-of-scope input */
  THESIS_E_ARG = 2,       /* invalid argument */
  THESIS_E_DEVICE = 3,    /* HIP runtime failure */
  THESIS_E_CAPACITY = 4,  /* caller buffer or device work list too small */
  THESIS_E_NODEVICE = 5   /* no GPU visible */
};

enum { THESIS_FP64 = 0, THESIS_FP32C = 1 };

typedef struct {
  int mode;      /* 0 Snapshots, 1 Solidification */
  int tracking;  /* 0 None, 1 Volume, 2 Surface */
  int nbeams;
  int xnum, ynum, znum;
  long long pnum;
  int nsnap;
  double scan_end, dt, T_init, T_liq, alpha;
  double xmin, xmax, ymin, ymax, zmin, zmax;
} thesis_info;

typedef struct {
  double evals;        /* point-node evaluations of the last run */
  double wall_s;       /* wall time of the last run's mode driver */
  double steps;        /* time steps or snapshots */
  double nodes_built;  /* nodes built on the host */
  /* per kernel category: 0 node sums (grid, list and windowed box sums),
   * 1 per-point solidification builds, 2 tracking (column walks; windowed
   * tracker and event replay). Durations need option "timing". */
  double cat_evals[3];
  double cat_ms[3];
  double cat_launches[3];
  double evals_window;  /* evaluations of the windowed tracking's box sums (a superset of evals' tracked part) */
} thesis_stats;

/* Parse a ParamInput file and the files it names; set up the grid.
 * Replaces Init::GetFileNames + Init::ReadSimParams (Init.cpp:233-303,
 * called from bin/Main.cpp:63-66) and the Grid constructor (Grid.h:120-280).
 * Parsing needs no GPU; the device is opened on first use. */
int thesis_ctx_create(const char* paraminput, thesis_ctx** out);
int thesis_ctx_destroy(thesis_ctx* ctx);

/* Message of the last failure on this thread ("" if none). */
const char* thesis_last_error(void);

/* Options: "precision" (THESIS_FP64 | THESIS_FP32C), "abs_mode" (0: int
 * truncation of Grid.cpp:381/:535 as compiled by g++, 1: fabs as by clang),
 * "ppt" (points per thread of the FP32c grid kernel: 1, 2, 4), "progress",
 * "outputs" (0: thesis_run writes no files), "timing" (1: HIP-event timing of
 * every kernel launch, reported by thesis_get_stats), "phase_times" (1: the
 * window tracker reads the clock at each of its phases, reported by
 * thesis_get_window_stats; off by default, as every reading waits), "window" (steps per
 * tracked Solidification window, default 40; 0 runs one step per launch
 * sequence), "tracker_lds" (0: keep the window tracker's state in device
 * memory even where it fits in LDS), "async_sol" (1, the default: queued
 * solidifications are solved while windowed tracking continues, on the
 * stream "sol_stream" names; 0: at drain points only), "volume_switch" (default 24: a Volume
 * run whose windows outgrow the tracker's LDS that many times within the
 * first eighth of its scan steps starts again with one step per launch
 * sequence, which is faster for such melt pools; 0: never), "box_margin"
 * (default 5: cells of margin the windowed box planner leaves around the
 * melt pool and the beam path; a smaller margin makes windows stop early
 * more often and resume, with the same results), "depth_margin" (default 4:
 * layers the planner leaves below the deepest melt; 0 makes column walks
 * leave the boxes more often, with the same results), "surface_threads" (0, the
 * default: the Surface window tracker's workgroup size follows the number of
 * surface words; 64, 256 or 1024 fix it; results are the same),
 * "pipeline" (4, the default: tracked windows queued on the device at once,
 * 1 to 6; results are the same), "surface_rows" (2, the default: the Surface window tracker loads the
 * window's surface rows and per-step seed masks into LDS where they fit; 1:
 * the rows only; 0: neither; results are the same), "sol_stream" (1, the
 * default: queued solidifications of the windowed driver run on a stream of
 * their own; 0: on the box-sum stream, 2: the replay stream, 3: the trackers'
 * stream; results are the same), "lds_blocks_per_cu" (1, the default: blocks
 * per CU of the one-step-at-a-time list and BFS kernels, which stage the node
 * set in LDS once per block; 1..16; applies to every context of the process;
 * results are the same), "step_sol_cus" (32, the default: in the
 * one-step-at-a-time driver, queued solidifications run beside the steps in
 * ranges of at least "step_sol_batch" entries (2^18, the default), each on
 * this many CUs, which the steps' list and BFS kernels leave free while it
 * runs; 0: ranges on the whole device; results are the same),
 * "window_sol_batch" (8192, the default: queued solidifications that start a
 * range beside the windowed tracking; results are the same). */
int thesis_set_option(thesis_ctx* ctx, const char* key, double value);
/* Also write "<dir>/<Name>.<tag>.full.txt" full-precision dumps (NULL: off). */
int thesis_set_full_dump(thesis_ctx* ctx, const char* dir);

int thesis_get_info(const thesis_ctx* ctx, thesis_info* out);
int thesis_get_snapshot_times(const thesis_ctx* ctx, double* out, int cap);

/* Node list at time t: replaces Calc::Integrate_Serial (Calc.cpp:48-65, the
 * Calc.h:21-23 seam), including path compression with its minimum-size rule
 * (Calc.cpp:213-500) and wall images (Calc.cpp:502-590). out receives
 * n_nodes x 8 doubles (xb, yb, zb, 1/phix, 1/phiy, 1/phiz, dtau, expmod),
 * node-major. A build with is_sol = 0 advances the context's per-beam
 * segment hints (the reference's start_seg, Calc.cpp:70). If cap_nodes is too
 * small, n_nodes receives the required count and THESIS_E_CAPACITY is
 * returned. */
int thesis_build_nodes(thesis_ctx* ctx, double t, int is_sol, double* out, size_t cap_nodes, size_t* n_nodes);
/* Restart the per-beam segment hints at the first segment. */
int thesis_reset_hints(thesis_ctx* ctx);

/* T = T_init + node sum at grid points p_begin .. p_begin+count-1 (index
 * p = (i*ynum + j)*znum + k): replaces Grid::Calc_T (Grid.cpp:275-314, the
 * Grid.h:290 seam) without its bookkeeping. Nodes in thesis_build_nodes
 * layout, on the host; T_out on the host. */
int thesis_eval_T(thesis_ctx* ctx, const double* nodes, size_t n_nodes, int precision, long long p_begin,
                  long long count, double* T_out);
/* The same for an explicit list of grid point indices. */
int thesis_eval_T_points(thesis_ctx* ctx, const double* nodes, size_t n_nodes, int precision, const int* points,
                         size_t count, double* T_out);
/* Device-resident form for repeated evaluation: upload once, then evaluate
 * into a device array (count doubles) on the given stream. */
int thesis_upload_nodes(thesis_ctx* ctx, const double* nodes, size_t n_nodes);
int thesis_eval_T_dev(thesis_ctx* ctx, int precision, long long p_begin, long long count, double* T_dev,
                      void* stream);

/* Allocation-free form of thesis_upload_nodes + thesis_eval_T_dev. No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). nodes_dev holds n_nodes records in thesis_build_nodes
 * layout in device memory; T_dev receives count doubles for grid points
 * p_begin .. p_begin+count-1. FP32C packs the nodes on the device into a
 * library-owned static device array (up to 65536 nodes, 3 MiB per device,
 * held until the process ends); larger node sets fail with
 * THESIS_E_CAPACITY. Structured grids only (a custom point file:
 * THESIS_E_ARG). Calls nothing that allocates: no hipMalloc, no stream or
 * context device state is created (stream NULL = the null stream). All calls
 * on one device share that array, so calls that overlap on one device (on
 * different streams, or from several threads) race. */
int thesis_eval_T_static(const thesis_ctx* ctx, const double* nodes_dev, size_t n_nodes, int precision,
                         long long p_begin, long long count, double* T_dev, void* stream);

/* Solidification at grid points that crossed the liquidus between (t-dt,
 * T_last) and (t, T): replaces Grid::Solidify (Grid.cpp:316-338, the Grid.h:291
 * seam) with Calc_Solidification_time and Set_Solidficiaton_Primary. Needs a
 * Solidification-mode context. out7 receives 7 doubles per point: tSol, G,
 * Gx/G, Gy/G, Gz/G, dTdt, V. */
int thesis_eval_sol(thesis_ctx* ctx, double t, const int* points, size_t n, const double* T, const double* T_last,
                    double* out7);

/* The whole run for the configured mode and its output files
 * (<dir>/Data/<Name>.<tag>.csv): replaces Run::Simulate (Run.cpp:27-57, the
 * Run.h:31 seam) plus the final output of bin/Main.cpp:104-112. */
int thesis_run(thesis_ctx* ctx);
int thesis_get_stats(const thesis_ctx* ctx, thesis_stats* out);
/* Windowed tracking diagnostics of the last run (no reference counterpart):
 * out[0..44] = windows, windows ended early by a point outside the boxes,
 * box slots summed over windows, steps, windows tracked in LDS, neighbour
 * waves, the tracker's phase times in ms (option "phase_times"; setup, liquid staging, recheck
 * with seeds, the Surface commit's work before its barrier, waves, column
 * depths, commit, hand-off), waves that
 * overflowed their list and scanned the flags, largest frontier, largest
 * liquid set, then sums over windows of: boxes, box columns, box layers,
 * planner margin (cells), planner depth margin (layers), then host time in
 * ms: planning, staging, queueing trackers, building node sets ahead, waiting
 * for results, processing results, then 1 if a Volume run left the windows
 * for the per-step launch sequence ("volume_switch"), then per kernel
 * (windowed box sums, window tracker, window replay, solidification builds,
 * Surface column walks, other) the summed launch durations in ms (option
 * "timing") and the launch
 * counts, then the point-node evaluations of the windowed steps counted as
 * the reference counts them (evaluated points times their step's nodes),
 * then for Surface windows the column walks' times in ms, the slowest
 * workgroup of each window summed over windows (option "phase_times"):
 * loads, walks, hand-off. */
int thesis_get_window_stats(const thesis_ctx* ctx, double* out, int cap);

/* ---- Batched snapshots (no single counterpart in the reference: it solves
 * one deck per process; this runs many decks that share the context's
 * Material, Domain and Settings in one persistent launch).
 * Set s is a one-beam deck: beam5[5*s..] = Power, Efficiency, Width_X,
 * Width_Y, Depth_Z (the Beam file, Init.cpp:488-515); its Path file rows are
 * rows6[6*r..] for r in [row_off[s], row_off[s+1]) = Mode, X(mm), Y(mm),
 * Z(mm), Pmod, speed|time (Init.cpp:560-625); times[s] is its snapshot time
 * (NULL: each set's scan end, i.e. ScanFracs 100). Node sets are built on the
 * GPU, plain (decks with path compression or wall images are refused). */
int thesis_batch_set(thesis_ctx* ctx, int nsets, const double* beam5, const int* row_off, const double* rows6,
                     const double* times);
/* T of every set at every grid point into T_dev (nsets x pnum doubles, device),
 * one persistent launch on the given stream (NULL: the context's stream). */
int thesis_batch_eval_dev(thesis_ctx* ctx, int precision, double* T_dev, void* stream);
/* Time tile/occupancy candidates on the current batch, keep the fastest. */
int thesis_batch_autotune(thesis_ctx* ctx, int precision, double* T_dev, int* ppt, int* blocks);
/* Total nodes of the batch and the point-node evaluations of one eval. */
int thesis_batch_info(const thesis_ctx* ctx, long long* total_nodes, double* evals);
/* Node list of one set as built on the GPU (thesis_build_nodes layout). */
int thesis_batch_nodes(thesis_ctx* ctx, int set, double* out, size_t cap_nodes, size_t* n_nodes);

/* ---- Design gradients (no counterpart in the reference, which has no
 * derivatives). For a single-beam context, the design is the beam (Power,
 * Efficiency, Width_X, Width_Y, Depth_Z) and the Path-file rows (6 numbers per
 * row, file units); it starts as the deck's own. t < 0 means the scan end of
 * the current design (ScanFracs 100), which then moves with the path. Decks
 * with path compression or wall images are refused. */
int thesis_design_get(const thesis_ctx* ctx, double* beam5, double* rows6, int cap_rows, int* nrows);
int thesis_design_set(thesis_ctx* ctx, const double* beam5, const double* rows6, int nrows);
/* T at every grid point for the current design (device array of pnum). */
int thesis_forward_dev(thesis_ctx* ctx, int precision, double t, double* T_dev, void* stream);
/* Gradient of a loss L(T) given dL/dT on the device (pnum doubles): grad
 * receives 5 + 2*nrows doubles -- dL/d(Power, Efficiency, Width_X, Width_Y,
 * Depth_Z), then dL/dx, dL/dy of every path row in file units (mm). Node
 * sensitivities are FP32 sums with FP64 carries on the GPU; the chain through
 * the quadrature schedule (branch decisions frozen) is FP64 on the host. */
int thesis_adjoint_dev(thesis_ctx* ctx, double t, const double* dLdT_dev, double* grad);
/* The host half of thesis_adjoint_dev: given node sensitivities S (n_nodes x 7,
 * host: dL/dxb, dL/dyb, dL/dzb, dL/dphix, dL/dphiy, dL/dphiz, dL/d ln amplitude)
 * for the design's node list at t, write the design gradient. No GPU needed. */
int thesis_design_chain(thesis_ctx* ctx, double t, const double* S, size_t n_nodes, double* grad);
/* Node list of the current design at t (thesis_build_nodes layout). No GPU needed. */
int thesis_design_nodes(thesis_ctx* ctx, double t, double* out, size_t cap_nodes, size_t* n_nodes);

#ifdef __cplusplus
}
#endif
#endif
