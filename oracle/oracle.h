/* FP64 CPU oracle: a plain-C restatement of 3DThesis's Calc/Grid/Melt/Run path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker or the timed
 * CPU baseline. The product (3dthesis_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates. Parity pins are
 * listed in oracle.c's header comment.
 */
#ifndef THESIS_ORACLE_H
#define THESIS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int smode;          /* 0 = line, 1 = spot */
  double sx, sy, sz;  /* segment end point (m) */
  double sqmod;       /* power modulation */
  double sparam;      /* speed (line) or dwell (spot) */
  double seg_time;    /* cumulative end time */
} or_seg;

typedef struct {
  double ax, ay, az, eff, q, nond_dt;
} or_beam;

typedef struct {
  size_t size, cap;
  double *xb, *yb, *zb, *phix, *phiy, *phiz, *dtau, *expmod;
} or_nodes;

enum { OR_MODE_SNAPSHOTS = 0, OR_MODE_SOLIDIFICATION = 1 };
enum { OR_TRACK_NONE = 0, OR_TRACK_VOLUME = 1, OR_TRACK_SURFACE = 2, OR_TRACK_STORK = 3 };
/* abs() semantics at Grid.cpp:381 and :535. 0 = int truncation (g++ build),
 * 1 = floating fabs (clang build). */
enum { OR_ABS_INT = 0, OR_ABS_FABS = 1 };

typedef struct {
  char name[256], data_dir[1024];
  /* mode */
  int mode, tracking;
  int nsnap;
  double *snap_times;
  double dt, radius_check;
  int out_freq, secondary;
  /* material */
  double T_init, T_liq, kon, cps, rho, a, cet_N0, cet_n, cet_a;
  /* beams and paths (CSR) */
  int nbeams;
  or_beam *beams;
  int *path_off;   /* nbeams+1 */
  or_seg *segs;
  double all_end;
  /* domain */
  int xnum, ynum, znum, pnum;
  double xmin, xmax, ymin, ymax, zmin, zmax, xres, yres, zres;
  int use_bcs, custom_points;
  /* output */
  int o_x, o_y, o_z, o_T, o_Thist, o_tSol, o_G, o_Gx, o_Gy, o_Gz, o_V, o_dTdt,
      o_eqFrac, o_depth, o_numMelt, o_RDF, o_mpstats;
  /* settings */
  double sol_tol, t_hist, p_hist, r_max;
  int sol_iter, compress, thnum;
  /* oracle-only knobs */
  int abs_mode;
  int parse_error;  /* nonzero: the reference would have exited or crashed */
  char parse_msg[512];
  /* wall planes (DBL_MAX: unset) and image rounds (INT_MAX: unset), Init.cpp:679-684 */
  double bc_xmin, bc_xmax, bc_ymin, bc_ymax, bc_zmin;
  int bc_refl;
  /* custom point coordinates (m), pnum of them (Init.cpp:816-857) */
  double *pts_x, *pts_y, *pts_z;
} or_sim;

/* Init::GetFileNames + Init::ReadSimParams (Init.cpp:233-303). Returns NULL on
 * allocation failure; check sim->parse_error. */
or_sim* or_load(const char* paraminput);
void or_free(or_sim* sim);

/* Calc::Integrate_Serial (Calc.cpp:48-65): GaussIntegrate (Calc.cpp:67-211),
 * or GaussCompressIntegrate (Calc.cpp:213-500) with the minimum-size rule,
 * then AddBCs (Calc.cpp:502-590). hint: per-beam start segment (Calc.cpp:70);
 * updated when !is_sol. out must be empty. */
void or_nodes_clear(or_nodes* n);
void or_nodes_free(or_nodes* n);
void or_build_nodes(const or_sim* sim, int* hint, double t, int is_sol, or_nodes* out);

/* Point coordinates (Grid.h:310-327) and Grid::Calc_T's sum (Grid.cpp:286-298). */
void or_point_xyz(const or_sim* sim, int p, double* x, double* y, double* z);
double or_calc_T_xyz(const or_sim* sim, const or_nodes* nodes, double x, double y, double z);

/* Calc_Solidficiaton_Primary (Grid.cpp:406-453): out[5] = Gx,Gy,Gz,Laplace,S. */
void or_sol_primary_xyz(const or_nodes* nodes, double x, double y, double z, double out[5]);

/* Full mode driver (Run.cpp:27-709 + Main.cpp:104-115). Writes the reference
 * CSVs into sim->data_dir and, when full_dir != NULL, full-precision dumps
 * "<full_dir>/<Name>.<tag>.full.txt" (rows: p then every output column, %.17g)
 * and, with the T_hist output, "<full_dir>/<Name>.T_hist.full.txt" (rows:
 * p t T, %.17g, per point in recording order), with RDF "<Name>.RDF.full.txt"
 * (p tm tl cr per melt event) and with Stork tracking "<Name>.RRDF.full.txt"
 * (ci cj ck t_prev t_cur and 16 vertex temperatures per record).
 * evals (optional) receives the number of point-node evaluations performed.
 * Returns 0 on success. */
int or_run(or_sim* sim, const char* full_dir, double* evals);

/* Snapshot-at-t helper: T at every grid point from one node build
 * (Run.cpp:103-113). T must hold pnum doubles. Returns node count. */
long or_snapshot_T(const or_sim* sim, double t, double* T, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
