/* FP64 CPU oracle for the 3DThesis Calc/Grid/Melt/Run path (plain C11 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the "port"
 * CPU baseline. The product never links, loads or calls this file.
 *
 * It restates the reference algorithm function by function; each function
 * names the reference file:line it follows. Parity pins (tests/test_oracle.py):
 *   - node lists, T fields and full mode outputs against golden vectors written
 *     by the reference itself (oracle/_ref, fixtures under tests/golden), bit-for-bit for the
 *     node lists and to the last printed digit for the CSVs;
 *   - closed forms: stationary isotropic spot (erf form), per-node energy
 *     integral, sum of quadrature weights = powered time;
 *   - finite differences of Calc_T for the gradient/Laplacian accumulators;
 *   - tracking: liquid sets equal the seeded connected components of {T>=T_L}
 *     found by exhaustive evaluation on small grids.
 * Known, documented divergence: the reference keeps a function-static segment
 * hint (Calc.cpp:70) that concurrent and per-point builds overwrite; here the
 * per-point (solidification) builds start from the current step's hint and do
 * not write it back. The two agree except when a solidification time lands
 * exactly on a segment end time.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <errno.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------- */
/* istream emulation for the keyword-brace files (Init.cpp:33-187, :560-605) */

typedef struct {
  char* s;
  size_t n, pos;
  int eof, fail;
} or_in;

static int in_open(or_in* in, const char* path) {
  memset(in, 0, sizeof(*in));
  FILE* f = fopen(path, "rb");
  if (!f) { in->fail = 1; return 0; }
  fseek(f, 0, SEEK_END);
  long len = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (len < 0) len = 0;
  in->s = (char*)malloc((size_t)len + 1);
  in->n = fread(in->s, 1, (size_t)len, f);
  in->s[in->n] = 0;
  fclose(f);
  return 1;
}
static void in_close(or_in* in) { free(in->s); in->s = NULL; }
static int is_ws(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}
/* operator>>(std::string&) */
static int in_word(or_in* in, char* buf, size_t cap) {
  if (in->fail) return 0;
  if (in->eof) { in->fail = 1; return 0; }
  while (in->pos < in->n && is_ws(in->s[in->pos])) in->pos++;
  if (in->pos >= in->n) { in->eof = 1; in->fail = 1; return 0; }
  size_t k = 0;
  while (in->pos < in->n && !is_ws(in->s[in->pos])) {
    if (k + 1 < cap) buf[k++] = in->s[in->pos];
    in->pos++;
  }
  buf[k] = 0;
  if (in->pos >= in->n) in->eof = 1;
  return 1;
}
/* std::getline(is, line) with the line discarded */
static int in_getline(or_in* in) {
  if (in->fail) return 0;
  if (in->eof) { in->fail = 1; return 0; }
  size_t count = 0;
  while (in->pos < in->n) {
    char c = in->s[in->pos++];
    count++;
    if (c == '\n') return 1;
  }
  in->eof = 1;
  if (count == 0) { in->fail = 1; return 0; }
  return 1;
}
/* operator>>(double&) / operator>>(int&) */
static int in_num(or_in* in, double* v, int as_int) {
  if (in->fail) return 0;
  if (in->eof) { in->fail = 1; return 0; }
  while (in->pos < in->n && is_ws(in->s[in->pos])) in->pos++;
  if (in->pos >= in->n) { in->eof = 1; in->fail = 1; return 0; }
  char* end = NULL;
  const char* st = in->s + in->pos;
  if (as_int) {
    long x = strtol(st, &end, 10);
    *v = (double)x;
  } else {
    *v = strtod(st, &end);
  }
  if (end == st) { in->fail = 1; return 0; }
  in->pos += (size_t)(end - st);
  if (in->pos >= in->n) in->eof = 1;
  return 1;
}

#define OR_MAXW 8
#define OR_MAXS 16
#define OR_WLEN 512
typedef struct {
  int nmain;
  const char* main[OR_MAXW];
  int nsub[OR_MAXW];
  const char* sub[OR_MAXW][OR_MAXS];
  char val[OR_MAXW][OR_MAXS][OR_WLEN];
  int count[OR_MAXW];
} or_kw;

static int kw_main(const or_kw* kw, const char* w) {
  for (int i = 0; i < kw->nmain; i++)
    if (strcmp(w, kw->main[i]) == 0) return i;
  return -1;
}

/* Init::Keywords_Lv1 (Init.cpp:33-68) when lv1 != 0, else Keywords_Lv2
 * (string flavour, Init.cpp:77-131). */
static void kw_read(or_kw* kw, const char* path, int lv1) {
  for (int i = 0; i < kw->nmain; i++) {
    kw->count[i] = 0;
    for (int j = 0; j < kw->nsub[i]; j++) kw->val[i][j][0] = 0;
  }
  or_in in;
  if (!in_open(&in, path)) return;
  char w[OR_WLEN], w2[OR_WLEN];
  for (;;) {
    if (!in_word(&in, w, sizeof w)) break;
    if (!in_getline(&in)) break;
    int m = kw_main(kw, w);
    if (m < 0) continue;
    if (lv1) kw->count[m]++;
    if (!in_word(&in, w, sizeof w)) break;
    if (!in_getline(&in)) break;
    if (strcmp(w, "{") != 0) continue;
    int stop = 0;
    for (;;) {
      if (!in_word(&in, w, sizeof w)) { stop = 1; break; }
      if (strcmp(w, "}") != 0) {
        if (!in_word(&in, w2, sizeof w2)) { stop = 1; break; }
        if (!in_getline(&in)) { stop = 1; break; }
        if (!lv1)
          for (int j = 0; j < kw->nsub[m]; j++)
            if (strcmp(w, kw->sub[m][j]) == 0) {
              strncpy(kw->val[m][j], w2, OR_WLEN - 1);
              kw->val[m][j][OR_WLEN - 1] = 0;
            }
      } else {
        if (!in_getline(&in)) stop = 1;
        break;
      }
    }
    if (stop) break;
  }
  in_close(&in);
}

static void set_err(or_sim* sim, const char* msg) {
  if (!sim->parse_error) {
    sim->parse_error = 1;
    snprintf(sim->parse_msg, sizeof sim->parse_msg, "%s", msg);
  }
}
/* Init::SetValues overloads (Init.cpp:189-222): empty -> default, else stod/stoi. */
static double sv_d(or_sim* sim, const char* v, double def) {
  if (!v[0]) return def;
  char* e;
  double x = strtod(v, &e);
  if (e == v) set_err(sim, "std::stod: invalid argument");
  return x;
}
static int sv_i(or_sim* sim, const char* v, int def) {
  if (!v[0]) return def;
  char* e;
  long x = strtol(v, &e, 10);
  if (e == v) set_err(sim, "std::stoi: invalid argument");
  return (int)x;
}
static int sv_b(or_sim* sim, const char* v, int def) { return sv_i(sim, v, def) != 0; }

/* ------------------------------------------------------------------------- */
/* readers (Init.cpp:305-865)                                                */

static void read_beam(or_sim* sim, or_beam* b, const char* path) {  /* Init.cpp:488-515 */
  or_kw kw = {0};
  kw.nmain = 2;
  kw.main[0] = "Shape";     kw.nsub[0] = 3;
  kw.sub[0][0] = "Width_X"; kw.sub[0][1] = "Width_Y"; kw.sub[0][2] = "Depth_Z";
  kw.main[1] = "Intensity"; kw.nsub[1] = 2;
  kw.sub[1][0] = "Power";   kw.sub[1][1] = "Efficiency";
  kw_read(&kw, path, 0);
  b->ax = sv_d(sim, kw.val[0][0], 10.0e-6);
  b->ay = sv_d(sim, kw.val[0][1], 10.0e-6);
  b->az = sv_d(sim, kw.val[0][2], 1.0e-6);
  b->q = sv_d(sim, kw.val[1][0], 1200);
  b->eff = sv_d(sim, kw.val[1][1], 1.0);
  if (b->eff == DBL_MAX) b->eff = 1.0;      /* Init.cpp:862-865 */
  b->q = b->q * b->eff * 2.0;
}

typedef struct { or_seg* v; int n, cap; } seg_vec;
static void seg_push(seg_vec* s, or_seg x) {
  if (s->n == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 64;
    s->v = (or_seg*)realloc(s->v, sizeof(or_seg) * (size_t)s->cap);
  }
  s->v[s->n++] = x;
}

static void read_path(seg_vec* path, const char* file) {  /* Init.cpp:560-625 */
  const double convert = 1e-3;
  or_in in;
  or_seg seg;
  int opened = in_open(&in, file);
  if (opened) {
    seg.smode = 1; seg.sx = seg.sy = seg.sz = 0.0; seg.sqmod = 0.0; seg.sparam = 0.0;
    seg.seg_time = 0.0;
    seg_push(path, seg);
    while (in_getline(&in)) {
      double v[6];
      int ok = 1;
      for (int c = 0; c < 6 && ok; c++) ok = in_num(&in, &v[c], c == 0);
      if (!ok) break;
      seg.smode = (int)v[0];
      seg.sx = v[1] * convert; seg.sy = v[2] * convert; seg.sz = v[3] * convert;
      seg.sqmod = v[4]; seg.sparam = v[5];
      seg.seg_time = 0.0;
      seg_push(path, seg);
    }
    in_close(&in);
  }
  if (path->n == 0) return;  /* reference indexes path[0] regardless: undefined */
  path->v[0].seg_time = 0;
  for (int s = 1; s < path->n; s++) {
    or_seg* c = &path->v[s];
    const or_seg* p = &path->v[s - 1];
    if (c->smode) {
      c->seg_time = p->seg_time + c->sparam;
    } else {
      double dx = c->sx - p->sx, dy = c->sy - p->sy, dz = c->sz - p->sz;
      double dist = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
      c->seg_time = p->seg_time + dist / c->sparam;
    }
  }
}

/* Init::FileRead_Points (Init.cpp:816-857): the first line is skipped, then
 * x y z triples in mm (divided by 1000) until extraction fails. */
static void read_points(or_sim* sim, const char* file) {
  or_in in;
  int n = 0, cap = 0;
  if (in_open(&in, file)) {
    while (in_getline(&in)) {
      double v[3];
      int ok = 1;
      for (int c = 0; c < 3 && ok; c++) ok = in_num(&in, &v[c], 0);
      if (!ok) break;
      if (n == cap) {
        cap = cap ? 2 * cap : 256;
        sim->pts_x = (double*)realloc(sim->pts_x, sizeof(double) * (size_t)cap);
        sim->pts_y = (double*)realloc(sim->pts_y, sizeof(double) * (size_t)cap);
        sim->pts_z = (double*)realloc(sim->pts_z, sizeof(double) * (size_t)cap);
      }
      sim->pts_x[n] = v[0] / 1000.0;
      sim->pts_y[n] = v[1] / 1000.0;
      sim->pts_z[n] = v[2] / 1000.0;
      n++;
    }
    in_close(&in);
  }
  sim->pnum = n;
}

static int file_exists(const char* p) {
  FILE* f = fopen(p, "rb");
  if (!f) return 0;
  fclose(f);
  return 1;
}

/* Init::FileRead_Beams / FileRead_Paths wildcard rule (Init.cpp:445-487, :517-559). */
static int wildcard_files(or_sim* sim, const char* file, char (*out)[1024], int maxn) {
  const char* star = strchr(file, '*');
  if (!star) {
    snprintf(out[0], 1024, "%s", file);
    return 1;
  }
  if (strchr(star + 1, '*')) { set_err(sim, "Too many '*'"); return 0; }
  int n = 0;
  for (int num = 1; n < maxn; num++) {
    char cand[1024];
    snprintf(cand, sizeof cand, "%.*s%d%s", (int)(star - file), file, num, star + 1);
    if (!file_exists(cand)) break;
    snprintf(out[n++], 1024, "%s", cand);
  }
  return n;
}

static void scan_bounds(or_sim* sim) {  /* Util::Calc_ScanBounds, Util.cpp:82-136 */
  double xmin = DBL_MAX, xmax = -DBL_MAX, ymin = DBL_MAX, ymax = -DBL_MAX,
         zmin = DBL_MAX, zmax = -DBL_MAX;
  double psx = 0, psy = 0, psz = 0, psq = 0;
  int first = 1;
  for (int b = 0; b < sim->nbeams; b++) {
    for (int s = sim->path_off[b]; s < sim->path_off[b + 1]; s++) {
      const or_seg* g = &sim->segs[s];
      if (g->sqmod > 0.0) {
        if (first) {
          xmin = xmax = g->sx; ymin = ymax = g->sy; zmin = zmax = g->sz;
        } else {
          if (g->sx < xmin) xmin = g->sx;
          if (g->sx > xmax) xmax = g->sx;
          if (g->sy < ymin) ymin = g->sy;
          if (g->sy > ymax) ymax = g->sy;
          if (g->sz < zmin) zmin = g->sz;
          if (g->sz > zmax) zmax = g->sz;
          if (psq == 0.0 && g->smode == 0) {
            if (psx < xmin) xmin = psx;
            if (psx > xmax) xmax = psx;
            if (psy < ymin) ymin = psy;
            if (psy > ymax) ymax = psy;
            if (psz < zmin) zmin = psz;
            if (psz > zmax) zmax = psz;
          }
        }
      }
      first = 0;
      psx = g->sx; psy = g->sy; psz = g->sz; psq = g->sqmod;
    }
  }
  if (sim->xmin == -DBL_MAX) sim->xmin = xmin - 500e-6;
  if (sim->xmax == DBL_MAX) sim->xmax = xmax + 500e-6;
  if (sim->ymin == -DBL_MAX) sim->ymin = ymin - 500e-6;
  if (sim->ymax == DBL_MAX) sim->ymax = ymax + 500e-6;
  if (sim->zmin == -DBL_MAX) sim->zmin = zmin - 250e-6;
  if (sim->zmax == DBL_MAX) sim->zmax = zmax;
}

static void domain_params(or_sim* sim) {  /* Init::SetDomainParams, Init.cpp:866-908 */
  if (sim->custom_points) return;
  if (sim->xnum == INT_MAX) {
    sim->xnum = 1 + (int)(0.5 + (sim->xmax - sim->xmin) / sim->xres);
    sim->xmax = sim->xmin + (sim->xnum - 1) * sim->xres;
  } else if (sim->xnum != 1) {
    sim->xres = (sim->xmax - sim->xmin) / (sim->xnum - 1);
  } else {
    sim->xres = DBL_MAX;
  }
  if (sim->ynum == INT_MAX) {
    sim->ynum = 1 + (int)(0.5 + (sim->ymax - sim->ymin) / sim->yres);
    sim->ymax = sim->ymin + (sim->ynum - 1) * sim->yres;
  } else if (sim->ynum != 1) {
    sim->yres = (sim->ymax - sim->ymin) / (sim->ynum - 1);
  } else {
    sim->yres = DBL_MAX;
  }
  if (sim->znum == INT_MAX) {
    sim->znum = 1 + (int)(0.5 + (sim->zmax - sim->zmin) / sim->zres);
    sim->zmax = sim->zmin + (sim->znum - 1) * sim->zres;
  } else if (sim->ynum != 1) {  /* the reference tests ynum here (Init.cpp:898) */
    sim->zres = (sim->zmax - sim->zmin) / (sim->znum - 1);
  } else {
    sim->zres = DBL_MAX;
  }
  sim->pnum = sim->xnum * sim->ynum * sim->znum;
}

static void calc_rmax(or_sim* sim) {  /* Util::Calc_RMax, Util.cpp:145-166 */
  sim->t_hist = 1.0 / sim->t_hist;
  if (sim->r_max < 0.0) {
    for (int b = 0; b < sim->nbeams; b++) {
      const or_beam* bm = &sim->beams[b];
      if (sim->t_hist < exp(3.0 / 2.0)) sim->r_max = bm->ax * sqrt(log(sim->t_hist) / 3.0);
      else sim->r_max = bm->ax * pow(sim->t_hist, (1.0 / 3.0)) / sqrt(2.0 * exp(1.0));
      double beta = pow(3.0 / 3.14159, 1.5) * bm->q / (sim->rho * sim->cps);
      double temp_diff = sim->T_liq - sim->T_init;
      double x = temp_diff * sim->p_hist;
      double r2;
      if (beta / (x * bm->ax * bm->ax * bm->ax) < exp(3.0 / 2.0))
        r2 = bm->ax * sqrt(log(beta / (x * bm->ax * bm->ax * bm->ax)) / 3.0);
      else
        r2 = pow(beta / x, (1.0 / 3.0)) / sqrt(2.0 * exp(1.0));
      sim->r_max = sim->r_max > r2 ? sim->r_max : r2;
    }
  }
}

static int parse_tracking(const char* s) {
  if (strcmp(s, "None") == 0) return OR_TRACK_NONE;
  if (strcmp(s, "Volume") == 0) return OR_TRACK_VOLUME;
  if (strcmp(s, "Surface") == 0) return OR_TRACK_SURFACE;
  if (strcmp(s, "Stork") == 0) return OR_TRACK_STORK;
  return -1;
}

/* Comma list of Times / ScanFracs (Init.cpp:362-380). */
static void parse_times(or_sim* sim, const char* s, int fracs, double scale) {
  const char* p = s;
  for (;;) {
    const char* c = strchr(p, ',');
    size_t len = c ? (size_t)(c - p) : strlen(p);
    if (!c && len == 0) break;
    char tok[OR_WLEN];
    snprintf(tok, sizeof tok, "%.*s", (int)len, p);
    char* e;
    double v = strtod(tok, &e);
    if (e == tok) { set_err(sim, "std::stod: invalid argument"); return; }
    sim->snap_times = (double*)realloc(sim->snap_times, sizeof(double) * (size_t)(sim->nsnap + 1));
    sim->snap_times[sim->nsnap++] = fracs ? v / 100.0 * scale : v;
    if (!c) break;
    p = c + 1;
  }
}

static int in_rmax(const or_sim* sim, double x, double y);

/* GaussCompressIntegrate's R_max skip (Calc.cpp:290-302) rereads the same path
 * vertex until it lies inside the R_max box, so the reference does not
 * terminate once its walk reaches a vertex (after t = 0) outside the box.
 * Such decks are refused instead. */
static void compress_vertex_check(or_sim* sim) {
  for (int b = 0; b < sim->nbeams; b++)
    for (int s = sim->path_off[b] + 1; s < sim->path_off[b + 1]; s++)
      if (sim->segs[s].seg_time > 0.0 && !in_rmax(sim, sim->segs[s].sx, sim->segs[s].sy)) {
        set_err(sim, "path compression with a path vertex outside the R_max box (the reference does not terminate)");
        return;
      }
}

or_sim* or_load(const char* paraminput) {  /* Init.cpp:233-303 */
  or_sim* sim = (or_sim*)calloc(1, sizeof(or_sim));
  if (!sim) return NULL;
  sim->abs_mode = OR_ABS_INT;
  /* data directory (Init.cpp:235-238) */
  const char* slash = strrchr(paraminput, '/');
  if (slash) snprintf(sim->data_dir, sizeof sim->data_dir, "%.*s/Data", (int)(slash - paraminput), paraminput);
  else snprintf(sim->data_dir, sizeof sim->data_dir, "Data");
  mkdir(sim->data_dir, 0777);

  or_kw kw = {0};
  kw.nmain = 2;
  kw.main[0] = "Simulation"; kw.nsub[0] = 5;
  kw.sub[0][0] = "Name"; kw.sub[0][1] = "Mode"; kw.sub[0][2] = "Material";
  kw.sub[0][3] = "Beam"; kw.sub[0][4] = "Path";
  kw.main[1] = "Options"; kw.nsub[1] = 3;
  kw.sub[1][0] = "Domain"; kw.sub[1][1] = "Output"; kw.sub[1][2] = "Settings";
  kw_read(&kw, paraminput, 0);
  char f_mode[OR_WLEN], f_mat[OR_WLEN], f_beam[OR_WLEN], f_path[OR_WLEN], f_dom[OR_WLEN],
      f_out[OR_WLEN], f_set[OR_WLEN];
  snprintf(sim->name, sizeof sim->name, "%s", kw.val[0][0][0] ? kw.val[0][0] : "N/A");
  snprintf(f_mode, sizeof f_mode, "%s", kw.val[0][1]);
  snprintf(f_mat, sizeof f_mat, "%s", kw.val[0][2]);
  snprintf(f_beam, sizeof f_beam, "%s", kw.val[0][3]);
  snprintf(f_path, sizeof f_path, "%s", kw.val[0][4]);
  snprintf(f_dom, sizeof f_dom, "%s", kw.val[1][0]);
  snprintf(f_out, sizeof f_out, "%s", kw.val[1][1]);
  snprintf(f_set, sizeof f_set, "%s", kw.val[1][2]);

  /* beams (Init.cpp:445-487) */
  static char files[256][1024];
  int nb = wildcard_files(sim, f_beam, files, 256);
  sim->nbeams = nb;
  sim->beams = (or_beam*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(or_beam));
  for (int b = 0; b < nb; b++) read_beam(sim, &sim->beams[b], files[b]);

  /* material (Init.cpp:410-443), Util::Calc_NonD_dt (Util.cpp:138-143) */
  or_kw km = {0};
  km.nmain = 2;
  km.main[0] = "Constants"; km.nsub[0] = 5;
  km.sub[0][0] = "T_0"; km.sub[0][1] = "T_L"; km.sub[0][2] = "k"; km.sub[0][3] = "c"; km.sub[0][4] = "p";
  km.main[1] = "CET"; km.nsub[1] = 3;
  km.sub[1][0] = "N0"; km.sub[1][1] = "n"; km.sub[1][2] = "a";
  kw_read(&km, f_mat, 0);
  sim->T_init = sv_d(sim, km.val[0][0], 1273.0);
  sim->T_liq = sv_d(sim, km.val[0][1], 1610.0);
  sim->kon = sv_d(sim, km.val[0][2], 26.6);
  sim->cps = sv_d(sim, km.val[0][3], 600.00);
  sim->rho = sv_d(sim, km.val[0][4], 7451.0);
  sim->cet_N0 = sv_d(sim, km.val[1][0], DBL_MAX);
  sim->cet_n = sv_d(sim, km.val[1][1], DBL_MAX);
  sim->cet_a = sv_d(sim, km.val[1][2], DBL_MAX);
  sim->a = sim->kon / (sim->rho * sim->cps);
  for (int b = 0; b < nb; b++) sim->beams[b].nond_dt = sim->beams[b].ax * sim->beams[b].ax / sim->a;

  /* paths (Init.cpp:517-625), Util::Calc_AllScansEndTime (Util.cpp:76-80) */
  int np = wildcard_files(sim, f_path, files, 256);
  seg_vec all = {0};
  sim->path_off = (int*)calloc((size_t)np + 1, sizeof(int));
  for (int b = 0; b < np; b++) {
    seg_vec one = {0};
    read_path(&one, files[b]);
    for (int s = 0; s < one.n; s++) seg_push(&all, one.v[s]);
    sim->path_off[b + 1] = all.n;
    free(one.v);
  }
  sim->segs = all.v;
  if (np != nb) set_err(sim, "beam and path counts differ");
  if (np < nb) sim->nbeams = np;
  sim->all_end = 0;
  for (int b = 0; b < sim->nbeams; b++) {
    int last = sim->path_off[b + 1] - 1;
    if (last < sim->path_off[b] + 1) { set_err(sim, "path has no segments"); continue; }
    if (sim->segs[last].seg_time > sim->all_end) sim->all_end = sim->segs[last].seg_time;
  }

  /* mode (Init.cpp:305-408) */
  or_kw k1 = {0};
  k1.nmain = 2;
  k1.main[0] = "Snapshots"; k1.main[1] = "Solidification";
  kw_read(&k1, f_mode, 1);
  int sum = k1.count[0] + k1.count[1];
  if (sum > 1) set_err(sim, "Too many modes");
  else if (sum == 0) set_err(sim, "No modes");
  if (k1.count[0] > 0) {
    sim->mode = OR_MODE_SNAPSHOTS;
    or_kw ks = {0};
    ks.nmain = 1; ks.main[0] = "Snapshots"; ks.nsub[0] = 3;
    ks.sub[0][0] = "Times"; ks.sub[0][1] = "ScanFracs"; ks.sub[0][2] = "Tracking";
    kw_read(&ks, f_mode, 0);
    if (ks.val[0][0][0] && ks.val[0][1][0]) set_err(sim, "Too many time types");
    if (ks.val[0][0][0]) parse_times(sim, ks.val[0][0], 0, 0.0);
    if (ks.val[0][1][0]) parse_times(sim, ks.val[0][1], 1, sim->all_end);
    sim->tracking = parse_tracking(ks.val[0][2][0] ? ks.val[0][2] : "None");
  } else {
    sim->mode = OR_MODE_SOLIDIFICATION;
    or_kw ks = {0};
    ks.nmain = 1; ks.main[0] = "Solidification"; ks.nsub[0] = 5;
    ks.sub[0][0] = "Tracking"; ks.sub[0][1] = "Timestep"; ks.sub[0][2] = "OutputFrequency";
    ks.sub[0][3] = "CheckRadius"; ks.sub[0][4] = "Secondary";
    kw_read(&ks, f_mode, 0);
    sim->tracking = parse_tracking(ks.val[0][0][0] ? ks.val[0][0] : "None");
    sim->dt = sv_d(sim, ks.val[0][1], 1e-5);
    sim->out_freq = sv_i(sim, ks.val[0][2], INT_MAX);
    sim->radius_check = sv_d(sim, ks.val[0][3], 1.0);
    sim->secondary = sv_b(sim, ks.val[0][4], 0);
  }
  if (sim->tracking < 0) set_err(sim, "tracking mode outside the oracle's scope");
  /* Run.cpp:58-77 has no Stork snapshots: the reference prints an error and runs nothing */
  if (sim->mode == OR_MODE_SNAPSHOTS && sim->tracking == OR_TRACK_STORK)
    set_err(sim, "Stork tracking needs Solidification mode (the reference runs nothing for snapshots)");

  /* domain (Init.cpp:627-707) + scan bounds + sizing */
  or_kw kd = {0};
  kd.nmain = 5;
  const char* axes[3] = {"X", "Y", "Z"};
  for (int a = 0; a < 3; a++) {
    kd.main[a] = axes[a]; kd.nsub[a] = 4;
    kd.sub[a][0] = "Min"; kd.sub[a][1] = "Max"; kd.sub[a][2] = "Num"; kd.sub[a][3] = "Res";
  }
  kd.main[3] = "BoundaryConditions"; kd.nsub[3] = 6;
  kd.sub[3][0] = "X_min"; kd.sub[3][1] = "X_max"; kd.sub[3][2] = "Y_min";
  kd.sub[3][3] = "Y_max"; kd.sub[3][4] = "Z_min"; kd.sub[3][5] = "Reflections";
  kd.main[4] = "Custom"; kd.nsub[4] = 1; kd.sub[4][0] = "File";
  kw_read(&kd, f_dom, 0);
  sim->xmin = sv_d(sim, kd.val[0][0], -DBL_MAX); sim->xmax = sv_d(sim, kd.val[0][1], DBL_MAX);
  sim->xnum = sv_i(sim, kd.val[0][2], INT_MAX);  sim->xres = sv_d(sim, kd.val[0][3], 50e-6);
  sim->ymin = sv_d(sim, kd.val[1][0], -DBL_MAX); sim->ymax = sv_d(sim, kd.val[1][1], DBL_MAX);
  sim->ynum = sv_i(sim, kd.val[1][2], INT_MAX);  sim->yres = sv_d(sim, kd.val[1][3], 50e-6);
  sim->zmin = sv_d(sim, kd.val[2][0], -DBL_MAX); sim->zmax = sv_d(sim, kd.val[2][1], DBL_MAX);
  sim->znum = sv_i(sim, kd.val[2][2], INT_MAX);  sim->zres = sv_d(sim, kd.val[2][3], 50e-6);
  sim->bc_xmin = sv_d(sim, kd.val[3][0], DBL_MAX); sim->bc_xmax = sv_d(sim, kd.val[3][1], DBL_MAX);
  sim->bc_ymin = sv_d(sim, kd.val[3][2], DBL_MAX); sim->bc_ymax = sv_d(sim, kd.val[3][3], DBL_MAX);
  sim->bc_zmin = sv_d(sim, kd.val[3][4], DBL_MAX); sim->bc_refl = sv_i(sim, kd.val[3][5], INT_MAX);
  /* Init.cpp:689-694 */
  sim->use_bcs = sim->bc_xmin != DBL_MAX || sim->bc_xmax != DBL_MAX || sim->bc_ymin != DBL_MAX ||
                 sim->bc_ymax != DBL_MAX || sim->bc_zmin != DBL_MAX;
  if (kd.val[4][0][0]) {  /* Init.cpp:696-703 */
    sim->custom_points = 1;
    read_points(sim, kd.val[4][0]);
  }
  scan_bounds(sim);
  domain_params(sim);

  /* output (Init.cpp:709-771) */
  or_kw ko = {0};
  ko.nmain = 4;
  ko.main[0] = "Grid"; ko.nsub[0] = 3;
  ko.sub[0][0] = "x"; ko.sub[0][1] = "y"; ko.sub[0][2] = "z";
  ko.main[1] = "Temperature"; ko.nsub[1] = 2;
  ko.sub[1][0] = "T"; ko.sub[1][1] = "T_hist";
  ko.main[2] = "Solidification"; ko.nsub[2] = 12;
  const char* so[12] = {"tSol", "G", "Gx", "Gy", "Gz", "V", "dTdt", "eqFrac", "depth", "numMelt", "RDF", "MP_Stats"};
  for (int j = 0; j < 12; j++) ko.sub[2][j] = so[j];
  ko.main[3] = "Solidification+"; ko.nsub[3] = 4;
  ko.sub[3][0] = "H"; ko.sub[3][1] = "Hx"; ko.sub[3][2] = "Hy"; ko.sub[3][3] = "Hz";
  kw_read(&ko, f_out, 0);
  sim->o_x = sv_b(sim, ko.val[0][0], 1); sim->o_y = sv_b(sim, ko.val[0][1], 1);
  sim->o_z = sv_b(sim, ko.val[0][2], 1);
  sim->o_T = sv_b(sim, ko.val[1][0], 1); sim->o_Thist = sv_b(sim, ko.val[1][1], 0);
  sim->o_tSol = sv_b(sim, ko.val[2][0], 0); sim->o_G = sv_b(sim, ko.val[2][1], 0);
  sim->o_Gx = sv_b(sim, ko.val[2][2], 0); sim->o_Gy = sv_b(sim, ko.val[2][3], 0);
  sim->o_Gz = sv_b(sim, ko.val[2][4], 0); sim->o_V = sv_b(sim, ko.val[2][5], 0);
  sim->o_dTdt = sv_b(sim, ko.val[2][6], 0); sim->o_eqFrac = sv_b(sim, ko.val[2][7], 0);
  sim->o_depth = sv_b(sim, ko.val[2][8], 0); sim->o_numMelt = sv_b(sim, ko.val[2][9], 0);
  sim->o_RDF = sv_b(sim, ko.val[2][10], 0); sim->o_mpstats = sv_b(sim, ko.val[2][11], 0);
  if (sim->o_Thist && sim->tracking == OR_TRACK_STORK && !sim->custom_points)
    set_err(sim, "T_hist with Stork tracking on a grid is outside the oracle's scope");

  /* settings (Init.cpp:772-814) + Util::Calc_RMax */
  or_kw kt = {0};
  kt.nmain = 4;
  kt.main[0] = "Temperature"; kt.nsub[0] = 4;
  kt.sub[0][0] = "Sol_Tol"; kt.sub[0][1] = "Sol_Iter"; kt.sub[0][2] = "Cutoff_Peak"; kt.sub[0][3] = "Cutoff_T0TL";
  kt.main[1] = "Path"; kt.nsub[1] = 2;
  kt.sub[1][0] = "Buffer"; kt.sub[1][1] = "Compression";
  kt.main[2] = "Compute"; kt.nsub[2] = 2;
  kt.sub[2][0] = "MaxThreads"; kt.sub[2][1] = "PINT";
  kt.main[3] = "MPI"; kt.nsub[3] = 1; kt.sub[3][0] = "Overlap";
  kw_read(&kt, f_set, 0);
  sim->sol_tol = sv_d(sim, kt.val[0][0], 1e-3);
  sim->sol_iter = sv_i(sim, kt.val[0][1], 10);
  sim->t_hist = sv_d(sim, kt.val[0][2], 1e-9);
  sim->p_hist = sv_d(sim, kt.val[0][3], 1e-2);
  sim->r_max = sv_d(sim, kt.val[1][0], -1.0);
  sim->compress = sv_i(sim, kt.val[1][1], 0);
#ifdef _OPENMP
  sim->thnum = sv_i(sim, kt.val[2][0], omp_get_max_threads() / 2);
#else
  sim->thnum = sv_i(sim, kt.val[2][0], 1);
#endif
  calc_rmax(sim);
  if (sim->compress) compress_vertex_check(sim);
  return sim;
}

void or_free(or_sim* sim) {
  if (!sim) return;
  free(sim->snap_times);
  free(sim->beams);
  free(sim->path_off);
  free(sim->segs);
  free(sim->pts_x);
  free(sim->pts_y);
  free(sim->pts_z);
  free(sim);
}

/* ------------------------------------------------------------------------- */
/* node build                                                                 */

/* 8-digit Gauss-Legendre abscissae and weights used by the reference
 * (Calc.cpp:73-86), laid out by order; emitted from the last entry down. */
static const double GL_X2[2] = {-0.57735027, 0.57735027};
static const double GL_W2[2] = {1.0, 1.0};
static const double GL_X4[4] = {-0.86113631, -0.33998104, 0.33998104, 0.86113631};
static const double GL_W4[4] = {0.34785485, 0.65214515, 0.65214515, 0.34785485};
static const double GL_X8[8] = {-0.96028986, -0.79666648, -0.52553241, -0.18343464,
                                0.18343464, 0.52553241, 0.79666648, 0.96028986};
static const double GL_W8[8] = {0.10122854, 0.22238103, 0.31370665, 0.36268378,
                                0.36268378, 0.31370665, 0.22238103, 0.10122854};
static const double GL_X16[16] = {-0.98940093, -0.94457502, -0.8656312, -0.75540441,
                                  -0.61787624, -0.45801678, -0.28160355, -0.09501251,
                                  0.09501251, 0.28160355, 0.45801678, 0.61787624,
                                  0.75540441, 0.8656312, 0.94457502, 0.98940093};
static const double GL_W16[16] = {0.02715246, 0.06225352, 0.09515851, 0.12462897,
                                  0.14959599, 0.16915652, 0.18260342, 0.18945061,
                                  0.18945061, 0.18260342, 0.16915652, 0.14959599,
                                  0.12462897, 0.09515851, 0.06225352, 0.02715246};

void or_nodes_clear(or_nodes* n) { n->size = 0; }
void or_nodes_free(or_nodes* n) {
  free(n->xb); free(n->yb); free(n->zb); free(n->phix); free(n->phiy); free(n->phiz);
  free(n->dtau); free(n->expmod);
  memset(n, 0, sizeof(*n));
}
/* Util::AddToNodes (Util.cpp:40-50): stores inverted spreads and the log prefactor. */
static void add_node(or_nodes* n, double xb, double yb, double zb, double px, double py,
                     double pz, double dtau, double qmod) {
  if (n->size == n->cap) {
    n->cap = n->cap ? 2 * n->cap : 1024;
    double** arrs[8] = {&n->xb, &n->yb, &n->zb, &n->phix, &n->phiy, &n->phiz, &n->dtau, &n->expmod};
    for (int i = 0; i < 8; i++) *arrs[i] = (double*)realloc(*arrs[i], sizeof(double) * n->cap);
  }
  size_t i = n->size++;
  n->xb[i] = xb; n->yb[i] = yb; n->zb[i] = zb;
  n->phix[i] = 1.0 / px; n->phiy[i] = 1.0 / py; n->phiz[i] = 1.0 / pz;
  n->dtau[i] = dtau;
  n->expmod[i] = log(qmod) - log(px * py * pz) / 2.0;
}

static int in_rmax(const or_sim* sim, double x, double y) {  /* Util::InRMax, Util.cpp:168-172 */
  if ((x > (sim->xmax + sim->r_max)) || (x < (sim->xmin - sim->r_max))) return 0;
  if ((y > (sim->ymax + sim->r_max)) || (y < (sim->ymin - sim->r_max))) return 0;
  return 1;
}

/* Util::GetBeamLoc (Util.cpp:211-237) on path P (segment array). */
static void beam_loc(const or_sim* sim, const or_seg* P, int seg, double time,
                     double* xb, double* yb, double* zb, double* qmod) {
  if (P[seg].smode) {
    *xb = P[seg].sx; *yb = P[seg].sy; *zb = P[seg].sz;
  } else {
    const double dx = P[seg].sx - P[seg - 1].sx;
    const double dy = P[seg].sy - P[seg - 1].sy;
    const double dz = P[seg].sz - P[seg - 1].sz;
    const double tcur = time - P[seg - 1].seg_time;
    const double dt_cur = P[seg].seg_time - P[seg - 1].seg_time;
    *xb = P[seg - 1].sx + (tcur / dt_cur) * dx;
    *yb = P[seg - 1].sy + (tcur / dt_cur) * dy;
    *zb = P[seg - 1].sz + (tcur / dt_cur) * dz;
  }
  *qmod = in_rmax(sim, *xb, *yb) ? P[seg].sqmod : 0.0;
}

/* Util::GetRefTime (Util.cpp:192-209). */
static double ref_time(const or_seg* P, int seg, double tpp, const or_beam* b) {
  double r = 0.0;
  const double spp = (tpp / b->nond_dt) > 0.0 ? (tpp / b->nond_dt) : 0.0;
  if (P[seg].smode == 0) {
    double t0 = 0.58870501125 * b->ax / P[seg].sparam;
    r = t0 * sqrt(12.0 * spp + 1.0);
  } else if (P[seg].smode == 1) {
    r = P[seg].seg_time - P[seg - 1].seg_time;
    if (r == 0) r = 1.0e-9;
  }
  return r;
}

static const double* gl_x(int order) {
  return order == 16 ? GL_X16 : order == 8 ? GL_X8 : order == 4 ? GL_X4 : GL_X2;
}
static const double* gl_w(int order) {
  return order == 16 ? GL_W16 : order == 8 ? GL_W8 : order == 4 ? GL_W4 : GL_W2;
}

static double beam_beta(const or_sim* sim, const or_beam* bm) {  /* Calc.cpp:96 */
  return pow(3.0 / OR_PI, 1.5) * bm->q / (sim->rho * sim->cps);
}

/* Segment holding t, searched from the start hint (Calc.cpp:99-101). */
static int locate_seg(const or_seg* P, int nseg, double t, int seg) {
  while ((t > P[seg].seg_time) && (seg + 1 < nseg)) seg++;
  while ((t < P[seg - 1].seg_time) && (seg - 1 > 0)) seg--;
  return seg;
}

/* Calc::GaussIntegrate (Calc.cpp:67-211) for one beam from segment seg;
 * Util::t0calc always yields 0 (Util.cpp:187). */
static void gauss_beam(const or_sim* sim, int b, int seg, double t, int is_sol, or_nodes* out) {
  const or_beam* bm = &sim->beams[b];
  const or_seg* P = sim->segs + sim->path_off[b];
  const int nseg = sim->path_off[b + 1] - sim->path_off[b];
  const double beta = beam_beta(sim, bm);
  const double t0 = 0.0;
  double step_start = bm->nond_dt;
  if (is_sol) step_start *= bm->az;
  double step_max = step_start;
  int order = 16;
  const double end = P[nseg - 1].seg_time;

  double xb, yb, zb, qm;
  beam_loc(sim, P, seg, t, &xb, &yb, &zb, &qm);
  qm *= beta;
  if (t <= end && qm > 0.0)
    add_node(out, xb, yb, zb, bm->ax * bm->ax + 0.0, bm->ay * bm->ay + 0.0, bm->az * bm->az + 0.0, 0.0, qm);

  int more = 1;
  double t2 = t, t1 = t, tpp = 0.0;
  if (t > end) { tpp += t - end; t2 = end; }
  while (tpp >= 2 * step_max - step_start) {
    step_max *= 2.0;
    if (order != 2) order /= 2;
  }
  while (more) {
    int switch_seg = 0;
    const double rt = ref_time(P, seg, tpp, bm);
    const double use = (rt < step_max) ? rt : step_max;
    t1 = t2 - use;
    const double next_time = P[seg - 1].seg_time;
    if (t1 <= next_time) {
      t1 = next_time;
      if (next_time > t0) switch_seg = 1;
      else more = 0;
    }
    const double* X = gl_x(order);
    const double* W = gl_w(order);
    for (int a = order - 1; a >= 0; a--) {
      const double tp = 0.5 * ((t2 - t1) * X[a] + (t2 + t1));
      const double tau = t - tp;
      const double ct = 12.0 * sim->a * tau;
      beam_loc(sim, P, seg, tp, &xb, &yb, &zb, &qm);
      qm *= beta;
      const double dtau = 0.5 * (t2 - t1) * W[a];
      if (qm > 0.0 && dtau > 0.0)
        add_node(out, xb, yb, zb, bm->ax * bm->ax + ct, bm->ay * bm->ay + ct, bm->az * bm->az + ct, dtau, qm);
    }
    if (switch_seg) seg--;
    tpp += (t2 - t1);
    t2 = t1;
    if (tpp >= 2 * step_max - step_start) {
      step_max *= 2.0;
      if (order != 2) order /= 2;
    }
  }
}

/* Calc::GaussCompressIntegrate (Calc.cpp:213-500) for one beam from segment
 * seg_temp. Where the scan path turns back on itself faster than heat diffuses
 * (segment ends within the diffusion radius r2 of the beam position at t2),
 * the segments of one quadrature interval are merged into a single moving
 * source at their power-weighted centre. Two properties of the reference are
 * kept as they are:
 *   - a merged interval's nodes start from qmod = (sum q dt)/(sum dt) and the
 *     loop multiplies it by beta before every node (Calc.cpp:463), so node k
 *     of the interval carries qmod * beta^(k+1);
 *   - a merged node's zb is a field the reference never sets (Calc.cpp:452,
 *     an uninitialised int_seg); 0 here (the reference build printed 0 or
 *     subnormal stack residue for it).
 * The R_max skip loop (Calc.cpp:290-302) never moves to another segment, so
 * it does not end once it meets a path vertex outside the R_max box; or_load
 * refuses such decks (compress_vertex_check). */
static void compress_beam(const or_sim* sim, int b, int seg_temp, double t, int is_sol, or_nodes* out) {
  const or_beam* bm = &sim->beams[b];
  const or_seg* P = sim->segs + sim->path_off[b];
  const int nseg = sim->path_off[b + 1] - sim->path_off[b];
  const double beta = beam_beta(sim, bm);
  const double t0 = 0.0;
  double cs_start = bm->nond_dt;
  if (is_sol) cs_start *= bm->az;
  double cs_max = cs_start;
  double cs_use = cs_max;
  int order = 16;
  const double end = P[nseg - 1].seg_time;
  const double ax2 = bm->ax * bm->ax, ay2 = bm->ay * bm->ay, az2 = bm->az * bm->az;

  double xb, yb, zb, qm;
  beam_loc(sim, P, seg_temp, t, &xb, &yb, &zb, &qm);
  qm *= beta;
  if (t <= end && qm > 0.0) add_node(out, xb, yb, zb, ax2 + 0.0, ay2 + 0.0, az2 + 0.0, 0.0, qm);

  int tflag = 1;
  double t2 = t, t1 = t, tpp = 0.0;
  if (t > end) { tpp += t - end; t2 = end; }
  while (tpp >= 2 * cs_max - cs_start) {
    cs_max *= 2.0;
    if (order != 2) order /= 2;
  }
  while (tflag) {
    double r2, dist2, xp, yp, xs, ys, dx, dy, ts, dt;
    double sum_t = 0, sum_qt = 0, sum_qtx = 0, sum_qty = 0;
    int seg2;
    int ncomb = 0;
    int quit = 0;
    for (;;) {
      xs = P[seg_temp].sx;
      ys = P[seg_temp].sy;
      ts = P[seg_temp].seg_time;
      if (ts <= t0) { quit = 1; break; }
      if (in_rmax(sim, xs, ys)) break;
      /* not reached: or_load refuses paths with a vertex outside R_max */
      quit = 1;
      break;
    }
    if (quit) break;
    while (tpp >= 2 * cs_max - cs_start) {
      cs_max *= 2.0;
      if (order != 2) order /= 2;
    }
    const double rt = ref_time(P, seg_temp, tpp, bm);
    cs_use = (rt < cs_max) ? rt : cs_max;
    double q2;
    beam_loc(sim, P, seg_temp, t2, &xp, &yp, &zb, &q2);
    r2 = log(2.0) / 8.0 * (bm->ax * bm->ax) * (12.0 * (t - t2) * sim->a / (bm->ax * bm->ax) + 1.0);
    seg2 = seg_temp;
    t1 = P[seg_temp - 1].seg_time;
    if (t1 < t2 - cs_use) {
      ncomb = 0;
      t1 = t2 - cs_use;
    } else {
      int cflag = 1;
      while (cflag) {
        if (P[seg2 - 1].seg_time <= t0) { tflag = 0; break; }
        xs = P[seg2 - 1].sx;
        ys = P[seg2 - 1].sy;
        dist2 = (xs - xp) * (xs - xp) + (ys - yp) * (ys - yp);
        ts = P[seg2 - 1].seg_time;
        if (!in_rmax(sim, xp, yp)) break;
        if ((dist2 > r2) ||
            (P[seg2].sqmod != P[seg2 - 1].sqmod && (order > 2 || (t2 - ts) > (cs_max / 64.0)))) {
          if (P[seg2].smode) {
            sum_qtx += P[seg2].sx * P[seg2].sqmod * (t1 - ts);
            sum_qty += P[seg2].sy * P[seg2].sqmod * (t1 - ts);
            sum_qt += P[seg2].sqmod * (t1 - ts);
            sum_t += (t1 - ts);
            t1 = ts;
          } else {
            dx = P[seg2].sx - P[seg2 - 1].sx;
            dy = P[seg2].sy - P[seg2 - 1].sy;
            dt = P[seg2].seg_time - ts;
            double t_int = ts + dt * (sqrt(((xp - xs) * (xp - xs) + (yp - ys) * (yp - ys)) / (dx * dx + dy * dy)) -
                                      sqrt(r2 / (dx * dx + dy * dy)));
            if (t_int < t2 - cs_max) t_int = t2 - cs_max;
            sum_qtx += (xs + dx * ((t1 + t_int) / 2.0 - ts) / dt) * P[seg2].sqmod * (t1 - t_int);
            sum_qty += (ys + dy * ((t1 + t_int) / 2.0 - ts) / dt) * P[seg2].sqmod * (t1 - t_int);
            sum_qt += P[seg2].sqmod * (t1 - t_int);
            sum_t += (t1 - t_int);
            t1 = t_int;
          }
          cflag = 0;
        } else {
          if (ts < t2 - cs_max) {
            cflag = 0;
            t1 = t2 - cs_max;
          } else {
            t1 = ts;
            ncomb++;
          }
          if (P[seg2].smode) {
            dt = (P[seg2].seg_time - t1);
            sum_qtx += P[seg2].sx * P[seg2].sqmod * dt;
            sum_qty += P[seg2].sy * P[seg2].sqmod * dt;
            sum_qt += P[seg2].sqmod * dt;
            sum_t += dt;
          } else {
            dx = P[seg2].sx - P[seg2 - 1].sx;
            dy = P[seg2].sy - P[seg2 - 1].sy;
            dt = (P[seg2].seg_time - t1);
            sum_qtx += (xs + dx * ((P[seg2].seg_time + t1) / 2.0 - ts) / dt) * P[seg2].sqmod * dt;
            sum_qty += (ys + dy * ((P[seg2].seg_time + t1) / 2.0 - ts) / dt) * P[seg2].sqmod * dt;
            sum_qt += P[seg2].sqmod * dt;
            sum_t += dt;
          }
        }
        if (t1 == P[seg2 - 1].seg_time) {
          xs = P[seg2 - 1].sx;
          ys = P[seg2 - 1].sy;
          seg2--;
          if (!in_rmax(sim, xs, ys)) break;
        }
      }
      if (!ncomb) {
        t1 = P[seg_temp - 1].seg_time;
        seg2 = seg_temp - 1;
      }
    }
    const double* X = gl_x(order);
    const double* W = gl_w(order);
    if (ncomb) {
      if (sum_qt > 0.0) {
        const double cx = sum_qtx / sum_qt, cy = sum_qty / sum_qt;
        double q = sum_qt / sum_t;
        for (int a = order - 1; a >= 0; a--) {
          const double tp = 0.5 * ((t2 - t1) * X[a] + (t2 + t1));
          const double ct = 12.0 * sim->a * (t - tp);
          q *= beta;
          const double dtau = 0.5 * (t2 - t1) * W[a];
          if ((q > 0.0) && (dtau > 0.0)) add_node(out, cx, cy, 0.0, ax2 + ct, ay2 + ct, az2 + ct, dtau, q);
        }
      }
    } else {
      for (int a = order - 1; a >= 0; a--) {
        const double tp = 0.5 * ((t2 - t1) * X[a] + (t2 + t1));
        const double ct = 12.0 * sim->a * (t - tp);
        beam_loc(sim, P, seg_temp, tp, &xb, &yb, &zb, &qm);
        qm *= beta;
        const double dtau = 0.5 * (t2 - t1) * W[a];
        if ((qm > 0.0) && (dtau > 0.0)) add_node(out, xb, yb, zb, ax2 + ct, ay2 + ct, az2 + ct, dtau, qm);
      }
    }
    tpp += t2 - t1;
    t2 = t1;
    if (tpp >= 2 * cs_max - cs_start) {
      cs_max *= 2.0;
      if (order != 2) order /= 2;
    }
    seg_temp = seg2;
    if (t1 <= t0) tflag = 0;
  }
}

/* Mirror images of every node in the wall planes (Calc::AddBCs,
 * Calc.cpp:502-590). Image coordinates (a list of integer triples, found
 * breadth-first, `Reflections` rounds) are reflected across the set planes
 * X_min, X_max, Y_min, Y_max, Z_min and the top surface Z = zmax; the image
 * one reflection above the top surface is never added. Nodes are replicated
 * image-major in the order the coordinates were found; a coordinate r on an
 * axis with planes lo/hi maps x to (r + n) hi - (r - n) lo + (1 - 2n) x,
 * n = |r mod 2| (an unset plane is DBL_MAX and only meets a zero factor). */
static void add_bcs(const or_sim* sim, or_nodes* n) {
  const double pl[6] = {sim->bc_xmin, sim->bc_xmax, sim->bc_ymin, sim->bc_ymax, sim->bc_zmin, sim->zmax};
  size_t cap = 64, nall = 1, ncheck = 1;
  int* all = (int*)malloc(sizeof(int) * 3 * cap);
  int* check = (int*)malloc(sizeof(int) * 3 * cap);
  int* fresh = (int*)malloc(sizeof(int) * 3 * cap);
  all[0] = all[1] = all[2] = 0;
  check[0] = check[1] = check[2] = 0;
  const int strength = sim->bc_refl != INT_MAX ? sim->bc_refl : 0;
  for (int iter = 0; ncheck > 0 && iter < strength; iter++) {
    size_t nfresh = 0;
    for (size_t i = 0; i < ncheck; i++) {
      for (int w = 0; w < 6; w++) {
        if (pl[w] == DBL_MAX) continue;
        int c[3] = {check[3 * i], check[3 * i + 1], check[3 * i + 2]};
        const int ax = w / 2;
        c[ax] = (w % 2 == 0) ? (-1 - c[ax]) : (1 - c[ax]);
        int seen = 0;
        for (size_t k = 0; k < nall && !seen; k++)
          seen = all[3 * k] == c[0] && all[3 * k + 1] == c[1] && all[3 * k + 2] == c[2];
        if (seen || (w == 5 && c[2] == 1)) continue;
        if (nall == cap || nfresh == cap) {
          cap *= 2;
          all = (int*)realloc(all, sizeof(int) * 3 * cap);
          check = (int*)realloc(check, sizeof(int) * 3 * cap);
          fresh = (int*)realloc(fresh, sizeof(int) * 3 * cap);
        }
        memcpy(all + 3 * nall++, c, sizeof(c));
        memcpy(fresh + 3 * nfresh++, c, sizeof(c));
      }
    }
    int* tmp = check;
    check = fresh;
    fresh = tmp;
    ncheck = nfresh;
  }
  const size_t org = n->size;
  or_nodes src = *n;
  or_nodes dst = {0};
  for (size_t i = 0; i < nall; i++) {
    const int xr = all[3 * i], yr = all[3 * i + 1], zr = all[3 * i + 2];
    const int nx = abs(xr % 2), ny = abs(yr % 2), nz = abs(zr % 2);
    for (size_t k = 0; k < org; k++) {
      const double xb = (xr + nx) * sim->bc_xmax - (xr - nx) * sim->bc_xmin + (1 - 2 * nx) * src.xb[k];
      const double yb = (yr + ny) * sim->bc_ymax - (yr - ny) * sim->bc_ymin + (1 - 2 * ny) * src.yb[k];
      const double zb = (zr + nz) * sim->zmax - (zr - nz) * sim->bc_zmin + (1 - 2 * nz) * src.zb[k];
      if (dst.size == dst.cap) {
        dst.cap = dst.cap ? 2 * dst.cap : (org > 0 ? org * nall : 1);
        double** arrs[8] = {&dst.xb, &dst.yb, &dst.zb, &dst.phix, &dst.phiy, &dst.phiz, &dst.dtau, &dst.expmod};
        for (int a = 0; a < 8; a++) *arrs[a] = (double*)realloc(*arrs[a], sizeof(double) * dst.cap);
      }
      const size_t d = dst.size++;
      dst.xb[d] = xb; dst.yb[d] = yb; dst.zb[d] = zb;
      dst.phix[d] = src.phix[k]; dst.phiy[d] = src.phiy[k]; dst.phiz[d] = src.phiz[k];
      dst.dtau[d] = src.dtau[k]; dst.expmod[d] = src.expmod[k];
    }
  }
  or_nodes_free(&src);
  *n = dst;
  free(all);
  free(check);
  free(fresh);
}

/* Calc::Integrate_Serial (Calc.cpp:48-65): the compressed build, replaced by
 * the plain one when that has no more nodes (main builds only), then wall
 * images. out must be empty. The reference keeps one static start segment in
 * each of GaussIntegrate and GaussCompressIntegrate (Calc.cpp:70, :216); both
 * are written by the same main builds, so one hint array stands for both. */
void or_build_nodes(const or_sim* sim, int* hint, double t, int is_sol, or_nodes* out) {
  int* seg = (int*)malloc(sizeof(int) * (size_t)(sim->nbeams > 0 ? sim->nbeams : 1));
  for (int b = 0; b < sim->nbeams; b++) {
    const or_seg* P = sim->segs + sim->path_off[b];
    const int nseg = sim->path_off[b + 1] - sim->path_off[b];
    seg[b] = locate_seg(P, nseg, t, hint[b]);
    if (!is_sol) hint[b] = seg[b];
  }
  if (sim->compress) {
    for (int b = 0; b < sim->nbeams; b++) compress_beam(sim, b, seg[b], t, is_sol, out);
    if (!is_sol) {
      or_nodes reg = {0};
      for (int b = 0; b < sim->nbeams; b++) gauss_beam(sim, b, seg[b], t, is_sol, &reg);
      if (reg.size <= out->size) {
        or_nodes tmp = *out;
        *out = reg;
        reg = tmp;
      }
      or_nodes_free(&reg);
    }
  } else {
    for (int b = 0; b < sim->nbeams; b++) gauss_beam(sim, b, seg[b], t, is_sol, out);
  }
  free(seg);
  if (sim->use_bcs) add_bcs(sim, out);
}

/* ------------------------------------------------------------------------- */
/* per-point kernels                                                          */

/* Grid::get_x/get_y/get_z (Grid.h:310-327) with i,j,k from p (Util.h:30-31). */
void or_point_xyz(const or_sim* sim, int p, double* x, double* y, double* z) {
  if (sim->custom_points) {  /* Grid.cpp:17-27, Grid.h:310-327 */
    *x = sim->pts_x[p];
    *y = sim->pts_y[p];
    *z = sim->pts_z[p];
    return;
  }
  const int i = p / (sim->ynum * sim->znum);
  const int j = (p / sim->znum) % sim->ynum;
  const int k = p % sim->znum;
  const int xn = (uint16_t)sim->xnum, yn = (uint16_t)sim->ynum, zn = (uint16_t)sim->znum;
  *x = (xn - 1) ? sim->xmin + ((double)(uint16_t)i * ((sim->xmax - sim->xmin) / (xn - 1))) : sim->xmax;
  *y = (yn - 1) ? sim->ymin + ((double)(uint16_t)j * ((sim->ymax - sim->ymin) / (yn - 1))) : sim->ymax;
  *z = (zn - 1) ? sim->zmin + ((double)(uint16_t)k * ((sim->zmax - sim->zmin) / (zn - 1))) : sim->zmax;
}

/* Grid::Calc_T's node sum (Grid.cpp:286-296), without the T_init offset. */
static double sum_T(const or_nodes* n, double xp, double yp, double zp) {
  double dT = 0;
  for (size_t it = 0; it < n->size; it++) {
    const double dx = xp - n->xb[it];
    const double dy = yp - n->yb[it];
    const double dz = zp - n->zb[it];
    const double phi = exp(-3.0 * ((dx * dx * n->phix[it]) + (dy * dy * n->phiy[it]) + (dz * dz * n->phiz[it])) + n->expmod[it]);
    dT += n->dtau[it] * phi;
  }
  return dT;
}

double or_calc_T_xyz(const or_sim* sim, const or_nodes* nodes, double x, double y, double z) {
  return sim->T_init + sum_T(nodes, x, y, z);
}

void or_sol_primary_xyz(const or_nodes* n, double xp, double yp, double zp, double out[5]) {
  double gx = 0, gy = 0, gz = 0, lap = 0, s = 0;
  for (size_t it = 0; it < n->size; it++) {
    const double dx = xp - n->xb[it];
    const double dy = yp - n->yb[it];
    const double dz = zp - n->zb[it];
    const double phi = exp(-3.0 * ((dx * dx * n->phix[it]) + (dy * dy * n->phiy[it]) + (dz * dz * n->phiz[it])) + n->expmod[it]);
    const double dT_seg = n->dtau[it] * phi;
    const double ddpx = (-6.0 * n->phix[it]);
    const double ddpy = (-6.0 * n->phiy[it]);
    const double ddpz = (-6.0 * n->phiz[it]);
    const double dpx = ddpx * dx, dpy = ddpy * dy, dpz = ddpz * dz;
    gx += dT_seg * dpx;
    gy += dT_seg * dpy;
    gz += dT_seg * dpz;
    lap += dT_seg * (dpx * dpx + dpy * dpy + dpz * dpz + ddpx + ddpy + ddpz);
    s += (n->dtau[it] == 0) ? phi : 0;
  }
  out[0] = gx; out[1] = gy; out[2] = gz; out[3] = lap; out[4] = s;
}

long or_snapshot_T(const or_sim* sim, double t, double* T, int nthreads) {
  int* hint = (int*)malloc(sizeof(int) * (size_t)sim->nbeams);
  for (int b = 0; b < sim->nbeams; b++) hint[b] = 1;
  or_nodes nodes = {0};
  or_build_nodes(sim, hint, t, 0, &nodes);
  const int pnum = sim->pnum;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int p = 0; p < pnum; p++) {
    double x, y, z;
    or_point_xyz(sim, p, &x, &y, &z);
    T[p] = sim->T_init + sum_T(&nodes, x, y, z);
  }
  long n = (long)nodes.size;
  or_nodes_free(&nodes);
  free(hint);
  return n;
}

/* ------------------------------------------------------------------------- */
/* grid state and mode drivers                                                */

typedef struct { int* v; size_t n, cap; } ivec;
static void iv_push(ivec* a, int x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 256;
    a->v = (int*)realloc(a->v, sizeof(int) * a->cap);
  }
  a->v[a->n++] = x;
}
static void iv_copy(ivec* dst, const ivec* src) {
  dst->n = 0;
  for (size_t i = 0; i < src->n; i++) iv_push(dst, src->v[i]);
}

typedef struct { double* v; int n, cap; } dvec;
typedef struct { ivec pool; int seg; double sn, cs; } mp_pool;

typedef struct {
  or_sim* sim;
  int do_sol, has_flags;
  double *T, *T_last, *tSol, *G, *V, *Gx, *Gy, *Gz, *dTdt, *eqFrac, *depth;
  uint16_t* numMelt;
  unsigned char *calc_flag, *out_flag;
  int* hint;          /* main-step segment hint per beam */
  double evals;
  const char* full_dir;
  /* T_hist (Grid.h:100-101): the untracked drivers evaluate every point at
   * every step, so one time list serves all points (hist_T is step-major);
   * the tracked drivers record per point, in Calc_T's set branch
   * (Grid.cpp:300-304), at the time t_now of the step being evaluated */
  int hist_n, hist_cap;
  double *hist_t, *hist_T;
  dvec *hp_t, *hp_T;
  double t_now;
  /* RDF (Grid.h:103-105): per point melt, liquidus and cooling-rate lists */
  dvec *rdf_tm, *rdf_tl, *rdf_cr;
  /* MP_Stats (Grid.h:107-109) and Melt::calc_mp_info's function statics
   * (Melt.cpp:364-369): per-segment local pools, oldest first */
  double *mp_w, *mp_l, *mp_d;
  mp_pool* mp;
  int mp_n, mp_cap, mp_seg;
  /* RRDF records (Grid.h:111-113) */
  unsigned *rr_idx;
  double *rr_ts, *rr_Ts;
  size_t rr_n, rr_cap;
} or_grid;

static void dv_push(dvec* a, double x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 4;
    a->v = (double*)realloc(a->v, sizeof(double) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}
static void dv_free_all(dvec* a, int n) {
  if (!a) return;
  for (int i = 0; i < n; i++) free(a[i].v);
  free(a);
}

/* Grid::add_T_hist / add_t_hist (Grid.h:377-378) for a step in which every
 * point was evaluated with set. */
static void hist_record(or_grid* g, double t) {
  const or_sim* sim = g->sim;
  if (!sim->o_Thist) return;
  if (g->hist_n == g->hist_cap) {
    g->hist_cap = g->hist_cap ? 2 * g->hist_cap : 64;
    g->hist_t = (double*)realloc(g->hist_t, sizeof(double) * (size_t)g->hist_cap);
    g->hist_T = (double*)realloc(g->hist_T, sizeof(double) * (size_t)g->hist_cap * (size_t)(sim->pnum ? sim->pnum : 1));
  }
  g->hist_t[g->hist_n] = t;
  memcpy(g->hist_T + (size_t)g->hist_n * (size_t)sim->pnum, g->T, sizeof(double) * (size_t)sim->pnum);
  g->hist_n++;
}

static double* zalloc(or_sim* sim, size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* Grid::Grid (Grid.h:120-280): which arrays exist decides which setters act. */
static void grid_init(or_grid* g, or_sim* sim) {
  memset(g, 0, sizeof(*g));
  g->sim = sim;
  const size_t n = (size_t)sim->pnum;
  g->T = zalloc(sim, n);
  g->has_flags = (sim->mode == OR_MODE_SOLIDIFICATION) || sim->tracking != OR_TRACK_NONE;
  if (g->has_flags) {
    g->calc_flag = (unsigned char*)calloc(n, 1);
    g->out_flag = (unsigned char*)calloc(n, 1);
  }
  if (sim->mode == OR_MODE_SOLIDIFICATION) {
    g->T_last = zalloc(sim, n);
    if (sim->o_tSol) { g->do_sol = 1; g->tSol = zalloc(sim, n); }
    if (sim->o_G) { g->do_sol = 1; g->G = zalloc(sim, n); }
    if (sim->o_V) { g->do_sol = 1; g->V = zalloc(sim, n); }
    if (sim->o_Gx) { g->do_sol = 1; g->Gx = zalloc(sim, n); }
    if (sim->o_Gy) { g->do_sol = 1; g->Gy = zalloc(sim, n); }
    if (sim->o_Gz) { g->do_sol = 1; g->Gz = zalloc(sim, n); }
    if (sim->o_dTdt) { g->do_sol = 1; g->dTdt = zalloc(sim, n); }
    if (sim->o_eqFrac) { g->do_sol = 1; g->eqFrac = zalloc(sim, n); }
    if (sim->o_depth && sim->tracking == OR_TRACK_SURFACE) g->depth = zalloc(sim, n);
    /* with RDF, numMelt reports the number of recorded melt times (Grid.h:232, :343-350) */
    if (sim->o_numMelt && !sim->o_RDF) g->numMelt = (uint16_t*)calloc(n, sizeof(uint16_t));
    if (sim->o_RDF) {
      g->do_sol = 1;
      g->rdf_tm = (dvec*)calloc(n ? n : 1, sizeof(dvec));
      g->rdf_tl = (dvec*)calloc(n ? n : 1, sizeof(dvec));
      g->rdf_cr = (dvec*)calloc(n ? n : 1, sizeof(dvec));
    }
    if (sim->o_mpstats) {
      g->mp_w = zalloc(sim, n);
      g->mp_l = zalloc(sim, n);
      g->mp_d = zalloc(sim, n);
    }
  }
  if (sim->o_Thist && sim->tracking != OR_TRACK_NONE && !sim->custom_points) {
    g->hp_t = (dvec*)calloc(n ? n : 1, sizeof(dvec));
    g->hp_T = (dvec*)calloc(n ? n : 1, sizeof(dvec));
  }
  g->hint = (int*)malloc(sizeof(int) * (size_t)(sim->nbeams > 0 ? sim->nbeams : 1));
  for (int b = 0; b < sim->nbeams; b++) g->hint[b] = 1;
}

static void grid_free(or_grid* g) {
  free(g->T); free(g->T_last); free(g->tSol); free(g->G); free(g->V); free(g->Gx);
  free(g->Gy); free(g->Gz); free(g->dTdt); free(g->eqFrac); free(g->depth);
  free(g->numMelt); free(g->calc_flag); free(g->out_flag); free(g->hint);
  free(g->hist_t); free(g->hist_T);
  const int n = g->sim->pnum;
  dv_free_all(g->rdf_tm, n); dv_free_all(g->rdf_tl, n); dv_free_all(g->rdf_cr, n);
  dv_free_all(g->hp_t, n); dv_free_all(g->hp_T, n);
  free(g->mp_w); free(g->mp_l); free(g->mp_d);
  for (int i = 0; i < g->mp_n; i++) free(g->mp[i].pool.v);
  free(g->mp);
  free(g->rr_idx); free(g->rr_ts); free(g->rr_Ts);
}

static int ijk(const or_sim* s, int i, int j, int k) { return i * (s->znum * s->ynum) + j * s->znum + k; }

/* Grid::Calc_T (Grid.cpp:275-314) */
static double calc_T(or_grid* g, const or_nodes* nodes, int set, int p) {
  const or_sim* sim = g->sim;
  double x, y, z;
  or_point_xyz(sim, p, &x, &y, &z);
  const double T = sim->T_init + sum_T(nodes, x, y, z);
  g->evals += (double)nodes->size;
  if (set) {
    g->T[p] = T;
    if (g->hp_t) {  /* add_T_hist / add_t_hist (Grid.cpp:301-302) */
      dv_push(&g->hp_T[p], T);
      dv_push(&g->hp_t[p], g->t_now);
    }
    if (g->calc_flag) g->calc_flag[p] = 1;
    if (T >= sim->T_liq) {
      if (g->G) g->G[p] = DBL_MIN;
      if (g->V) g->V[p] = DBL_MIN;
      if (g->eqFrac) g->eqFrac[p] = -1.0;
      if (g->out_flag) g->out_flag[p] = 1;
    }
  }
  return T;
}

/* abs() at Grid.cpp:381/:535 under the chosen compiler semantics. */
static double ref_abs(const or_sim* sim, double x) {
  if (sim->abs_mode == OR_ABS_FABS) return fabs(x);
  int v;
  if (x != x || x >= 2147483648.0 || x <= -2147483649.0) v = INT_MIN;  /* x86 cvttsd2si */
  else v = (int)x;
  if (v == INT_MIN) return (double)INT_MIN;
  return (double)(v < 0 ? -v : v);
}

/* Grid::Calc_Solidification_time (Grid.cpp:340-404) */
static double sol_time(or_grid* g, double t, int p) {
  const or_sim* sim = g->sim;
  int run = 1, iter = 0;
  double T2 = g->T[p], T1 = g->T_last[p], T0 = sim->T_liq;
  double t2 = t, t1 = t - sim->dt, t0 = 0.0;
  int* hint = (int*)malloc(sizeof(int) * (size_t)sim->nbeams);
  or_nodes nodes = {0};
  while (run) {
    const double m = (T2 - T1) / (t2 - t1);
    t0 = t1 + ((T0 - T1) / m);
    or_nodes_clear(&nodes);
    memcpy(hint, g->hint, sizeof(int) * (size_t)sim->nbeams);
    or_build_nodes(sim, hint, t0, 0, &nodes);
    const double Tt = calc_T(g, &nodes, 0, p);
    const double err = ref_abs(sim, 1.0 - Tt / T0);
    iter++;
    if ((err < sim->sol_tol) || (iter >= sim->sol_iter)) run = 0;
    else if (Tt > T0) { t1 = t0; T1 = Tt; }
    else { t2 = t0; T2 = Tt; }
  }
  or_nodes_free(&nodes);
  free(hint);
  return t0;
}

/* Grid::Solidify + Set_Solidficiaton_Primary (Grid.cpp:316-338, :522-548) */
static void solidify(or_grid* g, double t, int p) {
  const or_sim* sim = g->sim;
  if (g->numMelt) g->numMelt[p] += 1;
  if (!g->do_sol) return;
  const double ts = sol_time(g, t, p);
  if (g->tSol) g->tSol[p] = ts;
  if (g->rdf_tl) dv_push(&g->rdf_tl[p], ts);  /* Grid.cpp:322 */
  int* hint = (int*)malloc(sizeof(int) * (size_t)sim->nbeams);
  memcpy(hint, g->hint, sizeof(int) * (size_t)sim->nbeams);
  or_nodes nodes = {0};
  or_build_nodes(sim, hint, ts, 1, &nodes);
  double x, y, z, pr[5];
  or_point_xyz(sim, p, &x, &y, &z);
  or_sol_primary_xyz(&nodes, x, y, z, pr);
  g->evals += (double)nodes.size;
  or_nodes_free(&nodes);
  free(hint);
  const double Gt = sqrt(pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2]);
  const double dTdt = ref_abs(sim, sim->a * pr[3] + pr[4]);
  const double V = dTdt / Gt;
  const double eq = (1.0 - exp((-4 * OR_PI * sim->cet_N0 / 3) *
                               pow(Gt * (sim->cet_n + 1) / (pow(sim->cet_a * V, 1 / sim->cet_n)), -3)));
  if (g->G) g->G[p] = Gt;
  if (g->Gx) g->Gx[p] = pr[0] / Gt;
  if (g->Gy) g->Gy[p] = pr[1] / Gt;
  if (g->Gz) g->Gz[p] = pr[2] / Gt;
  if (g->dTdt) g->dTdt[p] = dTdt;
  if (g->V) g->V[p] = V;
  if (g->eqFrac) g->eqFrac[p] = eq;
  if (g->rdf_cr) dv_push(&g->rdf_cr[p], dTdt);  /* Grid.cpp:547 */
}

/* Grid::Output (Grid.cpp:72-99) and a full-precision twin. */
static void grid_output(or_grid* g, const char* tag) {
  const or_sim* sim = g->sim;
  const char* names[16];
  int cols[16], nc = 0;
  enum { CX, CY, CZ, CT, CTS, CG, CV, CGX, CGY, CGZ, CDT, CEQ, CDEP, CNM, CMW, CML, CMD };
  if (sim->o_x) { names[nc] = "x"; cols[nc++] = CX; }
  if (sim->o_y) { names[nc] = "y"; cols[nc++] = CY; }
  if (sim->o_z) { names[nc] = "z"; cols[nc++] = CZ; }
  if (sim->o_T) { names[nc] = "T"; cols[nc++] = CT; }
  if (sim->mode == OR_MODE_SOLIDIFICATION) {
    if (sim->o_tSol) { names[nc] = "tSol"; cols[nc++] = CTS; }
    if (sim->o_G) { names[nc] = "G"; cols[nc++] = CG; }
    if (sim->o_V) { names[nc] = "V"; cols[nc++] = CV; }
    if (sim->o_Gx) { names[nc] = "Gx"; cols[nc++] = CGX; }
    if (sim->o_Gy) { names[nc] = "Gy"; cols[nc++] = CGY; }
    if (sim->o_Gz) { names[nc] = "Gz"; cols[nc++] = CGZ; }
    if (sim->o_dTdt) { names[nc] = "dTdt"; cols[nc++] = CDT; }
    if (sim->o_eqFrac) { names[nc] = "eqFrac"; cols[nc++] = CEQ; }
    if (sim->o_depth && sim->tracking == OR_TRACK_SURFACE) { names[nc] = "depth"; cols[nc++] = CDEP; }
    if (sim->o_numMelt) { names[nc] = "numMelt"; cols[nc++] = CNM; }
    if (sim->o_mpstats) {
      names[nc] = "MP_width"; cols[nc++] = CMW;
      names[nc] = "MP_length"; cols[nc++] = CML;
      names[nc] = "MP_depth"; cols[nc++] = CMD;
    }
  }
  if (nc == 0) return;
  char path[2048], fpath[2048];
  snprintf(path, sizeof path, "%s/%s.%s.csv", sim->data_dir, sim->name, tag);
  FILE* f = fopen(path, "w");
  FILE* ff = NULL;
  if (g->full_dir) {
    snprintf(fpath, sizeof fpath, "%s/%s.%s.full.txt", g->full_dir, sim->name, tag);
    ff = fopen(fpath, "w");
  }
  if (f) {
    fprintf(f, "%s", names[0]);
    for (int c = 1; c < nc; c++) fprintf(f, ",%s", names[c]);
    fprintf(f, "\n");
  }
  const double zstep = (sim->zmax - sim->zmin) / ((uint16_t)sim->znum - 1);
  for (int p = 0; p < sim->pnum; p++) {
    if (g->out_flag && !g->out_flag[p]) continue;
    if (ff) fprintf(ff, "%d", p);
    for (int c = 0; c < nc; c++) {
      double v = 0, x, y, z;
      int is_int = 0;
      switch (cols[c]) {
        case CX: or_point_xyz(sim, p, &v, &y, &z); break;
        case CY: or_point_xyz(sim, p, &x, &v, &z); break;
        case CZ: or_point_xyz(sim, p, &x, &y, &v); break;
        case CT: v = g->T[p]; break;
        case CTS: v = g->tSol[p]; break;
        case CG: v = g->G[p]; break;
        case CV: v = g->V[p]; break;
        case CGX: v = g->Gx[p]; break;
        case CGY: v = g->Gy[p]; break;
        case CGZ: v = g->Gz[p]; break;
        case CDT: v = g->dTdt[p]; break;
        case CEQ: v = g->eqFrac[p]; break;
        case CDEP: v = g->depth[p] * zstep; break;
        case CNM: v = g->numMelt ? g->numMelt[p] : (uint16_t)g->rdf_tm[p].n; is_int = 1; break;
        case CMW: v = g->mp_w[p]; break;
        case CML: v = g->mp_l[p]; break;
        case CMD: v = g->mp_d[p]; break;
      }
      if (f) {
        if (c) fputc(',', f);
        if (is_int) fprintf(f, "%d", (int)v);
        else fprintf(f, "%.6g", v);
      }
      if (ff) fprintf(ff, " %.17g", v);
    }
    if (f) fputc('\n', f);
    if (ff) fputc('\n', ff);
  }
  if (f) fclose(f);
  if (ff) fclose(ff);
}

static void zpad(char* buf, size_t cap, int num, int width) { snprintf(buf, cap, "%0*d", width, num); }

/* Run::Snapshots_NoTracking (Run.cpp:80-124) */
static void snapshots_none(or_grid* g) {
  or_sim* sim = g->sim;
  if (g->out_flag) memset(g->out_flag, 1, (size_t)sim->pnum);
  or_nodes nodes = {0};
  for (int s = 0; s < sim->nsnap; s++) {
    const double t = sim->snap_times[s];
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    const int pnum = sim->pnum;
    const int nth = sim->thnum > 0 ? sim->thnum : 1;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int p = 0; p < pnum; p++) {
      double x, y, z;
      or_point_xyz(sim, p, &x, &y, &z);
      const double T = sim->T_init + sum_T(&nodes, x, y, z);
      g->T[p] = T;
      if (g->calc_flag) g->calc_flag[p] = 1;
      if (T >= sim->T_liq && g->out_flag) g->out_flag[p] = 1;
    }
    g->evals += (double)nodes.size * pnum;
    hist_record(g, t);
    char tag[64], num[16];
    zpad(num, sizeof num, s, 2);
    snprintf(tag, sizeof tag, "Snapshot.%s", num);
    grid_output(g, tag);
  }
  or_nodes_free(&nodes);
}

/* Melt::beam_trace + beam_trace_perimeter (Melt.cpp:20-183) */
static void beam_trace(or_grid* g, ivec* pts, double t_start, double t_end) {
  const or_sim* sim = g->sim;
  const int zk = sim->znum - 1;
  /* perimeter seeding, evaluated at the end time only (Melt.cpp:20-96) */
  for (int b = 0; b < sim->nbeams; b++) {
    const or_seg* P = sim->segs + sim->path_off[b];
    const int nseg = sim->path_off[b + 1] - sim->path_off[b];
    const or_beam* bm = &sim->beams[b];
    const double rc = sim->radius_check;
    const double t_sub = (rc * rc - 1) * (bm->ax * bm->ax) / (12.0 * sim->a);
    const double ts = (t_end - t_sub < 0.0) ? 0.0 : t_end - t_sub;
    const double rc2 = rc * rc * bm->ax * bm->ax;
    int s0 = 0, s1 = 0;
    while (P[s0].seg_time < ts && (s0 + 1) != nseg) s0++;
    while (P[s1].seg_time < t_end && (s1 + 1) != nseg) s1++;
#define PERIM(FIXED, XFIXED, LIM, X, Y)                                        \
    for (int var = 0; var < (LIM); ++var) {                                    \
      int ii = (XFIXED) ? (FIXED) : var, jj = (XFIXED) ? var : (FIXED);        \
      double xp = sim->xmin + ii * sim->xres, yp = sim->ymin + jj * sim->yres; \
      double ddx = xp - (X), ddy = yp - (Y);                                   \
      if (ddx * ddx + ddy * ddy < rc2) {                                       \
        int p = ijk(sim, ii, jj, zk);                                          \
        if (!g->calc_flag[p]) { iv_push(pts, p); g->calc_flag[p] = 1; }        \
      }                                                                        \
    }
#define PERIM_ALL(X, Y)                                                                   \
    do {                                                                                  \
      int xg = (int)floor(((X) - sim->xmin) / sim->xres);                                 \
      int yg = (int)floor(((Y) - sim->ymin) / sim->yres);                                 \
      if (xg < 0) { PERIM(0, 1, sim->ynum, X, Y) }                                        \
      if (xg > sim->xnum - 1) { PERIM(sim->xnum - 1, 1, sim->ynum, X, Y) }                \
      if (yg < 0) { PERIM(0, 0, sim->xnum, X, Y) }                                        \
      if (yg > sim->ynum - 1) { PERIM(sim->ynum - 1, 0, sim->xnum, X, Y) }                \
    } while (0)
    for (int s = (s0 - 1 > 0 ? s0 - 1 : 0); s <= s1; s++) PERIM_ALL(P[s].sx, P[s].sy);
    if (P[s1].smode == 0 && t_end < P[nseg - 1].seg_time) {
      double xb, yb, zb, qm;
      beam_loc(sim, P, s1, t_end, &xb, &yb, &zb, &qm);
      PERIM_ALL(xb, yb);
    }
#undef PERIM_ALL
#undef PERIM
  }
  /* vertex and current-position seeding (Melt.cpp:98-183) */
  for (int b = 0; b < sim->nbeams; b++) {
    const or_seg* P = sim->segs + sim->path_off[b];
    const int nseg = sim->path_off[b + 1] - sim->path_off[b];
    int s0 = 0, s1 = 0;
    while (P[s0].seg_time < t_start && (s0 + 1) != nseg) s0++;
    while (P[s1].seg_time < t_end && (s1 + 1) != nseg) s1++;
#define SEED4(X, Y)                                                              \
    do {                                                                         \
      int xg = (int)floor(((X) - sim->xmin) / sim->xres);                        \
      int yg = (int)floor(((Y) - sim->ymin) / sim->yres);                        \
      if (xg < 0 || xg > sim->xnum - 1 || yg < 0 || yg > sim->ynum - 1) break;   \
      for (int ddx = 0; ddx <= 1; ddx++)                                         \
        for (int ddy = 0; ddy <= 1; ddy++) {                                     \
          int xx = xg + ddx, yy = yg + ddy;                                      \
          if (xx < 0 || xx > sim->xnum - 1 || yy < 0 || yy > sim->ynum - 1) continue; \
          int p = ijk(sim, xx, yy, zk);                                          \
          if (!g->calc_flag[p]) { iv_push(pts, p); g->calc_flag[p] = 1; }        \
        }                                                                        \
    } while (0)
    for (int s = (s0 - 1 > 0 ? s0 - 1 : 0); s <= s1; s++) SEED4(P[s].sx, P[s].sy);
    if (P[s1].smode == 0 && t_end < P[nseg - 1].seg_time) {
      double xb, yb, zb, qm;
      beam_loc(sim, P, s1, t_end, &xb, &yb, &zb, &qm);
      SEED4(xb, yb);
    }
#undef SEED4
  }
}

/* Melt::neighbor_check (Melt.cpp:185-269), one BFS wave, sequential. */
static void neighbor_check(or_grid* g, ivec* test, ivec* liq, ivec* reset, const or_nodes* nodes, int surface) {
  const or_sim* sim = g->sim;
  ivec next = {0};
  for (size_t it = 0; it < test->n; it++) {
    const int p = test->v[it];
    const int i = p / (sim->ynum * sim->znum);
    const int j = (p / sim->znum) % sim->ynum;
    const int k = p % sim->znum;
    const int dk0 = surface ? 0 : -1;
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++)
        for (int dk = dk0; dk < 1; dk++) {
          const int ni = i + di, nj = j + dj, nk = k + dk;
          if (ni < 0 || ni > sim->xnum - 1 || nj < 0 || nj > sim->ynum - 1 || nk < 0) continue;
          const int q = ijk(sim, ni, nj, nk);
          if (g->calc_flag[q]) continue;
          g->calc_flag[q] = 1;
          if (calc_T(g, nodes, 1, q) >= sim->T_liq) {
            iv_push(liq, q);
            iv_push(&next, q);
          }
          iv_push(reset, q);
        }
  }
  test->n = 0;
  iv_copy(test, &next);
  free(next.v);
}

/* Run::Snapshots_Volume (Run.cpp:126-193); Surface tracking aliases it. */
static void snapshots_volume(or_grid* g) {
  or_sim* sim = g->sim;
  ivec liq = {0}, reset = {0}, test = {0}, seeds = {0};
  or_nodes nodes = {0};
  for (int s = 0; s < sim->nsnap; s++) {
    for (size_t r = 0; r < reset.n; r++) {
      g->out_flag[reset.v[r]] = 0;
      g->calc_flag[reset.v[r]] = 0;
    }
    reset.n = 0;
    const double t = sim->snap_times[s];
    g->t_now = t;
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    seeds.n = 0;
    beam_trace(g, &seeds, 0, t);
    iv_copy(&test, &liq);
    for (size_t it = 0; it < seeds.n; it++) {
      const int p = seeds.v[it];
      iv_push(&reset, p);
      if (calc_T(g, &nodes, 1, p) >= sim->T_liq) {
        iv_push(&test, p);
        iv_push(&liq, p);
      }
    }
    while (test.n) neighbor_check(g, &test, &liq, &reset, &nodes, 0);
    char tag[64], num[16];
    zpad(num, sizeof num, s, 2);
    snprintf(tag, sizeof tag, "Snapshot.%s", num);
    grid_output(g, tag);
  }
  or_nodes_free(&nodes);
  free(liq.v); free(reset.v); free(test.v); free(seeds.v);
}

static void step_output(or_grid* g, const char* prefix, int itert) {
  char tag[64], num[16];
  zpad(num, sizeof num, itert, 7);
  snprintf(tag, sizeof tag, "%s%s", prefix, num);
  grid_output(g, tag);
}

/* Run::Solidify_NoTracking (Run.cpp:409-467) */
static void solidify_none(or_grid* g) {
  or_sim* sim = g->sim;
  int itert = 0;
  double t = 0.0;
  or_nodes nodes = {0};
  for (int p = 0; p < sim->pnum; p++) g->T[p] = sim->T_init;
  for (;;) {
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    for (int p = 0; p < sim->pnum; p++) g->T_last[p] = g->T[p];
    int liq = 0;
    for (int p = 0; p < sim->pnum; p++) {
      if (!(calc_T(g, &nodes, 1, p) >= sim->T_liq)) {
        if (g->T_last[p] >= sim->T_liq) solidify(g, t, p);
      } else {
        liq++;
      }
    }
    hist_record(g, t);
    if (itert && (itert % sim->out_freq == 0)) step_output(g, "", itert);
    if (t > sim->all_end && liq == 0) {
      if (sim->out_freq != INT_MAX) {
        itert = ((itert / sim->out_freq) + 1) * sim->out_freq;
        step_output(g, "", itert);
      }
      break;
    }
    t += sim->dt;
    itert++;
  }
  or_nodes_free(&nodes);
}

/* Run::Solidify_Volume (Run.cpp:469-573) */
static void solidify_volume(or_grid* g) {
  or_sim* sim = g->sim;
  ivec liq = {0}, reset = {0}, last = {0}, test = {0}, seeds = {0};
  or_nodes nodes = {0};
  double t = 0.0;
  int itert = 0;
  for (;;) {
    g->t_now = t;
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    for (size_t r = 0; r < reset.n; r++) {
      const int p = reset.v[r];
      g->T_last[p] = g->T[p];
      g->calc_flag[p] = 0;
    }
    reset.n = 0;
    iv_copy(&last, &liq);
    iv_copy(&reset, &liq);
    liq.n = 0;
    for (size_t it = 0; it < last.n; it++) {
      const int p = last.v[it];
      if (!(calc_T(g, &nodes, 1, p) >= sim->T_liq)) solidify(g, t, p);
      else iv_push(&liq, p);
    }
    iv_copy(&test, &liq);
    seeds.n = 0;
    if (itert && t < sim->all_end + sim->dt) beam_trace(g, &seeds, t - sim->dt, t);
    for (size_t it = 0; it < seeds.n; it++) {
      const int p = seeds.v[it];
      iv_push(&reset, p);
      if (calc_T(g, &nodes, 1, p) >= sim->T_liq) {
        iv_push(&test, p);
        iv_push(&liq, p);
      }
    }
    while (test.n) neighbor_check(g, &test, &liq, &reset, &nodes, 0);
    if (itert && (itert % sim->out_freq == 0)) step_output(g, "", itert);
    if (t > sim->all_end && liq.n == 0) {
      if (sim->out_freq != INT_MAX) {
        itert = ((itert / sim->out_freq) + 1) * sim->out_freq;
        step_output(g, "", itert);
      }
      break;
    }
    t += sim->dt;
    itert++;
  }
  or_nodes_free(&nodes);
  free(liq.v); free(reset.v); free(last.v); free(test.v); free(seeds.v);
}

/* Melt::calc_depth (Melt.cpp:271-317) */
static void calc_depth(or_grid* g, int* depths, ivec* liq, ivec* reset, const or_nodes* nodes,
                       const or_nodes* nodes_last, double t) {
  const or_sim* sim = g->sim;
  if (sim->znum == 1) return;
  for (size_t it = 0; it < liq->n; it++) {
    const int p = liq->v[it];
    const int i = p / (sim->ynum * sim->znum);
    const int j = (p / sim->znum) % sim->ynum;
    const int dnum = i * sim->ynum + j;
    int depth = depths[dnum];
    int flag = 1;
    for (;;) {
      depth++;
      if (depth == sim->znum) break;
      const int q = ijk(sim, i, j, sim->znum - 1 - depth);
      iv_push(reset, q);
      if (!(calc_T(g, nodes, 1, q) >= sim->T_liq)) break;
      flag = 0;
    }
    for (;;) {
      depth--;
      if (depth == 0 || flag == 0) break;
      const int q = ijk(sim, i, j, sim->znum - 1 - depth);
      iv_push(reset, q);
      if (!(calc_T(g, nodes, 1, q) >= sim->T_liq)) {
        if (depth != depths[dnum]) g->T_last[q] = calc_T(g, nodes_last, 0, q);
        solidify(g, t, q);
      } else {
        break;
      }
    }
    if (depth == sim->znum) depth--;
    depths[dnum] = depth;
  }
}

/* Melt::calc_depth_max (Melt.cpp:319-357) */
static void calc_depth_max(or_grid* g, const int* depths, double* depth_max, const ivec* liq) {
  const or_sim* sim = g->sim;
  if (sim->znum == 1 || sim->o_depth == 0) return;
  for (size_t it = 0; it < liq->n; it++) {
    const int p = liq->v[it];
    const int i = p / (sim->ynum * sim->znum);
    const int j = (p / sim->znum) % sim->ynum;
    const int dnum = i * sim->ynum + j;
    const int dl = depths[dnum];
    double depth;
    if (dl == sim->znum - 1) {
      depth = dl;
    } else {
      const int pl = ijk(sim, i, j, sim->znum - 1 - dl);
      const int ps = ijk(sim, i, j, sim->znum - 1 - (dl + 1));
      depth = dl + (sim->T_liq - g->T[pl]) / (g->T[ps] - g->T[pl]);
    }
    if (depth > depth_max[dnum]) {
      depth_max[dnum] = depth;
      for (int d = dl; d >= 0; d--)
        if (g->depth) g->depth[ijk(sim, i, j, sim->znum - 1 - d)] = depth;
    }
  }
}

/* 2x2 surface points around the grid cell holding (x, y) (Melt.cpp:384-409);
 * the y clamp compares with xnum as the reference does. */
static void mp_cell_points(const or_sim* sim, double x, double y, ivec* out) {
  int xg = (int)floor((x - sim->xmin) / sim->xres);
  int yg = (int)floor((y - sim->ymin) / sim->yres);
  const int zg = sim->znum - 1;
  if (xg < 0) xg = -1;
  if (xg > sim->xnum - 1) xg = sim->xnum - 1;
  if (yg < 0) yg = -1;
  if (yg > sim->xnum - 1) yg = sim->ynum - 1;
  for (int dx = 0; dx <= 1; dx++)
    for (int dy = 0; dy <= 1; dy++) {
      const int xx = xg + dx, yy = yg + dy;
      if (xx < 0 || xx > sim->xnum - 1 || yy < 0 || yy > sim->ynum - 1) continue;
      iv_push(out, zg + sim->znum * yy + sim->znum * sim->ynum * xx);
    }
}

/* Melt::local_neighbor_check (Melt.cpp:540-580): claim (clear the calc flag
 * of) the liquid, evaluated 3x3 neighbours at the same layer. */
static void mp_neighbor_check(or_grid* g, ivec* test, ivec* liq) {
  const or_sim* sim = g->sim;
  ivec next = {0};
  for (size_t it = 0; it < test->n; it++) {
    const int p = test->v[it];
    const int i = p / (sim->ynum * sim->znum);
    const int j = (p / sim->znum) % sim->ynum;
    const int k = p % sim->znum;
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++) {
        const int ni = i + di, nj = j + dj;
        if (ni < 0 || ni > sim->xnum - 1 || nj < 0 || nj > sim->ynum - 1) continue;
        const int q = ijk(sim, ni, nj, k);
        if (g->calc_flag[q] && g->T[q] >= sim->T_liq) {
          g->calc_flag[q] = 0;
          iv_push(&next, q);
          iv_push(liq, q);
        }
      }
  }
  free(test->v);
  *test = next;
}

/* Melt::calc_mp_info (Melt.cpp:359-538): width and length of each local
 * pool in the frame of its path segment (first beam's path), depth from the
 * deepest column of the whole grid times the x spacing; running maxima on
 * every point of the pool's columns. */
static void calc_mp_info(or_grid* g, const int* depths, double t) {
  const or_sim* sim = g->sim;
  if (!sim->o_mpstats) return;
  const or_seg* path = sim->segs + sim->path_off[0];
  const int npath = sim->path_off[1] - sim->path_off[0];
  while ((t > path[g->mp_seg].seg_time) && (g->mp_seg + 1 < npath)) {
    const int seg = ++g->mp_seg;
    if (path[seg].sqmod > 0) {
      const int sp = path[seg].smode == 0 ? seg - 1 : seg;
      if (g->mp_n == g->mp_cap) {
        g->mp_cap = g->mp_cap ? 2 * g->mp_cap : 16;
        g->mp = (mp_pool*)realloc(g->mp, sizeof(mp_pool) * (size_t)g->mp_cap);
      }
      mp_pool* e = &g->mp[g->mp_n++];
      memset(e, 0, sizeof(*e));
      mp_cell_points(sim, path[sp].sx, path[sp].sy, &e->pool);
      e->seg = seg;
      double angle;
      if (path[seg].smode == 1) angle = 0.0;
      else angle = atan2(path[seg].sy - path[seg - 1].sy, path[seg].sx - path[seg - 1].sx);
      e->sn = sin(angle);
      e->cs = cos(angle);
    }
  }
  if (!g->mp_n) return;
  const int seg = g->mp_seg;
  if (g->mp[g->mp_n - 1].seg == seg && path[seg].smode == 0) {
    double xb, yb, zb, qm;
    beam_loc(sim, path, seg, t, &xb, &yb, &zb, &qm);
    mp_cell_points(sim, xb, yb, &g->mp[g->mp_n - 1].pool);
  }
  const size_t ncol = (size_t)sim->xnum * (size_t)sim->ynum;
  int dmax = depths[0];
  for (size_t c = 1; c < ncol; c++)
    if (dmax < depths[c]) dmax = depths[c];
  for (int e = g->mp_n - 1; e >= 0; e--) {
    mp_pool* P = &g->mp[e];
    ivec liq = {0};
    while (P->pool.n) mp_neighbor_check(g, &P->pool, &liq);
    if (!liq.n) {
      free(P->pool.v);
      free(liq.v);
      memmove(g->mp + e, g->mp + e + 1, sizeof(mp_pool) * (size_t)(g->mp_n - e - 1));
      g->mp_n--;
      continue;
    }
    free(P->pool.v);
    P->pool = liq;
    double minX = DBL_MAX, maxX = -DBL_MAX, minY = DBL_MAX, maxY = -DBL_MAX;
    for (size_t it = 0; it < liq.n; it++) {
      double x, y, z;
      or_point_xyz(sim, liq.v[it], &x, &y, &z);
      const double xr = x * P->cs + y * P->sn;
      minX = (minX < xr) ? minX : xr;  /* std::min(x_rot, minRotX) */
      maxX = (xr < maxX) ? maxX : xr;  /* std::max(x_rot, maxRotX) */
      const double yr = -x * P->sn + y * P->cs;
      minY = (minY < yr) ? minY : yr;
      maxY = (yr < maxY) ? maxY : yr;
    }
    const double width = maxY - minY, length = maxX - minX;
    const double depth = sim->xres * dmax;
    for (size_t it = 0; it < liq.n; it++) {
      const int p = liq.v[it];
      const int i = p / (sim->ynum * sim->znum);
      const int j = (p / sim->znum) % sim->ynum;
      const int dnum = i * sim->ynum + j;
      for (int d = 0; d <= depths[dnum]; d++) {
        const int q = ijk(sim, i, j, sim->znum - 1 - d);
        g->mp_w[q] = (width < g->mp_w[q]) ? g->mp_w[q] : width;  /* std::max(new, old) */
        g->mp_l[q] = (length < g->mp_l[q]) ? g->mp_l[q] : length;
        g->mp_d[q] = (depth < g->mp_d[q]) ? g->mp_d[q] : depth;
      }
    }
  }
}

/* Run::Stork (Run.cpp:711-1079): Surface-style melt-pool search without
 * solidification; liquid columns mark their points, and every grid cell whose
 * count of liquid corners changes (or is partial) is recorded with the
 * temperatures of its 8 corners at t - dt and t (cell order c ascending, as
 * one thread produces it). */
static void stork(or_grid* g) {
  or_sim* sim = g->sim;
  const int np = sim->pnum;
  const size_t ncol = (size_t)sim->xnum * (size_t)sim->ynum;
  int* depths = (int*)calloc(ncol ? ncol : 1, sizeof(int));
  const int cx = sim->xnum - 1, cy = sim->ynum - 1, cz = sim->znum - 1;
  const long cn = (long)cx * cy * cz > 0 ? (long)cx * cy * cz : 0;
  unsigned char* c_a = (unsigned char*)calloc((size_t)(cn ? cn : 1), 1);
  unsigned char* c_b = (unsigned char*)calloc((size_t)(cn ? cn : 1), 1);
  double* T_a = (double*)malloc(sizeof(double) * (size_t)(np ? np : 1));
  double* T_b = (double*)malloc(sizeof(double) * (size_t)(np ? np : 1));
  unsigned char* f_a = (unsigned char*)malloc((size_t)(np ? np : 1));
  unsigned char* f_b = (unsigned char*)malloc((size_t)(np ? np : 1));
  unsigned char* is_liq = (unsigned char*)calloc((size_t)(np ? np : 1), 1);
  for (int p = 0; p < np; p++) { T_a[p] = T_b[p] = sim->T_init; f_a[p] = f_b[p] = 1; }
  ivec liq = {0}, reset = {0}, last = {0}, test = {0}, seeds = {0}, next = {0};
  or_nodes nodes = {0}, nodes_last = {0};
  double t = 0.0;
  int itert = 0;
  for (;;) {
    unsigned char* c_prev = (itert % 2) ? c_a : c_b;
    unsigned char* c_cur = (itert % 2) ? c_b : c_a;
    double* T_prev = (itert % 2) ? T_a : T_b;
    double* T_cur = (itert % 2) ? T_b : T_a;
    unsigned char* f_prev = (itert % 2) ? f_a : f_b;
    unsigned char* f_cur = (itert % 2) ? f_b : f_a;
    memset(c_cur, 0, (size_t)(cn ? cn : 1));
    memset(f_cur, 0, (size_t)(np ? np : 1));
    memset(is_liq, 0, (size_t)(np ? np : 1));
    for (size_t r = 0; r < reset.n; r++) {
      const int p = reset.v[r];
      g->T_last[p] = g->T[p];
      g->calc_flag[p] = 0;
    }
    reset.n = 0;
    iv_copy(&last, &liq);
    liq.n = 0;
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    for (size_t it = 0; it < last.n; it++) {
      const int p = last.v[it];
      T_cur[p] = calc_T(g, &nodes, 1, p);
      f_cur[p] = 1;
      iv_push(&reset, p);
      if (T_cur[p] >= sim->T_liq) {
        iv_push(&liq, p);
      } else {
        const int i = p / (sim->ynum * sim->znum);
        const int j = (p / sim->znum) % sim->ynum;
        depths[i * sim->ynum + j] = 0;
      }
    }
    iv_copy(&test, &liq);
    seeds.n = 0;
    if (itert && t < sim->all_end + sim->dt) beam_trace(g, &seeds, t - sim->dt, t);
    for (size_t it = 0; it < seeds.n; it++) {
      const int p = seeds.v[it];
      iv_push(&reset, p);
      if (f_cur[p] == 0) {
        T_cur[p] = calc_T(g, &nodes, 1, p);
        f_cur[p] = 1;
        if (T_cur[p] >= sim->T_liq) {
          iv_push(&test, p);
          iv_push(&liq, p);
        }
      }
    }
    while (test.n) {
      next.n = 0;
      for (size_t it = 0; it < test.n; it++) {
        const int p = test.v[it];
        const int i = p / (sim->ynum * sim->znum);
        const int j = (p / sim->znum) % sim->ynum;
        const int k = p % sim->znum;
        for (int di = 0; di < 3; di++)
          for (int dj = 0; dj < 3; dj++) {
            const int ni = i + di - 1, nj = j + dj - 1;
            if (ni < 0 || ni > sim->xnum - 1 || nj < 0 || nj > sim->ynum - 1) continue;
            const int q = ijk(sim, ni, nj, k);
            if (g->calc_flag[q]) continue;
            g->calc_flag[q] = 1;
            T_cur[q] = calc_T(g, &nodes, 1, q);
            f_cur[q] = 1;
            if (T_cur[q] >= sim->T_liq) {
              iv_push(&liq, q);
              iv_push(&next, q);
            }
            iv_push(&reset, q);
          }
      }
      iv_copy(&test, &next);
    }
    for (size_t it = 0; it < liq.n; it++) {
      const int p = liq.v[it];
      const int i = p / (sim->ynum * sim->znum);
      const int j = (p / sim->znum) % sim->ynum;
      const int dnum = i * sim->ynum + j;
      int depth = depths[dnum];
      int deeper = 0;
      for (;;) {
        depth++;
        if (depth == sim->znum) break;  /* the reference prints WARNING::MAXIMUM EXCEEDED */
        const int q = ijk(sim, i, j, sim->znum - 1 - depth);
        T_cur[q] = calc_T(g, &nodes, 0, q);
        f_cur[q] = 1;
        if (!(T_cur[q] >= sim->T_liq)) break;
        deeper = 1;
      }
      for (;;) {
        depth--;
        if (depth == 0 || deeper) break;
        const int q = ijk(sim, i, j, sim->znum - 1 - depth);
        T_cur[q] = calc_T(g, &nodes, 0, q);
        f_cur[q] = 1;
        if (T_cur[q] >= sim->T_liq) break;
      }
      if (depth == sim->znum - 1) depth--;
      depths[dnum] = depth;
      for (int d = 0; d <= depth; d++) is_liq[ijk(sim, i, j, sim->znum - 1 - d)] = 1;
    }
    if (last.n || liq.n) {
      for (long c = 0; c < cn; c++) {
        const int ci = (int)((c / cy) / cz), cj = (int)((c / cz) % cy), ck = (int)(c % cz);
        for (int di = 0; di < 2; di++)
          for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) c_cur[c] += is_liq[ijk(sim, ci + di, cj + dj, ck + dk)];
        if (!(c_cur[c] != c_prev[c] || c_prev[c] % 8 || c_cur[c] % 8)) continue;
        if (g->rr_n == g->rr_cap) {
          g->rr_cap = g->rr_cap ? 2 * g->rr_cap : 1024;
          g->rr_idx = (unsigned*)realloc(g->rr_idx, sizeof(unsigned) * 3 * g->rr_cap);
          g->rr_ts = (double*)realloc(g->rr_ts, sizeof(double) * 2 * g->rr_cap);
          g->rr_Ts = (double*)realloc(g->rr_Ts, sizeof(double) * 16 * g->rr_cap);
        }
        const size_t r = g->rr_n++;
        g->rr_idx[3 * r] = (unsigned)ci;
        g->rr_idx[3 * r + 1] = (unsigned)cj;
        g->rr_idx[3 * r + 2] = (unsigned)ck;
        g->rr_ts[2 * r] = t - sim->dt;
        g->rr_ts[2 * r + 1] = t;
        int n = 0;
        for (int di = 0; di < 2; di++)
          for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) {
              const int p = ijk(sim, ci + di, cj + dj, ck + dk);
              if (f_prev[p] == 0) { f_prev[p] = 1; T_prev[p] = calc_T(g, &nodes_last, 0, p); }
              if (f_cur[p] == 0) { f_cur[p] = 1; T_cur[p] = calc_T(g, &nodes, 0, p); }
              g->rr_Ts[16 * r + n] = T_prev[p];
              g->rr_Ts[16 * r + 8 + n] = T_cur[p];
              n++;
            }
      }
    }
    if (t > sim->all_end && liq.n == 0) break;
    t += sim->dt;
    itert++;
    or_nodes tmp = nodes_last;
    nodes_last = nodes;
    nodes = tmp;
  }
  or_nodes_free(&nodes);
  or_nodes_free(&nodes_last);
  free(depths); free(c_a); free(c_b); free(T_a); free(T_b); free(f_a); free(f_b); free(is_liq);
  free(liq.v); free(reset.v); free(last.v); free(test.v); free(seeds.v); free(next.v);
}

/* Run::Solidify_Surface (Run.cpp:575-709) */
static void solidify_surface(or_grid* g) {
  or_sim* sim = g->sim;
  ivec liq = {0}, reset = {0}, last = {0}, test = {0}, seeds = {0};
  const size_t ncol = (size_t)sim->xnum * (size_t)sim->ynum;
  int* depths = (int*)calloc(ncol, sizeof(int));
  double* depth_max = (double*)calloc(ncol, sizeof(double));
  or_nodes nodes = {0}, nodes_last = {0};
  double t = 0.0;
  int itert = 0;
  for (;;) {
    g->t_now = t;
    or_nodes_clear(&nodes);
    or_build_nodes(sim, g->hint, t, 0, &nodes);
    for (size_t r = 0; r < reset.n; r++) {
      const int p = reset.v[r];
      /* Run.cpp:607, Grid.h:400-404: a melt time opens when none is open */
      if (g->rdf_tm && g->T[p] >= sim->T_liq && g->rdf_tm[p].n == g->rdf_tl[p].n) dv_push(&g->rdf_tm[p], t - sim->dt);
      g->T_last[p] = g->T[p];
      g->calc_flag[p] = 0;
    }
    reset.n = 0;
    iv_copy(&last, &liq);
    iv_copy(&reset, &liq);
    liq.n = 0;
    for (size_t it = 0; it < last.n; it++) {
      const int p = last.v[it];
      if (!(calc_T(g, &nodes, 1, p) >= sim->T_liq)) {
        solidify(g, t, p);
        const int i = p / (sim->ynum * sim->znum);
        const int j = (p / sim->znum) % sim->ynum;
        const int dnum = i * sim->ynum + j;
        const int depth = depths[dnum];
        for (int d = depth; d > 0; d--) {
          const int q = ijk(sim, i, j, sim->znum - 1 - d);
          calc_T(g, &nodes, 1, q);
          if (d != depth) g->T_last[q] = calc_T(g, &nodes_last, 0, q);
          solidify(g, t, q);
          iv_push(&reset, q);
        }
        depths[dnum] = 0;
      } else {
        iv_push(&liq, p);
      }
    }
    iv_copy(&test, &last);
    seeds.n = 0;
    if (itert && t < sim->all_end + sim->dt) beam_trace(g, &seeds, t - sim->dt, t);
    for (size_t it = 0; it < seeds.n; it++) {
      const int p = seeds.v[it];
      iv_push(&reset, p);
      if (calc_T(g, &nodes, 1, p) >= sim->T_liq) {
        iv_push(&test, p);
        iv_push(&liq, p);
      }
    }
    while (test.n) neighbor_check(g, &test, &liq, &reset, &nodes, 1);
    calc_depth(g, depths, &liq, &reset, &nodes, &nodes_last, t);
    calc_depth_max(g, depths, depth_max, &liq);
    calc_mp_info(g, depths, t);
    if (itert && (itert % sim->out_freq == 0)) step_output(g, "Solidification.", itert);
    if (t > sim->all_end && liq.n == 0) {
      if (sim->out_freq != INT_MAX) {
        itert = ((itert / sim->out_freq) + 1) * sim->out_freq;
        step_output(g, "Solidification.", itert);
      }
      break;
    }
    t += sim->dt;
    itert++;
    /* nodes_last = nodes */
    or_nodes tmp = nodes_last;
    nodes_last = nodes;
    nodes = tmp;
  }
  or_nodes_free(&nodes);
  or_nodes_free(&nodes_last);
  free(depths); free(depth_max);
  free(liq.v); free(reset.v); free(last.v); free(test.v); free(seeds.v);
}

/* Grid::Output_T_hist (Grid.cpp:101-174). With the tracking keyword None:
 * one table "<Name>T.hist.csv" (no dot after the name), a column per recorded
 * time of point 0, times and temperatures through std::to_string ("%f"),
 * coordinates through operator<< (6 significant digits). Otherwise
 * "<Name>.Points.csv" (p,x,y,z of the output-flagged points) and one
 * "<Name>.T_hist.<p>.csv" (t,T) per flagged point. */
static void thist_output(or_grid* g) {
  const or_sim* sim = g->sim;
  char path[2048];
  const int np = sim->pnum;
  if (sim->tracking == OR_TRACK_NONE) {
    snprintf(path, sizeof path, "%s/%sT.hist.csv", sim->data_dir, sim->name);
    FILE* f = fopen(path, "w");
    if (f) {
      fprintf(f, "x,y,z");
      for (int e = 0; e < g->hist_n; e++) fprintf(f, ",%f", g->hist_t[e]);
      fprintf(f, "\n");
      for (int p = 0; p < np; p++) {
        double x, y, z;
        or_point_xyz(sim, p, &x, &y, &z);
        fprintf(f, "%.6g,%.6g,%.6g", x, y, z);
        for (int e = 0; e < g->hist_n; e++) fprintf(f, ",%f", g->hist_T[(size_t)e * np + p]);
        fprintf(f, "\n");
      }
      fclose(f);
    }
  } else {
    snprintf(path, sizeof path, "%s/%s.Points.csv", sim->data_dir, sim->name);
    FILE* f = fopen(path, "w");
    if (f) {
      fprintf(f, "p,x,y,z\n");
      for (int p = 0; p < np; p++) {
        if (g->out_flag && !g->out_flag[p]) continue;
        double x, y, z;
        or_point_xyz(sim, p, &x, &y, &z);
        fprintf(f, "%d,%.6g,%.6g,%.6g\n", p, x, y, z);
      }
      fclose(f);
    }
    for (int p = 0; p < np; p++) {
      if (g->out_flag && !g->out_flag[p]) continue;
      snprintf(path, sizeof path, "%s/%s.T_hist.%d.csv", sim->data_dir, sim->name, p);
      FILE* h = fopen(path, "w");
      if (!h) continue;
      fprintf(h, "t,T\n");
      if (g->hp_t)
        for (int e = 0; e < g->hp_t[p].n; e++) fprintf(h, "%.6g,%.6g\n", g->hp_t[p].v[e], g->hp_T[p].v[e]);
      else
        for (int e = 0; e < g->hist_n; e++) fprintf(h, "%.6g,%.6g\n", g->hist_t[e], g->hist_T[(size_t)e * np + p]);
      fclose(h);
    }
  }
  if (g->full_dir) {
    snprintf(path, sizeof path, "%s/%s.T_hist.full.txt", g->full_dir, sim->name);
    FILE* f = fopen(path, "w");
    if (f) {
      for (int p = 0; p < np; p++) {
        if (g->hp_t) {
          for (int e = 0; e < g->hp_t[p].n; e++)
            fprintf(f, "%d %.17g %.17g\n", p, g->hp_t[p].v[e], g->hp_T[p].v[e]);
          continue;
        }
        for (int e = 0; e < g->hist_n; e++)
          fprintf(f, "%d %.17g %.17g\n", p, g->hist_t[e], g->hist_T[(size_t)e * np + p]);
      }
      fclose(f);
    }
  }
}

/* Grid::Output_RDF (Grid.cpp:176-201): one row per recorded melt time of
 * each output-flagged point. A melt still open at the end has no liquidus
 * time in the reference (it reads past its list); 0 is written there. */
static void rdf_output(or_grid* g) {
  const or_sim* sim = g->sim;
  char path[2048];
  snprintf(path, sizeof path, "%s/%s.RDF.Final.csv", sim->data_dir, sim->name);
  FILE* f = fopen(path, "w");
  FILE* ff = NULL;
  if (g->full_dir) {
    snprintf(path, sizeof path, "%s/%s.RDF.full.txt", g->full_dir, sim->name);
    ff = fopen(path, "w");
  }
  if (f) fprintf(f, "x,y,z,tm,tl,cr\n");
  for (int p = 0; p < sim->pnum; p++) {
    if (g->out_flag && !g->out_flag[p]) continue;
    double x, y, z;
    or_point_xyz(sim, p, &x, &y, &z);
    for (int e = 0; e < g->rdf_tm[p].n; e++) {
      const double tm = g->rdf_tm[p].v[e];
      const double tl = e < g->rdf_tl[p].n ? g->rdf_tl[p].v[e] : 0.0;
      const double cr = e < g->rdf_cr[p].n ? g->rdf_cr[p].v[e] : 0.0;
      if (f) fprintf(f, "%.6g,%.6g,%.6g,%.6g,%.6g,%.6g\n", x, y, z, tm, tl, cr);
      if (ff) fprintf(ff, "%d %.17g %.17g %.17g\n", p, tm, tl, cr);
    }
  }
  if (f) fclose(f);
  if (ff) fclose(ff);
}

/* Grid::Output_RRDF_csv (Grid.cpp:203-239) */
static void rrdf_output(or_grid* g) {
  const or_sim* sim = g->sim;
  char path[2048];
  snprintf(path, sizeof path, "%s/%s.RRDF.csv", sim->data_dir, sim->name);
  FILE* f = fopen(path, "w");
  FILE* ff = NULL;
  if (g->full_dir) {
    snprintf(path, sizeof path, "%s/%s.RRDF.full.txt", g->full_dir, sim->name);
    ff = fopen(path, "w");
  }
  if (f) {
    fprintf(f, "x,y,z,t_prev,t_cur");
    fprintf(f, ",T000_prev,T001_prev,T010_prev,T011_prev,T100_prev,T101_prev,T110_prev,T111_prev");
    fprintf(f, ",T000_cur,T001_cur,T010_cur,T011_cur,T100_cur,T101_cur,T110_cur,T111_cur\n");
  }
  for (size_t r = 0; r < g->rr_n; r++) {
    if (f) {
      fprintf(f, "%u,%u,%u,%.6g,%.6g", g->rr_idx[3 * r], g->rr_idx[3 * r + 1], g->rr_idx[3 * r + 2],
              g->rr_ts[2 * r], g->rr_ts[2 * r + 1]);
      for (int n = 0; n < 16; n++) fprintf(f, ",%.6g", g->rr_Ts[16 * r + n]);
      fprintf(f, "\n");
    }
    if (ff) {
      fprintf(ff, "%u %u %u %.17g %.17g", g->rr_idx[3 * r], g->rr_idx[3 * r + 1], g->rr_idx[3 * r + 2],
              g->rr_ts[2 * r], g->rr_ts[2 * r + 1]);
      for (int n = 0; n < 16; n++) fprintf(ff, " %.17g", g->rr_Ts[16 * r + n]);
      fprintf(ff, "\n");
    }
  }
  if (f) fclose(f);
  if (ff) fclose(ff);
}

int or_run(or_sim* sim, const char* full_dir, double* evals) {
  if (sim->parse_error) return 1;
  or_grid g;
  grid_init(&g, sim);
  g.full_dir = full_dir;
  /* Run.cpp:61, :391: custom point files always take the untracked drivers */
  const int untracked = sim->tracking == OR_TRACK_NONE || sim->custom_points;
  if (sim->mode == OR_MODE_SNAPSHOTS) {
    if (untracked) snapshots_none(&g);
    else snapshots_volume(&g);
  } else {
    if (untracked) solidify_none(&g);
    else if (sim->tracking == OR_TRACK_VOLUME) solidify_volume(&g);
    else if (sim->tracking == OR_TRACK_SURFACE) solidify_surface(&g);
    else stork(&g);
    /* Main.cpp:104-112: Stork writes its records instead of the grid table */
    if (sim->tracking != OR_TRACK_STORK) grid_output(&g, "Solidification.Final");
    else rrdf_output(&g);
  }
  if (sim->o_Thist) thist_output(&g);  /* Main.cpp:113-115 */
  if (sim->o_RDF && g.rdf_tm) rdf_output(&g);  /* Main.cpp:116-118 */
  if (evals) *evals = g.evals;
  grid_free(&g);
  return 0;
}
