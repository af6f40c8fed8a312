/*
  jur_oracle.c -- CPU FP64 restatement of JURASSIC's forward-model path.

  TEST INFRASTRUCTURE ONLY (see jur_oracle.h): the checker for the MI355X
  library, used by tests/, the smoke check and the CPU-baseline leg of the
  benchmark.  Compiled with gcc -O2 -ffp-contract=off so that no multiply-add
  is fused, as in the reference's x86-64 build.

  Each function restates the reference function named in its comment
  (file:line in slcs-jsc/jurassic) with the same formulas and operation
  order.  Data structures are this file's own.
*/
#define _GNU_SOURCE
#include "jur_oracle.h"

#include <math.h>
#include <omp.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "../jurassic_amd/csrc/ctm_coeffs.h"

/* Physical constants (src/jurassic.h:252-340). */
#define C1 1.19104259e-8
#define C2 1.43877506
#define KB 1.3806504e-23
#define NA 6.02214199e23
#define RE 6367.421
#define RI 8.3144598
#define G0 9.80665
#define P0 1013.25
#define TMIN 100.
#define TMAX 400.
#define TSUN 5780.
#define UMIN 0
#define UMAX 1e30
#define EPSMIN 0
#define EPSMAX 1

/* Interpolation macros (src/jurassic.h:164-177), same expression trees. */
#define LIN(x0, y0, x1, y1, x) ((y0) + ((y1) - (y0)) / ((x1) - (x0)) * ((x) - (x0)))
#define LOGX(x0, y0, x1, y1, x)                                              \
  (((x) / (x0) > 0 && (x1) / (x0) > 0)                                       \
       ? ((y0) + ((y1) - (y0)) * log((x) / (x0)) / log((x1) / (x0)))         \
       : LIN(x0, y0, x1, y1, x))
#define LOGY(x0, y0, x1, y1, x)                                              \
  (((y1) / (y0) > 0)                                                         \
       ? ((y0) * exp(log((y1) / (y0)) / ((x1) - (x0)) * ((x) - (x0))))       \
       : LIN(x0, y0, x1, y1, x))
#define POW2(x) ((x) * (x))
#define POW3(x) ((x) * (x) * (x))
#define DOTP(a, b) (a[0] * b[0] + a[1] * b[1] + a[2] * b[2])
#define NORM(a) sqrt(DOTP(a, a))
#define DIST2(a, b)                                                          \
  ((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +           \
   (a[2] - b[2]) * (a[2] - b[2]))
#define DIST(a, b) sqrt(DIST2(a, b))
#define DEG2RAD(d) ((d) * (M_PI / 180.0))
#define RAD2DEG(r) ((r) * (180.0 / M_PI))
#define REFRAC(p, T) (7.753e-05 * (p) / (T))
#define MAX(a, b) (((a) > (b)) ? (a) : (b))
#define MIN(a, b) (((a) < (b)) ? (a) : (b))

/* ------------------------------------------------------------------ */
/* Error text.                                                         */
/* ------------------------------------------------------------------ */

static __thread char orc_err[512];

static int fail(int code, const char *msg) {
  snprintf(orc_err, sizeof orc_err, "%s", msg);
  return code;
}

const char *orc_last_error(void) { return orc_err; }

/* ------------------------------------------------------------------ */
/* Tables: one block per (channel, emitter), sized to the file read.   */
/* ------------------------------------------------------------------ */

typedef struct {
  int np;                     /* -1: missing table (src/jurassic.c:4980-4983) */
  double p[TBLNP];
  int nt[TBLNP];
  double t[TBLNP][TBLNT];
  int nu[TBLNP][TBLNT];
  float *u[TBLNP][TBLNT];     /* column arrays of nu[ip][it] values */
  float *eps[TBLNP][TBLNT];
} orc_tab_t;

typedef struct {
  /* key */
  int nd, ng, tblfmt;
  double nu[ND];
  char emitter[NG][64];
  char tblbase[LEN];
  /* data */
  orc_tab_t *tab;             /* [nd][ng] */
  double st[TBLNS];
  double *sr;                 /* [TBLNS][nd] */
} orc_ctx_t;

static orc_ctx_t *g_ctx = NULL;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static void ctx_free(orc_ctx_t *c) {
  if (!c) return;
  if (c->tab) {
    for (int i = 0; i < c->nd * c->ng; i++)
      for (int ip = 0; ip < TBLNP; ip++)
        for (int it = 0; it < TBLNT; it++) {
          free(c->tab[i].u[ip][it]);
          free(c->tab[i].eps[ip][it]);
        }
    free(c->tab);
  }
  free(c->sr);
  free(c);
}

void orc_reset(void) {
  pthread_mutex_lock(&g_lock);
  ctx_free(g_ctx);
  g_ctx = NULL;
  pthread_mutex_unlock(&g_lock);
}

static int ctx_matches(const orc_ctx_t *c, const ctl_t *ctl) {
  if (c->nd != ctl->nd || c->ng != ctl->ng || c->tblfmt != ctl->tblfmt)
    return 0;
  if (strcmp(c->tblbase, ctl->tblbase) != 0) return 0;
  for (int id = 0; id < ctl->nd; id++)
    if (c->nu[id] != ctl->nu[id]) return 0;
  for (int ig = 0; ig < ctl->ng; ig++)
    if (strncmp(c->emitter[ig], ctl->emitter[ig], 63) != 0) return 0;
  return 1;
}

/* read_tbl (src/jurassic.c:4949-5110): ASCII rows "p T u eps" with the
   range filter and the monotonicity merge, or the binary layout. */
static int read_one_table(const ctl_t *ctl, int id, int ig, orc_tab_t *tb) {
  char filename[2 * LEN + 64];
  tb->np = -1;
  snprintf(filename, sizeof filename, "%s_%.4f_%s.%s", ctl->tblbase,
           ctl->nu[id], ctl->emitter[ig], ctl->tblfmt == 1 ? "tab" : "bin");
  FILE *in = fopen(filename, "r");
  if (!in) return 0; /* missing table: np = -1 -> eps = 0 downstream */
  if (ctl->tblfmt == 1) {
    /* Rows are accumulated into per-column scratch of TBLNU values. */
    float(*ubuf)[TBLNT][TBLNU] = malloc(sizeof(float[TBLNP][TBLNT][TBLNU]));
    float(*ebuf)[TBLNT][TBLNU] = malloc(sizeof(float[TBLNP][TBLNT][TBLNU]));
#define ASCII_FAIL(code, msg) { free(ubuf); free(ebuf); fclose(in); return fail(code, msg); }
    char line[LEN];
    double eps, press, temp, u;
    double eps_old = -999, press_old = -999, temp_old = -999, u_old = -999;
    while (fgets(line, LEN, in)) {
      if (sscanf(line, "%lg %lg %lg %lg", &press, &temp, &u, &eps) != 4)
        continue;
      if (u < UMIN || u > UMAX || eps < EPSMIN || eps > EPSMAX) continue;
      if (press != press_old) {
        press_old = press;
        if ((++tb->np) >= TBLNP) ASCII_FAIL(JUR_ERR_ARG, "Too many pressure levels!");
        tb->nt[tb->np] = -1;
      }
      if (temp != temp_old) {
        temp_old = temp;
        if ((++tb->nt[tb->np]) >= TBLNT) ASCII_FAIL(JUR_ERR_ARG, "Too many temperatures!");
        tb->nu[tb->np][tb->nt[tb->np]] = -1;
      }
      int ip = tb->np, it = tb->nt[ip];
      if ((eps > eps_old && u > u_old) || tb->nu[ip][it] < 0) {
        eps_old = eps;
        u_old = u;
        if ((++tb->nu[ip][it]) >= TBLNU) ASCII_FAIL(JUR_ERR_ARG, "Too many column densities!");
      }
      tb->p[ip] = press;
      tb->t[ip][it] = temp;
      ubuf[ip][it][tb->nu[ip][it]] = (float)u;
      ebuf[ip][it][tb->nu[ip][it]] = (float)eps;
    }
    tb->np++;
    for (int ip = 0; ip < tb->np; ip++) {
      tb->nt[ip]++;
      for (int it = 0; it < tb->nt[ip]; it++) {
        tb->nu[ip][it]++;
        int n = tb->nu[ip][it];
        tb->u[ip][it] = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        tb->eps[ip][it] = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        memcpy(tb->u[ip][it], ubuf[ip][it], (size_t)n * sizeof(float));
        memcpy(tb->eps[ip][it], ebuf[ip][it], (size_t)n * sizeof(float));
      }
    }
    free(ubuf);
    free(ebuf);
#undef ASCII_FAIL
  } else if (ctl->tblfmt == 2) {
#define RD(ptr, sz, cnt)                                                      \
  if (fread(ptr, sz, (size_t)(cnt), in) != (size_t)(cnt)) {                    \
    fclose(in);                                                               \
    return fail(JUR_ERR_IO, "Error while reading!");                          \
  }
    RD(&tb->np, sizeof(int), 1);
    if (tb->np > TBLNP) { fclose(in); return fail(JUR_ERR_ARG, "Too many pressure levels!"); }
    RD(tb->p, sizeof(double), tb->np);
    for (int ip = 0; ip < tb->np; ip++) {
      RD(&tb->nt[ip], sizeof(int), 1);
      if (tb->nt[ip] > TBLNT) { fclose(in); return fail(JUR_ERR_ARG, "Too many temperatures!"); }
      RD(tb->t[ip], sizeof(double), tb->nt[ip]);
      for (int it = 0; it < tb->nt[ip]; it++) {
        RD(&tb->nu[ip][it], sizeof(int), 1);
        int n = tb->nu[ip][it];
        if (n > TBLNU) { fclose(in); return fail(JUR_ERR_ARG, "Too many column densities!"); }
        tb->u[ip][it] = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        tb->eps[ip][it] = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        RD(tb->u[ip][it], sizeof(float), n);
        RD(tb->eps[ip][it], sizeof(float), n);
      }
    }
#undef RD
  } else {
    fclose(in);
    return fail(JUR_ERR_ARG, "Unknown look-up table format!");
  }
  fclose(in);
  return 0;
}

/* read_shape (src/jurassic.c:4907-4946). */
static int read_shape(const char *filename, double *x, double *y, int *n) {
  FILE *in = fopen(filename, "r");
  if (!in) return fail(JUR_ERR_IO, "Cannot open file!");
  char line[LEN];
  *n = 0;
  while (fgets(line, LEN, in))
    if (sscanf(line, "%lg %lg", &x[*n], &y[*n]) == 2)
      if ((++(*n)) > NSHAPE) { fclose(in); return fail(JUR_ERR_ARG, "Too many data points!"); }
  fclose(in);
  if (*n < 2) return fail(JUR_ERR_IO, "Could not read any data!");
  return 0;
}

/* init_srcfunc (src/jurassic.c:3631-3681): filter-weighted Planck table on
   a uniform 1200-point temperature grid, integrated with the fnu += dnu
   loop of the reference. */
static int init_srcfunc(const ctl_t *ctl, orc_ctx_t *c) {
  double *f = malloc(NSHAPE * sizeof(double));
  double *nu = malloc(NSHAPE * sizeof(double));
  int n, rc = 0;
  for (int id = 0; id < ctl->nd && rc == 0; id++) {
    char filename[2 * LEN + 32];
    snprintf(filename, sizeof filename, "%s_%.4f.filt", ctl->tblbase,
             ctl->nu[id]);
    rc = read_shape(filename, nu, f, &n);
    if (rc) break;
    double dnu = 1.0;
    for (int i = 1; i < n; i++) dnu = MIN(dnu, nu[i] - nu[i - 1]);
#pragma omp parallel for schedule(static)
    for (int it = 0; it < TBLNS; it++) {
      c->st[it] = LIN(0.0, TMIN, TBLNS - 1.0, TMAX, (double)it);
      double fsum = 0, acc = 0;
      for (double fnu = nu[0]; fnu <= nu[n - 1]; fnu += dnu) {
        const int i = orc_locate_irr(nu, n, fnu);
        const double ff = LIN(nu[i], f[i], nu[i + 1], f[i + 1], fnu);
        fsum += ff;
        acc += ff * (C1 * POW3(fnu) / expm1(C2 * fnu / c->st[it]));
      }
      c->sr[it * ctl->nd + id] = acc / fsum;
    }
  }
  free(f);
  free(nu);
  return rc;
}

/* Lazy per-ctl context: tables and source function (the reference keeps
   them in statics of formod_pencil bound to the first ctl,
   src/jurassic.c:3202-3217; here they are re-read when ctl changes). */
static int get_ctx(const ctl_t *ctl, orc_ctx_t **out) {
  pthread_mutex_lock(&g_lock);
  if (g_ctx && ctx_matches(g_ctx, ctl)) {
    *out = g_ctx;
    pthread_mutex_unlock(&g_lock);
    return 0;
  }
  ctx_free(g_ctx);
  g_ctx = NULL;
  orc_ctx_t *c = calloc(1, sizeof(orc_ctx_t));
  c->nd = ctl->nd;
  c->ng = ctl->ng;
  c->tblfmt = ctl->tblfmt;
  snprintf(c->tblbase, LEN, "%s", ctl->tblbase);
  for (int id = 0; id < ctl->nd; id++) c->nu[id] = ctl->nu[id];
  for (int ig = 0; ig < ctl->ng; ig++)
    snprintf(c->emitter[ig], 64, "%s", ctl->emitter[ig]);
  c->tab = calloc((size_t)(ctl->nd * ctl->ng + 1), sizeof(orc_tab_t));
  c->sr = calloc((size_t)TBLNS * (size_t)(ctl->nd + 1), sizeof(double));
  int rc = 0;
  for (int id = 0; id < ctl->nd && rc == 0; id++)
    for (int ig = 0; ig < ctl->ng && rc == 0; ig++)
      rc = read_one_table(ctl, id, ig, &c->tab[id * ctl->ng + ig]);
  if (rc == 0) rc = init_srcfunc(ctl, c);
  if (rc) {
    ctx_free(c);
    pthread_mutex_unlock(&g_lock);
    return rc;
  }
  g_ctx = c;
  *out = c;
  pthread_mutex_unlock(&g_lock);
  return 0;
}

int orc_srcfunc_table(const ctl_t *ctl, double *st, double *sr) {
  orc_ctx_t *c;
  int rc = get_ctx(ctl, &c);
  if (rc) return rc;
  memcpy(st, c->st, sizeof c->st);
  memcpy(sr, c->sr, (size_t)TBLNS * (size_t)ctl->nd * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------------ */
/* LOS storage.                                                        */
/* ------------------------------------------------------------------ */

orc_los_t *orc_los_alloc(int nd, int ng) {
  orc_los_t *l = calloc(1, sizeof(orc_los_t));
  int nd1 = nd > 0 ? nd : 1, ng1 = ng > 0 ? ng : 1;
  l->nd = nd;
  l->ng = ng;
  l->z = malloc(NLOS * sizeof(double));
  l->lon = malloc(NLOS * sizeof(double));
  l->lat = malloc(NLOS * sizeof(double));
  l->p = malloc(NLOS * sizeof(double));
  l->t = malloc(NLOS * sizeof(double));
  l->ds = malloc(NLOS * sizeof(double));
  l->sfeps = malloc((size_t)nd1 * sizeof(double));
  l->q = malloc((size_t)NLOS * ng1 * sizeof(double));
  l->k = malloc((size_t)NLOS * nd1 * sizeof(double));
  l->u = malloc((size_t)NLOS * ng1 * sizeof(double));
  l->cgp = malloc((size_t)NLOS * ng1 * sizeof(double));
  l->cgt = malloc((size_t)NLOS * ng1 * sizeof(double));
  l->cgu = malloc((size_t)NLOS * ng1 * sizeof(double));
  l->eps = calloc((size_t)NLOS * nd1, sizeof(double));
  l->src = calloc((size_t)NLOS * nd1, sizeof(double));
  return l;
}

void orc_los_free(orc_los_t *l) {
  if (!l) return;
  free(l->z); free(l->lon); free(l->lat); free(l->p); free(l->t); free(l->ds);
  free(l->sfeps); free(l->q); free(l->k); free(l->u); free(l->cgp);
  free(l->cgt); free(l->cgu); free(l->eps); free(l->src);
  free(l);
}

/* ------------------------------------------------------------------ */
/* Geometry, search and interpolation.                                 */
/* ------------------------------------------------------------------ */

/* cart2geo (src/jurassic.c:108-119). */
void orc_cart2geo(const double *x, double *z, double *lon, double *lat) {
  const double radius = NORM(x);
  *lat = RAD2DEG(asin(x[2] / radius));
  *lon = RAD2DEG(atan2(x[1], x[0]));
  *z = radius - RE;
}

/* geo2cart (src/jurassic.c:3500-3516). */
void orc_geo2cart(double z, double lon, double lat, double *x) {
  const double radius = z + RE;
  const double latrad = lat / 180. * M_PI;
  const double lonrad = lon / 180. * M_PI;
  const double coslat = cos(latrad);
  x[0] = radius * coslat * cos(lonrad);
  x[1] = radius * coslat * sin(lonrad);
  x[2] = radius * sin(latrad);
}

/* locate_irr (src/jurassic.c:4103-4129): bisection on an ascending or
   descending grid. */
int orc_locate_irr(const double *xx, int n, double x) {
  int ilo = 0, ihi = n - 1, i = (ihi + ilo) >> 1;
  if (xx[i] < xx[i + 1])
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if (xx[i] > x) ihi = i; else ilo = i;
    }
  else
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if (xx[i] <= x) ihi = i; else ilo = i;
    }
  return ilo;
}

/* locate_reg (src/jurassic.c:4133-4148): uniform grid, clamped. */
int orc_locate_reg(const double *xx, int n, double x) {
  const int i = (int)((x - xx[0]) / (xx[1] - xx[0]));
  if (i < 0) return 0;
  else if (i > n - 2) return n - 2;
  else return i;
}

/* locate_tbl (src/jurassic.c:4152-4170): bisection on an ascending float array. */
int orc_locate_tbl(const float *xx, int n, double x) {
  int ilo = 0, ihi = n - 1, i = (ihi + ilo) >> 1;
  while (ihi > ilo + 1) {
    i = (ihi + ilo) >> 1;
    if (xx[i] > x) ihi = i; else ilo = i;
  }
  return ilo;
}

/* intpol_atm (src/jurassic.c:3685-3706). */
void orc_intpol_atm(const ctl_t *ctl, const atm_t *atm, double z, double *p,
                    double *t, double *q, double *k) {
  const int ip = orc_locate_irr(atm->z, atm->np, z);
  *p = LOGY(atm->z[ip], atm->p[ip], atm->z[ip + 1], atm->p[ip + 1], z);
  *t = LIN(atm->z[ip], atm->t[ip], atm->z[ip + 1], atm->t[ip + 1], z);
  for (int ig = 0; ig < ctl->ng; ig++)
    q[ig] = LIN(atm->z[ip], atm->q[ig][ip], atm->z[ip + 1], atm->q[ig][ip + 1], z);
  for (int iw = 0; iw < ctl->nw; iw++)
    k[iw] = LIN(atm->z[ip], atm->k[iw][ip], atm->z[ip + 1], atm->k[iw][ip + 1], z);
}

/* gsl_stats_minmax semantics: NaN propagates. */
static void stats_minmax(double *mn, double *mx, const double *d, int n) {
  double a = d[0], b = d[0];
  for (int i = 0; i < n; i++) {
    if (d[i] < a) a = d[i];
    if (d[i] > b) b = d[i];
    if (isnan(d[i])) { a = b = d[i]; break; }
  }
  *mn = a;
  *mx = b;
}

/* gsl_stats_min_index semantics: first minimum, or first NaN. */
static int stats_min_index(const double *d, int n) {
  double m = d[0];
  int im = 0;
  for (int i = 0; i < n; i++) {
    if (d[i] < m) { m = d[i]; im = i; }
    if (isnan(d[i])) return i;
  }
  return im;
}

/* tangent_point (src/jurassic.c:5223-5263): end point for nadir/zenith,
   parabola through the three points around the minimum for limb. */
void orc_tangent_point(const orc_los_t *los, double *tpz, double *tplon,
                       double *tplat) {
  double dummy, v[3], v0[3], v2[3];
  const int ip = stats_min_index(los->z, los->np);
  if (ip <= 0 || ip >= los->np - 1) {
    *tpz = los->z[los->np - 1];
    *tplon = los->lon[los->np - 1];
    *tplat = los->lat[los->np - 1];
  } else {
    const double yy0 = los->z[ip - 1];
    const double yy1 = los->z[ip];
    const double yy2 = los->z[ip + 1];
    const double x1 = sqrt(POW2(los->ds[ip]) - POW2(yy1 - yy0));
    const double x2 = x1 + sqrt(POW2(los->ds[ip + 1]) - POW2(yy2 - yy1));
    const double a = 1 / (x1 - x2) * (-(yy0 - yy1) / x1 + (yy0 - yy2) / x2);
    const double b = -(yy0 - yy1) / x1 - a * x1;
    const double c = yy0;
    const double x = -b / (2 * a);
    *tpz = a * x * x + b * x + c;
    orc_geo2cart(los->z[ip - 1], los->lon[ip - 1], los->lat[ip - 1], v0);
    orc_geo2cart(los->z[ip + 1], los->lon[ip + 1], los->lat[ip + 1], v2);
    for (int i = 0; i < 3; i++) v[i] = LIN(0.0, v0[i], x2, v2[i], x);
    orc_cart2geo(v, &dummy, tplon, tplat);
  }
}

/* raytrace (src/jurassic.c:4201-4438). */
int orc_raytrace(const ctl_t *ctl, const atm_t *atm, obs_t *obs,
                 orc_los_t *los, int ir) {
  const double h = 0.02, zrefrac = 60;
  double ex0[3], ex1[3], k[NW], lat, lon, n, ng[3], norm, p, q[NG], t, x[3],
      xh[3], xobs[3], xvp[3], z = 1e99, zmax, zmin;
  int stop = 0;
  const int nd = ctl->nd, ngas = ctl->ng;

  los->np = 0;
  los->sft = -999;
  obs->tpz[ir] = obs->vpz[ir];
  obs->tplon[ir] = obs->vplon[ir];
  obs->tplat[ir] = obs->vplat[ir];

  stats_minmax(&zmin, &zmax, atm->z, atm->np);
  if (ctl->nsf > 0) {
    zmin = MAX(atm->sfz, zmin);
    if (atm->sfp > 0) {
      const int ip = orc_locate_irr(atm->p, atm->np, atm->sfp);
      const double zip = LIN(log(atm->p[ip]), atm->z[ip], log(atm->p[ip + 1]),
                             atm->z[ip + 1], log(atm->sfp));
      zmin = MAX(zip, zmin);
    }
  }

  if (obs->obsz[ir] < zmin) return fail(JUR_ERR_OBSERVER, "Observer below surface!");
  if (obs->vpz[ir] > zmax) return 0;

  orc_geo2cart(obs->obsz[ir], obs->obslon[ir], obs->obslat[ir], xobs);
  orc_geo2cart(obs->vpz[ir], obs->vplon[ir], obs->vplat[ir], xvp);
  for (int i = 0; i < 3; i++) ex0[i] = xvp[i] - xobs[i];
  norm = NORM(ex0);
  for (int i = 0; i < 3; i++) ex0[i] /= norm;
  for (int i = 0; i < 3; i++) x[i] = xobs[i];

  /* Entry point search for an observer above the atmosphere. */
  if (obs->obsz[ir] > zmax) {
    double dmax = norm, dmin = 0;
    while (fabs(dmin - dmax) > 0.001) {
      const double d = (dmax + dmin) / 2;
      for (int i = 0; i < 3; i++) x[i] = xobs[i] + d * ex0[i];
      orc_cart2geo(x, &z, &lon, &lat);
      if (z <= zmax && z > zmax - 0.001) break;
      if (z < zmax - 0.0005) dmax = d; else dmin = d;
    }
  }

  while (1) {
    double ds = ctl->rayds;
    if (ctl->raydz > 0) {
      norm = NORM(x);
      for (int i = 0; i < 3; i++) xh[i] = x[i] / norm;
      const double cosa = fabs(DOTP(ex0, xh));
      if (cosa != 0) ds = MIN(ctl->rayds, ctl->raydz / cosa);
    }

    orc_cart2geo(x, &z, &lon, &lat);

    if (z < zmin || z > zmax) {
      stop = (z < zmin ? 2 : 1);
      const double frac = ((z < zmin ? zmin : zmax) - los->z[los->np - 1]) /
                          (z - los->z[los->np - 1]);
      orc_geo2cart(los->z[los->np - 1], los->lon[los->np - 1],
                   los->lat[los->np - 1], xh);
      for (int i = 0; i < 3; i++) x[i] = xh[i] + frac * (x[i] - xh[i]);
      orc_cart2geo(x, &z, &lon, &lat);
      los->ds[los->np - 1] = ds * frac;
      ds = 0;
    }

    orc_intpol_atm(ctl, atm, z, &p, &t, q, k);

    los->lon[los->np] = lon;
    los->lat[los->np] = lat;
    los->z[los->np] = z;
    los->p[los->np] = p;
    los->t[los->np] = t;
    for (int ig = 0; ig < ngas; ig++) los->q[los->np * ngas + ig] = q[ig];
    for (int id = 0; id < nd; id++) los->k[los->np * nd + id] = k[ctl->window[id]];
    los->ds[los->np] = ds;

    if (ctl->ncl > 0 && atm->cldz > 0) {
      const double aux = exp(-0.5 * POW2((z - atm->clz) / atm->cldz));
      for (int id = 0; id < nd; id++) {
        const int icl = orc_locate_irr(ctl->clnu, ctl->ncl, ctl->nu[id]);
        los->k[los->np * nd + id] +=
            aux * LIN(ctl->clnu[icl], atm->clk[icl], ctl->clnu[icl + 1],
                      atm->clk[icl + 1], ctl->nu[id]);
      }
    }

    if ((++los->np) > NLOS) return fail(JUR_ERR_LOS, "Too many LOS points!");

    if (stop) {
      if (ctl->nsf > 0 && atm->sft > 0) t = atm->sft;
      los->sft = (stop == 2 ? t : -999);
      for (int id = 0; id < nd; id++) {
        los->sfeps[id] = 1.0;
        if (ctl->nsf > 0) {
          const int isf = orc_locate_irr(ctl->sfnu, ctl->nsf, ctl->nu[id]);
          los->sfeps[id] = LIN(ctl->sfnu[isf], atm->sfeps[isf],
                               ctl->sfnu[isf + 1], atm->sfeps[isf + 1],
                               ctl->nu[id]);
        }
      }
      break;
    }

    if (ctl->refrac && z <= zrefrac) n = 1 + REFRAC(p, t);
    else n = 1;

    for (int i = 0; i < 3; i++) ex1[i] = ex0[i] * n;

    if (ctl->refrac && z <= zrefrac) {
      for (int i = 0; i < 3; i++) xh[i] = x[i] + 0.5 * ds * ex0[i];
      orc_cart2geo(xh, &z, &lon, &lat);
      orc_intpol_atm(ctl, atm, z, &p, &t, q, k);
      n = REFRAC(p, t);
      for (int i = 0; i < 3; i++) {
        xh[i] += h;
        orc_cart2geo(xh, &z, &lon, &lat);
        orc_intpol_atm(ctl, atm, z, &p, &t, q, k);
        ng[i] = (REFRAC(p, t) - n) / h;
        xh[i] -= h;
      }
    } else
      for (int i = 0; i < 3; i++) ng[i] = 0;

    for (int i = 0; i < 3; i++) ex1[i] += ds * ng[i];
    norm = NORM(ex1);
    for (int i = 0; i < 3; i++) ex1[i] /= norm;
    for (int i = 0; i < 3; i++) x[i] += 0.5 * ds * (ex0[i] + ex1[i]);
    for (int i = 0; i < 3; i++) ex0[i] = ex1[i];
  }

  orc_tangent_point(los, &obs->tpz[ir], &obs->tplon[ir], &obs->tplat[ir]);

  /* Trapezoid segment lengths, column densities, Curtis-Godson means
     (src/jurassic.c:4410-4437). */
  for (int ip = los->np - 1; ip >= 1; ip--)
    los->ds[ip] = 0.5 * (los->ds[ip - 1] + los->ds[ip]);
  los->ds[0] *= 0.5;
  for (int ip = 0; ip < los->np; ip++)
    for (int ig = 0; ig < ngas; ig++)
      los->u[ip * ngas + ig] = 10 * los->q[ip * ngas + ig] * los->p[ip] /
                               (KB * los->t[ip]) * los->ds[ip];
  for (int ig = 0; ig < ngas; ig++) {
    los->cgu[ig] = los->u[ig];
    los->cgp[ig] = los->u[ig] * los->p[0];
    los->cgt[ig] = los->u[ig] * los->t[0];
  }
  for (int ip = 1; ip < los->np; ip++)
    for (int ig = 0; ig < ngas; ig++) {
      const int a = ip * ngas + ig, b = (ip - 1) * ngas + ig;
      los->cgu[a] = los->cgu[b] + los->u[a];
      los->cgp[a] = los->cgp[b] + los->u[a] * los->p[ip];
      los->cgt[a] = los->cgt[b] + los->u[a] * los->t[ip];
    }
  for (int ip = 0; ip < los->np; ip++)
    for (int ig = 0; ig < ngas; ig++) {
      const int a = ip * ngas + ig;
      los->cgp[a] /= los->cgu[a];
      los->cgt[a] /= los->cgu[a];
    }
  return 0;
}

/* hydrostatic (src/jurassic.c:3520-3582). */
void orc_hydrostatic(const ctl_t *ctl, atm_t *atm) {
  const double mmair = 28.96456e-3, mmh2o = 18.0153e-3;
  const int ipts = 20;
  double dzmin = 1e99, e = 0;
  int ipref = 0, ig_h2o = -1;
  if (ctl->hydz < 0) return;
  for (int ig = 0; ig < ctl->ng; ig++)
    if (strcasecmp(ctl->emitter[ig], "H2O") == 0) { ig_h2o = ig; break; }
  for (int ip = 0; ip < atm->np; ip++)
    if (fabs(atm->z[ip] - ctl->hydz) < dzmin) {
      dzmin = fabs(atm->z[ip] - ctl->hydz);
      ipref = ip;
    }
  for (int ip = ipref + 1; ip < atm->np; ip++) {
    double mean = 0;
    for (int i = 0; i < ipts; i++) {
      if (ig_h2o >= 0)
        e = LIN(0.0, atm->q[ig_h2o][ip - 1], ipts - 1.0, atm->q[ig_h2o][ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * G0 / RI /
              LIN(0.0, atm->t[ip - 1], ipts - 1.0, atm->t[ip], (double)i) / ipts;
    }
    atm->p[ip] = exp(log(atm->p[ip - 1]) - mean * 1000 * (atm->z[ip] - atm->z[ip - 1]));
  }
  for (int ip = ipref - 1; ip >= 0; ip--) {
    double mean = 0;
    for (int i = 0; i < ipts; i++) {
      if (ig_h2o >= 0)
        e = LIN(0.0, atm->q[ig_h2o][ip + 1], ipts - 1.0, atm->q[ig_h2o][ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * G0 / RI /
              LIN(0.0, atm->t[ip + 1], ipts - 1.0, atm->t[ip], (double)i) / ipts;
    }
    atm->p[ip] = exp(log(atm->p[ip + 1]) - mean * 1000 * (atm->z[ip] - atm->z[ip + 1]));
  }
}

/* ------------------------------------------------------------------ */
/* Continua (src/jurassic.c:886-2918).                                 */
/* ------------------------------------------------------------------ */

/* ctmco2 (src/jurassic.c:886-1744; formula :1724-1741). */
double orc_ctmco2(double nu, double p, double t, double u) {
  const double xw = nu / 2 + 1;
  if (xw >= 1 && xw < 2001) {
    const int iw = (int)xw;
    const double dw = xw - iw;
    const double ew = 1 - dw;
    const double cw296 = ew * JUR_CTM_CO2_296[iw - 1] + dw * JUR_CTM_CO2_296[iw];
    const double cw260 = ew * JUR_CTM_CO2_260[iw - 1] + dw * JUR_CTM_CO2_260[iw];
    const double cw230 = ew * JUR_CTM_CO2_230[iw - 1] + dw * JUR_CTM_CO2_230[iw];
    const double dt230 = t - 230;
    const double dt260 = t - 260;
    const double dt296 = t - 296;
    const double ctw = dt260 * 5.050505e-4 * dt296 * cw230 -
                       dt230 * 9.259259e-4 * dt296 * cw260 +
                       dt230 * 4.208754e-4 * dt260 * cw296;
    return u / NA / 1000 * p / P0 * ctw;
  }
  return 0;
}

/* ctmh2o (src/jurassic.c:1746-2794; formula :2762-2791). */
double orc_ctmh2o(double nu, double p, double t, double q, double u) {
  double sfac;
  const double xw = nu / 10 + 1;
  if (xw >= 1 && xw < 2001) {
    const int iw = (int)xw;
    const double dw = xw - iw;
    const double ew = 1 - dw;
    const double cw296 = ew * JUR_CTM_H2O_296[iw - 1] + dw * JUR_CTM_H2O_296[iw];
    const double cw260 = ew * JUR_CTM_H2O_260[iw - 1] + dw * JUR_CTM_H2O_260[iw];
    const double cwfrn = ew * JUR_CTM_H2O_FRN[iw - 1] + dw * JUR_CTM_H2O_FRN[iw];
    if (nu <= 820 || nu >= 960) {
      sfac = 1;
    } else {
      const double xx = (nu - 820) / 10;
      const int ix = (int)xx;
      const double dx = xx - ix;
      sfac = (1 - dx) * JUR_CTM_H2O_XFCREV[ix] + dx * JUR_CTM_H2O_XFCREV[ix + 1];
    }
    const double ctwslf = sfac * cw296 * pow(cw260 / cw296, (296 - t) / (296 - 260));
    const double vf2 = POW2(nu - 370);
    const double vf6 = POW3(vf2);
    const double fscal = 36100 / (vf2 + vf6 * 1e-8 + 36100) * -.25 + 1;
    const double ctwfrn = cwfrn * fscal;
    const double a1 = nu * u * tanh(.7193876 / t * nu);
    const double a2 = 296 / t;
    const double a3 = p / P0 * (q * ctwslf + (1 - q) * ctwfrn) * 1e-20;
    return a1 * a2 * a3;
  }
  return 0;
}

/* ctmn2 (src/jurassic.c:2796-2859; formula :2841-2858). */
double orc_ctmn2(double nu, double p, double t) {
  const double q_n2 = 0.79, t0 = 273.0, tr = 296.0;
  if (nu < JUR_CTM_N2_NU[0] || nu > JUR_CTM_N2_NU[97]) return 0;
  const int idx = orc_locate_reg(JUR_CTM_N2_NU, 98, nu);
  const double b = LIN(JUR_CTM_N2_NU[idx], JUR_CTM_N2_B[idx], JUR_CTM_N2_NU[idx + 1],
                       JUR_CTM_N2_B[idx + 1], nu);
  const double beta = LIN(JUR_CTM_N2_NU[idx], JUR_CTM_N2_BETA[idx],
                          JUR_CTM_N2_NU[idx + 1], JUR_CTM_N2_BETA[idx + 1], nu);
  return 0.1 * POW2(p / P0 * t0 / t) * exp(beta * (1 / tr - 1 / t)) * q_n2 * b *
         (q_n2 + (1 - q_n2) * (1.294 - 0.4545 * t / tr));
}

/* ctmo2 (src/jurassic.c:2861-2918; formula :2903-2917). */
double orc_ctmo2(double nu, double p, double t) {
  const double q_o2 = 0.21, t0 = 273, tr = 296;
  if (nu < JUR_CTM_O2_NU[0] || nu > JUR_CTM_O2_NU[89]) return 0;
  const int idx = orc_locate_reg(JUR_CTM_O2_NU, 90, nu);
  const double b = LIN(JUR_CTM_O2_NU[idx], JUR_CTM_O2_B[idx], JUR_CTM_O2_NU[idx + 1],
                       JUR_CTM_O2_B[idx + 1], nu);
  const double beta = LIN(JUR_CTM_O2_NU[idx], JUR_CTM_O2_BETA[idx],
                          JUR_CTM_O2_NU[idx + 1], JUR_CTM_O2_BETA[idx + 1], nu);
  return 0.1 * POW2(p / P0 * t0 / t) * exp(beta * (1 / tr - 1 / t)) * q_o2 * b;
}

static int find_emitter(const ctl_t *ctl, const char *name) {
  for (int ig = 0; ig < ctl->ng; ig++)
    if (strcasecmp(ctl->emitter[ig], name) == 0) return ig;
  return -1;
}

/* formod_continua (src/jurassic.c:3075-3117). */
static void formod_continua(const ctl_t *ctl, const orc_los_t *los, int ip,
                            double *beta) {
  const int nd = ctl->nd, ng = ctl->ng;
  for (int id = 0; id < nd; id++) beta[id] = los->k[ip * nd + id];
  if (ctl->ctm_co2) {
    const int ig = find_emitter(ctl, "CO2");
    if (ig >= 0)
      for (int id = 0; id < nd; id++)
        beta[id] += orc_ctmco2(ctl->nu[id], los->p[ip], los->t[ip],
                               los->u[ip * ng + ig]) / los->ds[ip];
  }
  if (ctl->ctm_h2o) {
    const int ig = find_emitter(ctl, "H2O");
    if (ig >= 0)
      for (int id = 0; id < nd; id++)
        beta[id] += orc_ctmh2o(ctl->nu[id], los->p[ip], los->t[ip],
                               los->q[ip * ng + ig], los->u[ip * ng + ig]) /
                    los->ds[ip];
  }
  if (ctl->ctm_n2)
    for (int id = 0; id < nd; id++)
      beta[id] += orc_ctmn2(ctl->nu[id], los->p[ip], los->t[ip]);
  if (ctl->ctm_o2)
    for (int id = 0; id < nd; id++)
      beta[id] += orc_ctmo2(ctl->nu[id], los->p[ip], los->t[ip]);
}

/* ------------------------------------------------------------------ */
/* Emissivity table lookups (src/jurassic.c:3710-3967).                */
/* ------------------------------------------------------------------ */

/* intpol_tbl_eps on one column (src/jurassic.c:3895-3929). */
double orc_intpol_col_eps(const float *u, const float *eps, int nu, double x) {
  if (x < u[0]) return LIN(0, 0, u[0], eps[0], x);
  else if (x > u[nu - 1]) {
    const double a = log(1 - eps[nu - 1]) / u[nu - 1];
    return 1 - exp(a * x);
  } else {
    const int idx = orc_locate_tbl(u, nu, x);
    return LIN(u[idx], eps[idx], u[idx + 1], eps[idx + 1], x);
  }
}

/* intpol_tbl_u on one column (src/jurassic.c:3933-3967). */
double orc_intpol_col_u(const float *u, const float *eps, int nu, double x) {
  if (x < eps[0]) return LIN(0, 0, eps[0], u[0], x);
  else if (x > eps[nu - 1]) {
    const double a = log(1 - eps[nu - 1]) / u[nu - 1];
    return log(1 - x) / a;
  } else {
    const int idx = orc_locate_tbl(eps, nu, x);
    return LIN(eps[idx], u[idx], eps[idx + 1], u[idx + 1], x);
  }
}

#define TB_EPS(tb, ip, it, x) orc_intpol_col_eps((tb)->u[ip][it], (tb)->eps[ip][it], (tb)->nu[ip][it], x)
#define TB_U(tb, ip, it, x) orc_intpol_col_u((tb)->u[ip][it], (tb)->eps[ip][it], (tb)->nu[ip][it], x)

/* intpol_tbl_ega (src/jurassic.c:3799-3891). tau_path is [nd][ng]. */
static void intpol_tbl_ega(const ctl_t *ctl, const orc_ctx_t *c,
                           const orc_los_t *los, int ip, double *tau_path,
                           double *tau_seg) {
  const int nd = ctl->nd, ngas = ctl->ng;
  double eps, u;
  for (int id = 0; id < nd; id++) {
    tau_seg[id] = 1;
    for (int ig = 0; ig < ngas; ig++) {
      const orc_tab_t *tb = &c->tab[id * ngas + ig];
      double *tp = &tau_path[id * ngas + ig];
      if (tb->np < 30) eps = 0;
      else if (*tp < 1e-9) eps = 1;
      else {
        const int ipr = orc_locate_irr(tb->p, tb->np, los->p[ip]);
        const int it0 = orc_locate_reg(tb->t[ipr], tb->nt[ipr], los->t[ip]);
        const int it1 = orc_locate_reg(tb->t[ipr + 1], tb->nt[ipr + 1], los->t[ip]);
        if (tb->nt[ipr] < 2 || tb->nt[ipr + 1] < 2 || tb->nu[ipr][it0] < 2 ||
            tb->nu[ipr][it0 + 1] < 2 || tb->nu[ipr + 1][it1] < 2 ||
            tb->nu[ipr + 1][it1 + 1] < 2)
          eps = 0;
        else {
          const double useg = los->u[ip * ngas + ig];
          u = TB_U(tb, ipr, it0, 1 - *tp);
          double eps00 = TB_EPS(tb, ipr, it0, u + useg);
          u = TB_U(tb, ipr, it0 + 1, 1 - *tp);
          double eps01 = TB_EPS(tb, ipr, it0 + 1, u + useg);
          u = TB_U(tb, ipr + 1, it1, 1 - *tp);
          double eps10 = TB_EPS(tb, ipr + 1, it1, u + useg);
          u = TB_U(tb, ipr + 1, it1 + 1, 1 - *tp);
          double eps11 = TB_EPS(tb, ipr + 1, it1 + 1, u + useg);
          eps00 = LIN(tb->t[ipr][it0], eps00, tb->t[ipr][it0 + 1], eps01, los->t[ip]);
          eps11 = LIN(tb->t[ipr + 1][it1], eps10, tb->t[ipr + 1][it1 + 1], eps11, los->t[ip]);
          eps00 = LIN(tb->p[ipr], eps00, tb->p[ipr + 1], eps11, los->p[ip]);
          eps00 = MAX(MIN(eps00, 1), 0);
          eps = 1 - (1 - eps00) / *tp;
        }
      }
      *tp *= (1 - eps);
      tau_seg[id] *= (1 - eps);
    }
  }
}

/* intpol_tbl_cga (src/jurassic.c:3710-3795): Curtis-Godson path means,
   LOGX in pressure. */
static void intpol_tbl_cga(const ctl_t *ctl, const orc_ctx_t *c,
                           const orc_los_t *los, int ip, double *tau_path,
                           double *tau_seg) {
  const int nd = ctl->nd, ngas = ctl->ng;
  double eps;
  for (int id = 0; id < nd; id++) {
    tau_seg[id] = 1;
    for (int ig = 0; ig < ngas; ig++) {
      const orc_tab_t *tb = &c->tab[id * ngas + ig];
      double *tp = &tau_path[id * ngas + ig];
      const int a = ip * ngas + ig;
      if (tb->np < 30) eps = 0;
      else if (*tp < 1e-9) eps = 1;
      else {
        const int ipr = orc_locate_irr(tb->p, tb->np, los->cgp[a]);
        const int it0 = orc_locate_reg(tb->t[ipr], tb->nt[ipr], los->cgt[a]);
        const int it1 = orc_locate_reg(tb->t[ipr + 1], tb->nt[ipr + 1], los->cgt[a]);
        if (tb->nt[ipr] < 2 || tb->nt[ipr + 1] < 2 || tb->nu[ipr][it0] < 2 ||
            tb->nu[ipr][it0 + 1] < 2 || tb->nu[ipr + 1][it1] < 2 ||
            tb->nu[ipr + 1][it1 + 1] < 2)
          eps = 0;
        else {
          double eps00 = TB_EPS(tb, ipr, it0, los->cgu[a]);
          double eps01 = TB_EPS(tb, ipr, it0 + 1, los->cgu[a]);
          double eps10 = TB_EPS(tb, ipr + 1, it1, los->cgu[a]);
          double eps11 = TB_EPS(tb, ipr + 1, it1 + 1, los->cgu[a]);
          eps00 = LIN(tb->t[ipr][it0], eps00, tb->t[ipr][it0 + 1], eps01, los->cgt[a]);
          eps11 = LIN(tb->t[ipr + 1][it1], eps10, tb->t[ipr + 1][it1 + 1], eps11, los->cgt[a]);
          eps00 = LOGX(tb->p[ipr], eps00, tb->p[ipr + 1], eps11, los->cgp[a]);
          eps00 = MAX(MIN(eps00, 1), 0);
          eps = 1 - (1 - eps00) / *tp;
        }
      }
      *tp *= (1 - eps);
      tau_seg[id] *= (1 - eps);
    }
  }
}

/* formod_srcfunc (src/jurassic.c:3483-3496). */
static void formod_srcfunc(const ctl_t *ctl, const orc_ctx_t *c, double t,
                           double *src) {
  const int it = orc_locate_reg(c->st, TBLNS, t);
  for (int id = 0; id < ctl->nd; id++)
    src[id] = LIN(c->st[it], c->sr[it * ctl->nd + id], c->st[it + 1],
                  c->sr[(it + 1) * ctl->nd + id], t);
}

double orc_planck(double t, double nu) { return C1 * POW3(nu) / expm1(C2 * nu / t); }

double orc_bright(double rad, double nu) { return C2 * nu / log1p(C1 * POW3(nu) / rad); }

/* sza (src/jurassic.c:5183-5219). */
double orc_sza(double sec, double lon, double lat) {
  const double D = sec / 86400 - 0.5;
  const double g = DEG2RAD(357.529 + 0.98560028 * D);
  const double q = 280.459 + 0.98564736 * D;
  const double L = DEG2RAD(q + 1.915 * sin(g) + 0.020 * sin(2 * g));
  const double e = DEG2RAD(23.439 - 0.00000036 * D);
  const double dec = asin(sin(e) * sin(L));
  const double ra = atan2(cos(e) * sin(L), cos(L));
  const double GMST = 18.697374558 + 24.06570982441908 * D;
  const double LST = GMST + lon / 15;
  const double h = LST / 12 * M_PI - ra;
  const double latr = DEG2RAD(lat);
  return RAD2DEG(acos(sin(latr) * sin(dec) + cos(latr) * cos(dec) * cos(h)));
}

/* formod_pencil (src/jurassic.c:3196-3338). */
static int formod_pencil(const ctl_t *ctl, const orc_ctx_t *c,
                         const atm_t *atm, obs_t *obs, int ir,
                         orc_los_t *los) {
  const int nd = ctl->nd, ngas = ctl->ng;
  double beta_ctm[ND], rad[ND], tau[ND], tau_refl[ND], tau_gas[ND], x0[3], x1[3];
  double *tau_path = malloc((size_t)(nd * ngas + 1) * sizeof(double));

  for (int id = 0; id < nd; id++) {
    rad[id] = 0;
    tau[id] = 1;
    for (int ig = 0; ig < ngas; ig++) tau_path[id * ngas + ig] = 1;
  }

  int rc = orc_raytrace(ctl, atm, obs, los, ir);
  if (rc) { free(tau_path); return rc; }

  for (int ip = 0; ip < los->np; ip++) {
    if (ctl->formod == 0) intpol_tbl_cga(ctl, c, los, ip, tau_path, tau_gas);
    else intpol_tbl_ega(ctl, c, los, ip, tau_path, tau_gas);
    formod_continua(ctl, los, ip, beta_ctm);
    formod_srcfunc(ctl, c, los->t[ip], &los->src[ip * nd]);
    for (int id = 0; id < nd; id++) {
      /* A skipped segment keeps eps = 0 here; in the reference it is left
         uninitialised and only the reflection term reads it. */
      los->eps[ip * nd + id] = 0;
      if (tau_gas[id] > 0) {
        los->eps[ip * nd + id] = 1 - tau_gas[id] * exp(-beta_ctm[id] * los->ds[ip]);
        rad[id] += los->src[ip * nd + id] * los->eps[ip * nd + id] * tau[id];
        tau[id] *= (1 - los->eps[ip * nd + id]);
      }
    }
  }

  if (ctl->sftype >= 1 && los->sft > 0) {
    double src_sf[ND];
    formod_srcfunc(ctl, c, los->sft, src_sf);
    for (int id = 0; id < nd; id++) rad[id] += los->sfeps[id] * src_sf[id] * tau[id];
    int refl = 0;
    if (ctl->sftype >= 2)
      for (int id = 0; id < nd; id++)
        if (los->sfeps[id] < 1) { refl = 1; break; }
    if (refl) {
      for (int id = 0; id < nd; id++) tau_refl[id] = 1;
      for (int ip = los->np - 1; ip >= 0; ip--)
        for (int id = 0; id < nd; id++) {
          rad[id] += los->src[ip * nd + id] * los->eps[ip * nd + id] * tau_refl[id] *
                     tau[id] * (1 - los->sfeps[id]);
          tau_refl[id] *= (1 - los->eps[ip * nd + id]);
        }
      if (ctl->sftype >= 3) {
        double sza2;
        if (ctl->sfsza < 0)
          sza2 = orc_sza(obs->time[ir], los->lon[los->np - 1], los->lat[los->np - 1]);
        else
          sza2 = ctl->sfsza;
        if (sza2 < 89.999) {
          orc_geo2cart(los->z[los->np - 1], los->lon[los->np - 1], los->lat[los->np - 1], x0);
          orc_geo2cart(los->z[0], los->lon[0], los->lat[0], x1);
          for (int i = 0; i < 3; i++) x1[i] -= x0[i];
          const double cosa = DOTP(x0, x1) / NORM(x0) / NORM(x1);
          const double rcos = cosa / cos(DEG2RAD(sza2));
          for (int id = 0; id < nd; id++)
            rad[id] += 6.764e-5 / (2. * M_PI) * orc_planck(TSUN, ctl->nu[id]) *
                       tau_refl[id] * (1 - los->sfeps[id]) * tau[id] * rcos;
        }
      }
    }
  }

  for (int id = 0; id < nd; id++) {
    obs->rad[id][ir] = rad[id];
    obs->tau[id][ir] = tau[id];
  }
  free(tau_path);
  return 0;
}

/* formod_fov (src/jurassic.c:3121-3192). */
static int formod_fov(const ctl_t *ctl, obs_t *obs) {
  if (ctl->fov[0] == '-') return 0;
  double *dz = malloc(NSHAPE * sizeof(double)), *w = malloc(NSHAPE * sizeof(double));
  int n, rc = read_shape(ctl->fov, dz, w, &n);
  if (rc) { free(dz); free(w); return rc; }
  obs_t *obs2 = malloc(sizeof(obs_t));
  orc_copy_obs(ctl, obs2, obs, 0);
  double rad[ND][2 * NFOV + 1], tau[ND][2 * NFOV + 1], z[2 * NFOV + 1];
  for (int ir = 0; ir < obs->nr && rc == 0; ir++) {
    int nz = 0;
    for (int ir2 = MAX(ir - NFOV, 0); ir2 < MIN(ir + 1 + NFOV, obs->nr); ir2++)
      if (obs->time[ir2] == obs->time[ir]) {
        z[nz] = obs2->vpz[ir2];
        for (int id = 0; id < ctl->nd; id++) {
          rad[id][nz] = obs2->rad[id][ir2];
          tau[id][nz] = obs2->tau[id][ir2];
        }
        nz++;
      }
    if (nz < 2) { rc = fail(JUR_ERR_FOV, "Cannot apply FOV convolution!"); break; }
    double wsum = 0;
    for (int id = 0; id < ctl->nd; id++) {
      obs->rad[id][ir] = 0;
      obs->tau[id][ir] = 0;
    }
    for (int i = 0; i < n; i++) {
      const double zfov = obs->vpz[ir] + dz[i];
      const int idx = orc_locate_irr(z, nz, zfov);
      for (int id = 0; id < ctl->nd; id++) {
        obs->rad[id][ir] += w[i] * LIN(z[idx], rad[id][idx], z[idx + 1], rad[id][idx + 1], zfov);
        obs->tau[id][ir] += w[i] * LIN(z[idx], tau[id][idx], z[idx + 1], tau[id][idx + 1], zfov);
      }
      wsum += w[i];
    }
    for (int id = 0; id < ctl->nd; id++) {
      obs->rad[id][ir] /= wsum;
      obs->tau[id][ir] /= wsum;
    }
  }
  free(obs2);
  free(dz);
  free(w);
  return rc;
}

int orc_formod_fov(const ctl_t *ctl, obs_t *obs) { return formod_fov(ctl, obs); }

/* formod (src/jurassic.c:3026-3071). */
int orc_formod(const ctl_t *ctl, atm_t *atm, obs_t *obs) {
  if (ctl->nd < 0 || ctl->nd > ND || ctl->ng < 0 || ctl->ng > NG ||
      obs->nr < 0 || obs->nr > NR || atm->np < 2 || atm->np > NP)
    return fail(JUR_ERR_ARG, "Dimension out of range!");
  if (ctl->formod == 2) return fail(JUR_ERR_UNSUPPORTED, "RFM is not supported!");
  orc_ctx_t *c;
  int rc = get_ctx(ctl, &c);
  if (rc) return rc;

  int *mask = malloc((size_t)(ND * NR) * sizeof(int));
  for (int id = 0; id < ctl->nd; id++)
    for (int ir = 0; ir < obs->nr; ir++)
      mask[id * NR + ir] = !isfinite(obs->rad[id][ir]);

  orc_hydrostatic(ctl, atm);

  orc_los_t *los = orc_los_alloc(ctl->nd, ctl->ng);
  for (int ir = 0; ir < obs->nr && rc == 0; ir++)
    rc = formod_pencil(ctl, c, atm, obs, ir, los);
  orc_los_free(los);

  if (rc == 0) rc = formod_fov(ctl, obs);

  if (rc == 0 && ctl->write_bbt)
    for (int id = 0; id < ctl->nd; id++)
      for (int ir = 0; ir < obs->nr; ir++)
        obs->rad[id][ir] = orc_bright(obs->rad[id][ir], ctl->nu[id]);

  for (int id = 0; id < ctl->nd; id++)
    for (int ir = 0; ir < obs->nr; ir++)
      if (mask[id * NR + ir]) obs->rad[id][ir] = NAN;
  free(mask);
  return rc;
}

int orc_formod_batch(const ctl_t *ctl, atm_t *atm, obs_t *obs, int n,
                     int nthreads) {
  if (n <= 0) return 0;
  int rc = orc_formod(ctl, &atm[0], &obs[0]); /* warm call initialises the context */
  if (rc) return rc;
  int bad = 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : bad)
  for (int i = 1; i < n; i++)
    if (orc_formod(ctl, &atm[i], &obs[i])) bad |= 1;
  return bad ? fail(JUR_ERR_ARG, "formod failed in batch") : 0;
}

/* ------------------------------------------------------------------ */
/* State and measurement vectors (src/jurassic.c:29-104, 2921-3009,     */
/* 4174-4197, 5896-5976).                                              */
/* ------------------------------------------------------------------ */

#define IDXP 0
#define IDXT 1
#define IDXQ(ig) (2 + (ig))
#define IDXK(iw) (2 + ctl->ng + (iw))
#define IDXCLZ (2 + ctl->ng + ctl->nw)
#define IDXCLDZ (3 + ctl->ng + ctl->nw)
#define IDXCLK(icl) (4 + ctl->ng + ctl->nw + (icl))
#define IDXSFZ (4 + ctl->ng + ctl->nw + ctl->ncl)
#define IDXSFP (5 + ctl->ng + ctl->nw + ctl->ncl)
#define IDXSFT (6 + ctl->ng + ctl->nw + ctl->ncl)
#define IDXSFEPS(isf) (7 + ctl->ng + ctl->nw + ctl->ncl + (isf))

static void add_x(double v, int q, int ip, double *x, int *iqa, int *ipa, size_t *n) {
  if (x) x[*n] = v;
  if (iqa) iqa[*n] = q;
  if (ipa) ipa[*n] = ip;
  (*n)++;
}

size_t orc_atm2x(const ctl_t *ctl, const atm_t *atm, double *x, int *iqa, int *ipa) {
  size_t n = 0;
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->retp_zmin && atm->z[ip] <= ctl->retp_zmax)
      add_x(atm->p[ip], IDXP, ip, x, iqa, ipa, &n);
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->rett_zmin && atm->z[ip] <= ctl->rett_zmax)
      add_x(atm->t[ip], IDXT, ip, x, iqa, ipa, &n);
  for (int ig = 0; ig < ctl->ng; ig++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retq_zmin[ig] && atm->z[ip] <= ctl->retq_zmax[ig])
        add_x(atm->q[ig][ip], IDXQ(ig), ip, x, iqa, ipa, &n);
  for (int iw = 0; iw < ctl->nw; iw++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retk_zmin[iw] && atm->z[ip] <= ctl->retk_zmax[iw])
        add_x(atm->k[iw][ip], IDXK(iw), ip, x, iqa, ipa, &n);
  if (ctl->ret_clz) add_x(atm->clz, IDXCLZ, 0, x, iqa, ipa, &n);
  if (ctl->ret_cldz) add_x(atm->cldz, IDXCLDZ, 0, x, iqa, ipa, &n);
  if (ctl->ret_clk)
    for (int icl = 0; icl < ctl->ncl; icl++) add_x(atm->clk[icl], IDXCLK(icl), 0, x, iqa, ipa, &n);
  if (ctl->ret_sfz) add_x(atm->sfz, IDXSFZ, 0, x, iqa, ipa, &n);
  if (ctl->ret_sfp) add_x(atm->sfp, IDXSFP, 0, x, iqa, ipa, &n);
  if (ctl->ret_sft) add_x(atm->sft, IDXSFT, 0, x, iqa, ipa, &n);
  if (ctl->ret_sfeps)
    for (int isf = 0; isf < ctl->nsf; isf++) add_x(atm->sfeps[isf], IDXSFEPS(isf), 0, x, iqa, ipa, &n);
  return n;
}

void orc_x2atm(const ctl_t *ctl, const double *x, atm_t *atm) {
  size_t n = 0;
#define TAKE(dst) ((dst) = x[n++])
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->retp_zmin && atm->z[ip] <= ctl->retp_zmax) TAKE(atm->p[ip]);
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->rett_zmin && atm->z[ip] <= ctl->rett_zmax) TAKE(atm->t[ip]);
  for (int ig = 0; ig < ctl->ng; ig++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retq_zmin[ig] && atm->z[ip] <= ctl->retq_zmax[ig]) TAKE(atm->q[ig][ip]);
  for (int iw = 0; iw < ctl->nw; iw++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retk_zmin[iw] && atm->z[ip] <= ctl->retk_zmax[iw]) TAKE(atm->k[iw][ip]);
  if (ctl->ret_clz) TAKE(atm->clz);
  if (ctl->ret_cldz) TAKE(atm->cldz);
  if (ctl->ret_clk)
    for (int icl = 0; icl < ctl->ncl; icl++) TAKE(atm->clk[icl]);
  if (ctl->ret_sfz) TAKE(atm->sfz);
  if (ctl->ret_sfp) TAKE(atm->sfp);
  if (ctl->ret_sft) TAKE(atm->sft);
  if (ctl->ret_sfeps)
    for (int isf = 0; isf < ctl->nsf; isf++) TAKE(atm->sfeps[isf]);
#undef TAKE
}

size_t orc_obs2y(const ctl_t *ctl, const obs_t *obs, double *y, int *ida, int *ira) {
  size_t m = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (isfinite(obs->rad[id][ir])) {
        if (y) y[m] = obs->rad[id][ir];
        if (ida) ida[m] = id;
        if (ira) ira[m] = ir;
        m++;
      }
  return m;
}

void orc_y2obs(const ctl_t *ctl, const double *y, obs_t *obs) {
  size_t m = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (isfinite(obs->rad[id][ir])) obs->rad[id][ir] = y[m++];
}

void orc_copy_atm(const ctl_t *ctl, atm_t *d, const atm_t *s, int init) {
  const size_t sz = (size_t)s->np * sizeof(double);
  d->np = s->np;
  memcpy(d->time, s->time, sz);
  memcpy(d->z, s->z, sz);
  memcpy(d->lon, s->lon, sz);
  memcpy(d->lat, s->lat, sz);
  memcpy(d->p, s->p, sz);
  memcpy(d->t, s->t, sz);
  for (int ig = 0; ig < ctl->ng; ig++) memcpy(d->q[ig], s->q[ig], sz);
  for (int iw = 0; iw < ctl->nw; iw++) memcpy(d->k[iw], s->k[iw], sz);
  d->clz = s->clz;
  d->cldz = s->cldz;
  for (int icl = 0; icl < ctl->ncl; icl++) d->clk[icl] = s->clk[icl];
  d->sfz = s->sfz;
  d->sfp = s->sfp;
  d->sft = s->sft;
  for (int isf = 0; isf < ctl->nsf; isf++) d->sfeps[isf] = s->sfeps[isf];
  if (init)
    for (int ip = 0; ip < d->np; ip++) {
      d->p[ip] = 0;
      d->t[ip] = 0;
      for (int ig = 0; ig < ctl->ng; ig++) d->q[ig][ip] = 0;
      for (int iw = 0; iw < ctl->nw; iw++) d->k[iw][ip] = 0;
      d->clz = 0;
      d->cldz = 0;
      for (int icl = 0; icl < ctl->ncl; icl++) d->clk[icl] = 0;
      d->sfz = 0;
      d->sfp = 0;
      d->sft = 0;
      for (int isf = 0; isf < ctl->nsf; isf++) d->sfeps[isf] = 1;
    }
}

void orc_copy_obs(const ctl_t *ctl, obs_t *d, const obs_t *s, int init) {
  const size_t sz = (size_t)s->nr * sizeof(double);
  d->nr = s->nr;
  memcpy(d->time, s->time, sz);
  memcpy(d->obsz, s->obsz, sz);
  memcpy(d->obslon, s->obslon, sz);
  memcpy(d->obslat, s->obslat, sz);
  memcpy(d->vpz, s->vpz, sz);
  memcpy(d->vplon, s->vplon, sz);
  memcpy(d->vplat, s->vplat, sz);
  memcpy(d->tpz, s->tpz, sz);
  memcpy(d->tplon, s->tplon, sz);
  memcpy(d->tplat, s->tplat, sz);
  for (int id = 0; id < ctl->nd; id++) memcpy(d->rad[id], s->rad[id], sz);
  for (int id = 0; id < ctl->nd; id++) memcpy(d->tau[id], s->tau[id], sz);
  if (init)
    for (int id = 0; id < ctl->nd; id++)
      for (int ir = 0; ir < d->nr; ir++)
        if (isfinite(d->rad[id][ir])) {
          d->rad[id][ir] = 0;
          d->tau[id][ir] = 0;
        }
}

/* Perturbation size of kernel() (src/jurassic.c:4046-4068). */
static int fd_step(const ctl_t *ctl, int iq, double x0, double *h) {
  if (iq == IDXP) *h = MAX(fabs(0.01 * x0), 1e-7);
  else if (iq == IDXT) *h = 1.0;
  else if (iq >= IDXQ(0) && iq < IDXQ(ctl->ng)) *h = MAX(fabs(0.01 * x0), 1e-15);
  else if (iq >= IDXK(0) && iq < IDXK(ctl->nw)) *h = 1e-4;
  else if (iq == IDXCLZ || iq == IDXCLDZ) *h = 1.0;
  else if (iq >= IDXCLK(0) && iq < IDXCLK(ctl->ncl)) *h = 1e-4;
  else if (iq == IDXSFZ) *h = 0.1;
  else if (iq == IDXSFP) *h = 10.0;
  else if (iq == IDXSFT) *h = 1.0;
  else if (iq >= IDXSFEPS(0) && iq < IDXSFEPS(ctl->nsf)) *h = 1e-2;
  else return fail(JUR_ERR_ARG, "Cannot set perturbation size!");
  return 0;
}

static int kernel_impl(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                       size_t m, size_t n, int nthreads, int central, double scale) {
  int rc = orc_formod(ctl, atm, obs);
  if (rc) return rc;
  double *x0 = malloc((n + 1) * sizeof(double));
  double *yy0 = malloc((m + 1) * sizeof(double));
  int *iqa = malloc((n + 1) * sizeof(int));
  if (orc_atm2x(ctl, atm, x0, iqa, NULL) != n || orc_obs2y(ctl, obs, yy0, NULL, NULL) != m) {
    free(x0); free(yy0); free(iqa);
    return fail(JUR_ERR_ARG, "Kernel matrix size mismatch!");
  }
  memset(k, 0, m * n * sizeof(double));
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : bad)
  for (size_t j = 0; j < n; j++) {
    double h;
    if (fd_step(ctl, iqa[j], x0[j], &h)) { bad |= 1; continue; }
    h *= scale;
    double *x1 = malloc((n + 1) * sizeof(double));
    double *yy1 = malloc((m + 1) * sizeof(double));
    double *yy2 = malloc((m + 1) * sizeof(double));
    atm_t *atm1 = malloc(sizeof(atm_t));
    obs_t *obs1 = malloc(sizeof(obs_t));
    memcpy(x1, x0, n * sizeof(double));
    x1[j] = x1[j] + h;
    orc_copy_atm(ctl, atm1, atm, 0);
    orc_copy_obs(ctl, obs1, obs, 0);
    orc_x2atm(ctl, x1, atm1);
    if (orc_formod(ctl, atm1, obs1)) bad |= 1;
    orc_obs2y(ctl, obs1, yy1, NULL, NULL);
    if (central) {
      memcpy(x1, x0, n * sizeof(double));
      x1[j] = x1[j] - h;
      orc_copy_atm(ctl, atm1, atm, 0);
      orc_copy_obs(ctl, obs1, obs, 0);
      orc_x2atm(ctl, x1, atm1);
      if (orc_formod(ctl, atm1, obs1)) bad |= 1;
      orc_obs2y(ctl, obs1, yy2, NULL, NULL);
      for (size_t i = 0; i < m; i++) k[i * n + j] = (yy1[i] - yy2[i]) / (2 * h);
    } else {
      for (size_t i = 0; i < m; i++) k[i * n + j] = (yy1[i] - yy0[i]) / h;
    }
    free(x1); free(yy1); free(yy2); free(atm1); free(obs1);
  }
  free(x0); free(yy0); free(iqa);
  return bad ? fail(JUR_ERR_ARG, "kernel column failed") : 0;
}

/* kernel (src/jurassic.c:4004-4099). */
int orc_kernel(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m,
               size_t n, int nthreads) {
  return kernel_impl(ctl, atm, obs, k, m, n, nthreads, 0, 1.0);
}

int orc_kernel_central(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                       size_t m, size_t n, int nthreads) {
  return kernel_impl(ctl, atm, obs, k, m, n, nthreads, 1, 1.0);
}

int orc_kernel_scaled(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                      size_t m, size_t n, int nthreads, int central, double scale) {
  return kernel_impl(ctl, atm, obs, k, m, n, nthreads, central, scale);
}

int orc_trace_ray(const ctl_t *ctl, const atm_t *atm, obs_t *obs, int ir,
                  double *z, double *p, double *t, double *ds, double *u,
                  double *lon, double *lat, double *sft) {
  orc_los_t *los = orc_los_alloc(ctl->nd, ctl->ng);
  int rc = orc_raytrace(ctl, atm, obs, los, ir);
  if (rc) { orc_los_free(los); return -rc; }
  int np = los->np;
  for (int ip = 0; ip < np; ip++) {
    if (z) z[ip] = los->z[ip];
    if (p) p[ip] = los->p[ip];
    if (t) t[ip] = los->t[ip];
    if (ds) ds[ip] = los->ds[ip];
    if (lon) lon[ip] = los->lon[ip];
    if (lat) lat[ip] = los->lat[ip];
    if (u)
      for (int ig = 0; ig < ctl->ng; ig++) u[ip * ctl->ng + ig] = los->u[ip * ctl->ng + ig];
  }
  if (sft) *sft = los->sft;
  orc_los_free(los);
  return np;
}

/* ------------------------------------------------------------------ */
/* Optimal estimation (src/retrieval.c:340-964).                       */
/* ------------------------------------------------------------------ */

/* Textbook Cholesky A = L L^T in place (lower triangle), row-major n x n.
   GSL's blocked variant rounds differently; results agree to rounding. */
static int chol_decomp(double *a, size_t n) {
  for (size_t j = 0; j < n; j++) {
    double s = a[j * n + j];
    for (size_t k = 0; k < j; k++) s -= a[j * n + k] * a[j * n + k];
    if (!(s > 0)) return fail(JUR_ERR_ARG, "matrix not positive definite");
    const double l = sqrt(s);
    a[j * n + j] = l;
    for (size_t i = j + 1; i < n; i++) {
      double v = a[i * n + j];
      for (size_t k = 0; k < j; k++) v -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = v / l;
    }
  }
  return 0;
}

static void chol_solve(const double *l, size_t n, const double *b, double *x) {
  for (size_t i = 0; i < n; i++) {
    double v = b[i];
    for (size_t k = 0; k < i; k++) v -= l[i * n + k] * x[k];
    x[i] = v / l[i * n + i];
  }
  for (size_t ii = n; ii-- > 0;) {
    double v = x[ii];
    for (size_t k = ii + 1; k < n; k++) v -= l[k * n + ii] * x[k];
    x[ii] = v / l[ii * n + ii];
  }
}

/* matrix_invert (src/retrieval.c:373-399): diagonal shortcut, else Cholesky. */
static int matrix_invert(double *a, size_t n) {
  int diag = 1;
  for (size_t i = 0; i < n && diag; i++)
    for (size_t j = i + 1; j < n; j++)
      if (a[i * n + j] != 0) { diag = 0; break; }
  if (diag) {
    for (size_t i = 0; i < n; i++) a[i * n + i] = 1 / a[i * n + i];
    return 0;
  }
  double *l = malloc(n * n * sizeof(double));
  memcpy(l, a, n * n * sizeof(double));
  int rc = chol_decomp(l, n);
  if (rc) { free(l); return rc; }
  double *e = calloc(n, sizeof(double)), *col = malloc(n * sizeof(double));
  for (size_t j = 0; j < n; j++) {
    memset(e, 0, n * sizeof(double));
    e[j] = 1;
    chol_solve(l, n, e, col);
    for (size_t i = 0; i < n; i++) a[i * n + j] = col[i];
  }
  free(l); free(e); free(col);
  return 0;
}

int orc_set_cov_apr(const orc_ret_t *ret, const ctl_t *ctl, const atm_t *atm,
                    const int *iqa, const int *ipa, size_t n, double *s_a) {
  double x0[3], x1[3];
  double *xa = malloc((n + 1) * sizeof(double));
  orc_atm2x(ctl, atm, xa, NULL, NULL);
  for (size_t i = 0; i < n; i++) {
    if (iqa[i] == IDXP) xa[i] = ret->err_press / 100 * xa[i];
    if (iqa[i] == IDXT) xa[i] = ret->err_temp;
    for (int ig = 0; ig < ctl->ng; ig++)
      if (iqa[i] == IDXQ(ig)) xa[i] = ret->err_q[ig] / 100 * xa[i];
    for (int iw = 0; iw < ctl->nw; iw++)
      if (iqa[i] == IDXK(iw)) xa[i] = ret->err_k[iw];
    if (iqa[i] == IDXCLZ) xa[i] = ret->err_clz;
    if (iqa[i] == IDXCLDZ) xa[i] = ret->err_cldz;
    for (int icl = 0; icl < ctl->ncl; icl++)
      if (iqa[i] == IDXCLK(icl)) xa[i] = ret->err_clk[icl];
    if (iqa[i] == IDXSFZ) xa[i] = ret->err_sfz;
    if (iqa[i] == IDXSFP) xa[i] = ret->err_sfp;
    if (iqa[i] == IDXSFT) xa[i] = ret->err_sft;
    for (int isf = 0; isf < ctl->nsf; isf++)
      if (iqa[i] == IDXSFEPS(isf)) xa[i] = ret->err_sfeps[isf];
  }
  for (size_t i = 0; i < n; i++)
    if (POW2(xa[i]) <= 0) { free(xa); return fail(JUR_ERR_ARG, "Check a priori data (zero standard deviation)!"); }
  memset(s_a, 0, n * n * sizeof(double));
  for (size_t i = 0; i < n; i++) s_a[i * n + i] = POW2(xa[i]);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++)
      if (i != j && iqa[i] == iqa[j]) {
        double cz = 0, ch = 0;
        if (iqa[i] == IDXP) { cz = ret->err_press_cz; ch = ret->err_press_ch; }
        if (iqa[i] == IDXT) { cz = ret->err_temp_cz; ch = ret->err_temp_ch; }
        for (int ig = 0; ig < ctl->ng; ig++)
          if (iqa[i] == IDXQ(ig)) { cz = ret->err_q_cz[ig]; ch = ret->err_q_ch[ig]; }
        for (int iw = 0; iw < ctl->nw; iw++)
          if (iqa[i] == IDXK(iw)) { cz = ret->err_k_cz[iw]; ch = ret->err_k_ch[iw]; }
        if (cz > 0 && ch > 0) {
          orc_geo2cart(0, atm->lon[ipa[i]], atm->lat[ipa[i]], x0);
          orc_geo2cart(0, atm->lon[ipa[j]], atm->lat[ipa[j]], x1);
          const double rho = exp(-DIST(x0, x1) / ch - fabs(atm->z[ipa[i]] - atm->z[ipa[j]]) / cz);
          s_a[i * n + j] = xa[i] * xa[j] * rho;
        }
      }
  free(xa);
  return 0;
}

/* set_cov_meas (src/retrieval.c:924-964): 1/sqrt(noise^2 + formod^2). */
static int set_cov_meas(const orc_ret_t *ret, const ctl_t *ctl, const obs_t *obs,
                        size_t m, double *sig_eps_inv) {
  size_t i = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (isfinite(obs->rad[id][ir])) {
        const double sn = ret->err_noise[id];
        const double sf = fabs(ret->err_formod[id] / 100 * obs->rad[id][ir]);
        sig_eps_inv[i] = 1 / sqrt(POW2(sn) + POW2(sf));
        if (sig_eps_inv[i] <= 0) return fail(JUR_ERR_ARG, "Check measurement errors (zero standard deviation)!");
        i++;
      }
  return i == m ? 0 : fail(JUR_ERR_ARG, "measurement size mismatch");
}

/* cost_function (src/retrieval.c:340-369). */
static double cost_function(const double *dx, const double *dy, const double *s_a_inv,
                            const double *sig_eps_inv, size_t n, size_t m) {
  double chisq_m = 0, chisq_a = 0;
  for (size_t i = 0; i < m; i++) chisq_m += POW2(dy[i] * sig_eps_inv[i]);
  for (size_t i = 0; i < n; i++) {
    double v = 0;
    for (size_t j = 0; j < n; j++) v += s_a_inv[i * n + j] * dx[j];
    chisq_a += dx[i] * v;
  }
  return (chisq_m + chisq_a) / (double)m;
}

typedef struct {
  const ctl_t *ctl;
  atm_t *atm_apr;
  obs_t *obs_meas;
  atm_t *atm_tmp;
  obs_t *obs_tmp;
  int nthreads;
} oe_fwd_t;

static int oe_eval(const ctl_t *ctl, orc_fwd_fn fwd, void *user, atm_t *atm_i,
                   obs_t *obs_i, const double *x, double *y, double *k, size_t m,
                   size_t n) {
  if (fwd) {
    fwd(user, x, y, k);
    return 0;
  }
  int rc = orc_formod(ctl, atm_i, obs_i);
  if (rc) return rc;
  orc_obs2y(ctl, obs_i, y, NULL, NULL);
  if (k) rc = orc_kernel(ctl, atm_i, obs_i, k, m, n, 0);
  return rc;
}

/* Error analysis (src/retrieval.c:668-726) from the last Jacobian k_i. */
static int oe_analysis(const orc_ret_t *ret, const ctl_t *ctl, const obs_t *obs_meas,
                       const double *k_i, const double *s_a_inv, const double *sig_eps_inv,
                       const int *iqa, size_t m, size_t n, const orc_ana_t *ana) {
  double *cov = malloc(n * n * sizeof(double)), *gain = malloc(n * m * sizeof(double));
  double *auxnm = malloc(n * m * sizeof(double)), *a = malloc(n * n * sizeof(double));
  double *sig_noise = malloc(m * sizeof(double)), *sig_formod = malloc(m * sizeof(double));
  int rc;
  /* cov^-1 = K^T S_eps^-1 K + S_a^-1 (matrix_product, :403-444) */
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++) {
      double v = 0;
      for (size_t r = 0; r < m; r++) v += (sig_eps_inv[r] * k_i[r * n + i]) * (sig_eps_inv[r] * k_i[r * n + j]);
      cov[i * n + j] = v + s_a_inv[i * n + j];
    }
  rc = matrix_invert(cov, n);
  if (rc) goto out;
  if (ana->cov) memcpy(ana->cov, cov, n * n * sizeof(double));
  if (ana->err_total)
    for (size_t i = 0; i < n; i++) ana->err_total[i] = sqrt(cov[i * n + i]);
  if (ana->corr)
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++)
        ana->corr[i * n + j] = cov[i * n + j] / sqrt(cov[i * n + i]) / sqrt(cov[j * n + j]);
  /* G = cov K^T S_eps^-1 */
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < m; j++) auxnm[i * m + j] = k_i[j * n + i] * POW2(sig_eps_inv[j]);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < m; j++) {
      double v = 0;
      for (size_t k = 0; k < n; k++) v += cov[i * n + k] * auxnm[k * m + j];
      gain[i * m + j] = v;
    }
  if (ana->gain) memcpy(ana->gain, gain, n * m * sizeof(double));
  /* noise and forward-model errors of set_cov_meas (:924-964), then
     G S G^T = (G S^1/2)(G S^1/2)^T (matrix_product with transpose = 2) */
  {
    size_t r = 0;
    for (int ir = 0; ir < obs_meas->nr; ir++)
      for (int id = 0; id < ctl->nd; id++)
        if (isfinite(obs_meas->rad[id][ir])) {
          sig_noise[r] = ret->err_noise[id];
          sig_formod[r] = fabs(ret->err_formod[id] / 100 * obs_meas->rad[id][ir]);
          r++;
        }
  }
  for (int which = 0; which < 2; which++) {
    const double *sg = which == 0 ? sig_noise : sig_formod;
    double *out = which == 0 ? ana->err_noise : ana->err_formod;
    if (!out) continue;
    for (size_t i = 0; i < n; i++) {
      double v = 0;
      for (size_t j = 0; j < m; j++) v += (gain[i * m + j] * sg[j]) * (gain[i * m + j] * sg[j]);
      out[i] = sqrt(v);
    }
  }
  /* A = G K (:712-716) and analyze_avk (:257-338) */
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++) {
      double v = 0;
      for (size_t k = 0; k < m; k++) v += gain[i * m + k] * k_i[k * n + j];
      a[i * n + j] = v;
    }
  if (ana->avk) memcpy(ana->avk, a, n * n * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    if (ana->cont) {
      double v = 0;
      for (size_t j = 0; j < n; j++)
        if (iqa[j] == iqa[i]) v += a[i * n + j];
      ana->cont[i] = v;
    }
    if (ana->res) ana->res[i] = 1 / a[i * n + i];
  }
out:
  free(cov); free(gain); free(auxnm); free(a); free(sig_noise); free(sig_formod);
  return rc;
}

int orc_optimal_estimation(const orc_ret_t *ret, const ctl_t *ctl,
                           obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                           atm_t *atm_i, double *chisq_hist, int *niter,
                           orc_fwd_fn fwd, void *user) {
  return orc_optimal_estimation_ana(ret, ctl, obs_meas, obs_i, atm_apr, atm_i, chisq_hist, niter, fwd, user,
                                    NULL);
}

int orc_optimal_estimation_ana(const orc_ret_t *ret, const ctl_t *ctl,
                               obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                               atm_t *atm_i, double *chisq_hist, int *niter,
                               orc_fwd_fn fwd, void *user, const orc_ana_t *ana) {
  const size_t m = orc_obs2y(ctl, obs_meas, NULL, NULL, NULL);
  int *iqa = malloc(sizeof(int) * (size_t)(NP * (NG + NW + 2) + 64));
  int *ipa = malloc(sizeof(int) * (size_t)(NP * (NG + NW + 2) + 64));
  const size_t n = orc_atm2x(ctl, atm_apr, NULL, iqa, ipa);
  if (m == 0 || n == 0) { free(iqa); free(ipa); return fail(JUR_ERR_ARG, "Check problem definition!"); }
  double chisq, disq = 0, lmpar = 0.001;
  int rc = 0, it = 0;
  double *a = malloc(n * n * sizeof(double)), *cov = malloc(n * n * sizeof(double));
  double *k_i = malloc(m * n * sizeof(double)), *s_a_inv = malloc(n * n * sizeof(double));
  double *b = malloc(n * sizeof(double)), *dx = malloc(n * sizeof(double));
  double *dy = malloc(m * sizeof(double)), *sig_eps_inv = malloc(m * sizeof(double));
  double *x_a = malloc(n * sizeof(double)), *x_i = malloc(n * sizeof(double));
  double *x_step = malloc(n * sizeof(double)), *y_aux = malloc(m * sizeof(double));
  double *y_i = malloc(m * sizeof(double)), *y_m = malloc(m * sizeof(double));

  orc_copy_atm(ctl, atm_i, atm_apr, 0);
  orc_copy_obs(ctl, obs_i, obs_meas, 0);
  orc_atm2x(ctl, atm_apr, x_a, NULL, NULL);
  orc_atm2x(ctl, atm_i, x_i, NULL, NULL);
  orc_obs2y(ctl, obs_meas, y_m, NULL, NULL);
  rc = oe_eval(ctl, fwd, user, atm_i, obs_i, x_i, y_i, NULL, m, n);
  if (rc) goto done;

  rc = orc_set_cov_apr(ret, ctl, atm_apr, iqa, ipa, n, s_a_inv);
  if (rc) goto done;
  rc = matrix_invert(s_a_inv, n);
  if (rc) goto done;
  rc = set_cov_meas(ret, ctl, obs_meas, m, sig_eps_inv);
  if (rc) goto done;

  for (size_t i = 0; i < n; i++) dx[i] = x_i[i] - x_a[i];
  for (size_t i = 0; i < m; i++) dy[i] = y_m[i] - y_i[i];
  chisq = cost_function(dx, dy, s_a_inv, sig_eps_inv, n, m);
  if (chisq_hist) chisq_hist[0] = chisq;

  rc = oe_eval(ctl, fwd, user, atm_i, obs_i, x_i, y_i, k_i, m, n);
  if (rc) goto done;

  for (it = 1; it <= ret->conv_itmax; it++) {
    const double chisq_old = chisq;
    if (it > 1 && it % ret->kernel_recomp == 0) {
      rc = oe_eval(ctl, fwd, user, atm_i, obs_i, x_i, y_i, k_i, m, n);
      if (rc) goto done;
    }
    if (it == 1 || it % ret->kernel_recomp == 0) {
      /* K^T S_eps^-1 K = (B^1/2 K)^T (B^1/2 K) (matrix_product, retrieval.c:403-444) */
      for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
          double v = 0;
          for (size_t r = 0; r < m; r++)
            v += (sig_eps_inv[r] * k_i[r * n + i]) * (sig_eps_inv[r] * k_i[r * n + j]);
          cov[i * n + j] = v;
        }
    }
    for (size_t i = 0; i < m; i++) y_aux[i] = dy[i] * POW2(sig_eps_inv[i]);
    for (size_t j = 0; j < n; j++) {
      double v = 0;
      for (size_t i = 0; i < m; i++) v += k_i[i * n + j] * y_aux[i];
      b[j] = v;
    }
    for (size_t i = 0; i < n; i++) {
      double v = 0;
      for (size_t j = 0; j < n; j++) v += s_a_inv[i * n + j] * dx[j];
      b[i] = -1.0 * v + b[i];
    }
    for (int it2 = 0; it2 < 20; it2++) {
      for (size_t i = 0; i < n * n; i++) a[i] = s_a_inv[i] * (1 + lmpar) + cov[i];
      rc = chol_decomp(a, n);
      if (rc) goto done;
      chol_solve(a, n, b, x_step);
      for (size_t i = 0; i < n; i++) x_i[i] += x_step[i];
      orc_copy_atm(ctl, atm_i, atm_apr, 0);
      orc_copy_obs(ctl, obs_i, obs_meas, 0);
      orc_x2atm(ctl, x_i, atm_i);
      for (int ip = 0; ip < atm_i->np; ip++) {
        atm_i->p[ip] = MIN(MAX(atm_i->p[ip], 5e-7), 5e4);
        atm_i->t[ip] = MIN(MAX(atm_i->t[ip], 100), 400);
        for (int ig = 0; ig < ctl->ng; ig++) atm_i->q[ig][ip] = MIN(MAX(atm_i->q[ig][ip], 0), 1);
        for (int iw = 0; iw < ctl->nw; iw++) atm_i->k[iw][ip] = MAX(atm_i->k[iw][ip], 0);
      }
      atm_i->clz = MAX(atm_i->clz, 0);
      atm_i->cldz = MAX(atm_i->cldz, 0.1);
      for (int icl = 0; icl < ctl->ncl; icl++) atm_i->clk[icl] = MAX(atm_i->clk[icl], 0);
      atm_i->sfz = MAX(atm_i->sfz, 0);
      atm_i->sfp = MAX(atm_i->sfp, 0);
      atm_i->sft = MIN(MAX(atm_i->sft, 100), 400);
      for (int isf = 0; isf < ctl->nsf; isf++) atm_i->sfeps[isf] = MIN(MAX(atm_i->sfeps[isf], 0), 1);

      if (fwd) {
        /* Linear test models see the clamped state, as formod would. */
        double *xc = malloc(n * sizeof(double));
        orc_atm2x(ctl, atm_i, xc, NULL, NULL);
        fwd(user, xc, y_i, NULL);
        free(xc);
      } else {
        rc = orc_formod(ctl, atm_i, obs_i);
        if (rc) goto done;
        orc_obs2y(ctl, obs_i, y_i, NULL, NULL);
      }
      for (size_t i = 0; i < n; i++) dx[i] = x_i[i] - x_a[i];
      for (size_t i = 0; i < m; i++) dy[i] = y_m[i] - y_i[i];
      chisq = cost_function(dx, dy, s_a_inv, sig_eps_inv, n, m);
      if (chisq > chisq_old) {
        lmpar *= 10;
        for (size_t i = 0; i < n; i++) x_i[i] -= x_step[i];
      } else {
        lmpar /= 10;
        break;
      }
    }
    if (chisq_hist) chisq_hist[it] = chisq;
    disq = 0;
    for (size_t i = 0; i < n; i++) disq += x_step[i] * b[i];
    disq /= (double)n;
    if ((it == 1 || it % ret->kernel_recomp == 0) && disq < ret->conv_dmin) break;
  }
  if (it > ret->conv_itmax) it = ret->conv_itmax;
  if (ana) rc = oe_analysis(ret, ctl, obs_meas, k_i, s_a_inv, sig_eps_inv, iqa, m, n, ana);
done:
  if (niter) *niter = it;
  free(a); free(cov); free(k_i); free(s_a_inv); free(b); free(dx); free(dy);
  free(sig_eps_inv); free(x_a); free(x_i); free(x_step); free(y_aux); free(y_i);
  free(y_m); free(iqa); free(ipa);
  return rc;
}

/* tblgen (src/tblgen.c:44-140).  The column density of 1 km at 1 ppmv
   (:64) turns the optical depths into cross sections (:94, :99); the grid
   spacing is recomputed from the end points (:106), and the integration
   starts at the index (int) of the filter start's fractional position
   times npts (:107-108), skipping points below the filter and stopping
   above it (:117-120).  The filter bracket is kept from point to point and
   from one column density to the next, and searched only when the point
   leaves it (:122-123).  Column densities run 1, 1.122, 1.122^2, ... up to
   1e30 by repeated multiplication (:111); a row is kept when its emissivity
   lies in [1e-6, 0.999999] and exceeds the previous column density's
   (:129-134), and the loop ends after the first emissivity above 0.999999
   (:137-138). */
int orc_tblgen(double press, double temp, const double *tau, int npts, double nu0, double nu1,
               const double *nuf, const double *filt, int nf, double *u_out, double *eps_out,
               int maxrow) {
  if (nf < 2 || npts < 2) return -1;
  const double uref = 1e-6 * press * 100 / (1.380658e-23 * temp) * 1000 / 1e4;
  double *a = malloc(sizeof(double) * (size_t)npts);
  if (!a) return -1;
  for (int i = 0; i < npts; i++) a[i] = tau[i] / uref;
  const double dnu = (nu1 - nu0) / ((double)npts - 1.0);
  const double r0 = (nuf[0] - nu0) / (nu1 - nu0) * (double)npts;
  const int i0 = (int)r0;
  int idx = 0, nrow = 0;
  double epsold = 0;
  for (double u = 1.0; u <= 1e30; u *= 1.122) {
    double epssum = 0, fsum = 0;
    for (int i = i0; i < npts; i++) {
      const double nu = nu0 + dnu * (double)i;
      if (nu < nuf[0])
        continue;
      else if (nu > nuf[nf - 1])
        break;
      else {
        if (nu < nuf[idx] || nu > nuf[idx + 1]) idx = orc_locate_irr(nuf, nf, nu);
        const double f = LIN(nuf[idx], filt[idx], nuf[idx + 1], filt[idx + 1], nu);
        fsum += f;
        epssum += f * exp(-a[i] * u);
      }
    }
    epssum = 1 - epssum / fsum;
    if (epssum >= 1e-6 && epssum <= 0.999999 && epssum > epsold) {
      if (nrow < maxrow) {
        u_out[nrow] = u;
        eps_out[nrow] = epssum;
      }
      nrow++;
    }
    epsold = epssum;
    if (epssum > 0.999999) break;
  }
  free(a);
  return nrow;
}
