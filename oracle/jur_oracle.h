/*
  jur_oracle.h -- CPU FP64 restatement of JURASSIC's forward model path.

  TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
  forward model: only the test suite, the smoke check and the CPU-baseline
  leg of the benchmark load it.  The product library (libjurassic_gpu.so)
  never links, loads or calls it.

  Every function restates the reference algorithm (slcs-jsc/jurassic) with
  the same formulas, branch structure and operation order, and cites the
  reference file:line it follows.  Differences in data structures (tables
  sized to the files actually read, an explicit context instead of static
  caches bound to the first ctl) do not change results.

  Parity pins (see tests/): raytrace, intpol_atm, tangent_point against the
  tangent-point columns of the reference's tests/{limb,nadir}_test/rad.org;
  table lookups against brute-force scans and closed forms; the EGA chain
  against a linear-table closed form; radiance accumulation against the
  isothermal identity rad = B(T)(1 - tau); init_srcfunc against direct
  quadrature; kernel() against an analytic column on a linear table;
  optimal_estimation against the closed-form linear-Gaussian solution.
  Parity unpinned (no independent check exists without the reference's
  tables or GSL): the continuum coefficient values beyond the knot and
  reference-temperature identities tested, and GSL's Cholesky/BLAS
  rounding (a textbook Cholesky is used).
*/
#ifndef JUR_ORACLE_H
#define JUR_ORACLE_H

#include "../include/jurassic_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Line of sight of one ray (src/jurassic.h:678-731), sized by ctl->nd/ng. */
typedef struct {
  int np;
  int nd, ng;
  double *z, *lon, *lat, *p, *t, *ds, *sfeps;
  double *q;    /* [NLOS][ng] */
  double *k;    /* [NLOS][nd] */
  double *u, *cgp, *cgt, *cgu;   /* [NLOS][ng] */
  double *eps, *src;             /* [NLOS][nd] */
  double sft;
} orc_los_t;

orc_los_t *orc_los_alloc(int nd, int ng);
void orc_los_free(orc_los_t *los);

/* Geometry and interpolation (src/jurassic.c:108-119, 3500-3516, 3685-3706,
   4103-4170, 5223-5263). */
void orc_cart2geo(const double *x, double *z, double *lon, double *lat);
void orc_geo2cart(double z, double lon, double lat, double *x);
int orc_locate_irr(const double *xx, int n, double x);
int orc_locate_reg(const double *xx, int n, double x);
int orc_locate_tbl(const float *xx, int n, double x);
void orc_intpol_atm(const ctl_t *ctl, const atm_t *atm, double z, double *p,
                    double *t, double *q, double *k);
int orc_raytrace(const ctl_t *ctl, const atm_t *atm, obs_t *obs,
                 orc_los_t *los, int ir);
void orc_tangent_point(const orc_los_t *los, double *tpz, double *tplon,
                       double *tplat);
void orc_hydrostatic(const ctl_t *ctl, atm_t *atm);

/* Continua (src/jurassic.c:886-2918). */
double orc_ctmco2(double nu, double p, double t, double u);
double orc_ctmh2o(double nu, double p, double t, double q, double u);
double orc_ctmn2(double nu, double p, double t);
double orc_ctmo2(double nu, double p, double t);

/* Planck function and brightness temperature (src/jurassic.h:126-127,183-184). */
double orc_planck(double t, double nu);
double orc_bright(double rad, double nu);

/* Solar zenith angle (src/jurassic.c:5183-5219). */
double orc_sza(double sec, double lon, double lat);

/* Single emissivity-table column lookups on raw arrays
   (src/jurassic.c:3895-3967); exposed for unit tests. */
double orc_intpol_col_eps(const float *u, const float *eps, int nu, double x);
double orc_intpol_col_u(const float *u, const float *eps, int nu, double x);

/* Forward model (src/jurassic.c:3026-3071).  Returns 0 or a JUR_ERR_* code
   (the reference exits instead). */
int orc_formod(const ctl_t *ctl, atm_t *atm, obs_t *obs);
/* formod_fov alone on obs (src/jurassic.c:3121-3192), for its closed-form
   pins; obs->rad and obs->tau are the pencil-beam values. */
int orc_formod_fov(const ctl_t *ctl, obs_t *obs);

/* The same over n independent packages, OpenMP across packages after one
   warm call (the CPU baseline). */
int orc_formod_batch(const ctl_t *ctl, atm_t *atm, obs_t *obs, int n,
                     int nthreads);

/* State/measurement vector mapping (src/jurassic.c:29-104, 4174-4197,
   5896-5976, 2921-3009). */
size_t orc_atm2x(const ctl_t *ctl, const atm_t *atm, double *x, int *iqa,
                 int *ipa);
void orc_x2atm(const ctl_t *ctl, const double *x, atm_t *atm);
size_t orc_obs2y(const ctl_t *ctl, const obs_t *obs, double *y, int *ida,
                 int *ira);
void orc_y2obs(const ctl_t *ctl, const double *y, obs_t *obs);
void orc_copy_atm(const ctl_t *ctl, atm_t *dst, const atm_t *src, int init);
void orc_copy_obs(const ctl_t *ctl, obs_t *dst, const obs_t *src, int init);

/* Finite-difference Jacobian (src/jurassic.c:4004-4099); k is m x n
   row-major.  nthreads <= 0 uses the OpenMP default. */
int orc_kernel(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m,
               size_t n, int nthreads);

/* Central differences with the same step sizes (test reference). */
int orc_kernel_central(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                       size_t m, size_t n, int nthreads);

/* One-sided (central = 0) or central differences with the kernel() step
   sizes times scale: small steps approach the derivative that an adjoint
   computes (test reference). */
int orc_kernel_scaled(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                      size_t m, size_t n, int nthreads, int central,
                      double scale);

/* Retrieval settings (subset of ret_t, src/retrieval.c:31-113). */
typedef struct {
  int kernel_recomp;
  int conv_itmax;
  double conv_dmin;
  double err_formod[ND];
  double err_noise[ND];
  double err_press, err_press_cz, err_press_ch;
  double err_temp, err_temp_cz, err_temp_ch;
  double err_q[NG], err_q_cz[NG], err_q_ch[NG];
  double err_k[NW], err_k_cz[NW], err_k_ch[NW];
  double err_clz, err_cldz, err_clk[NCL];
  double err_sfz, err_sfp, err_sft, err_sfeps[NSF];
} orc_ret_t;

/* Levenberg-Marquardt optimal estimation (src/retrieval.c:448-666 without
   file output and error analysis).  On return atm_i/obs_i hold the final
   state; chisq_hist[0..*niter] the normalised cost per outer iteration.
   If fwd is non-NULL it replaces formod/kernel (for tests on a linear
   model): fwd(user, x, y, K) evaluates y = F(x) and, if K != NULL, K. */
typedef void (*orc_fwd_fn)(void *user, const double *x, double *y,
                           double *k);
int orc_optimal_estimation(const orc_ret_t *ret, const ctl_t *ctl,
                           obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                           atm_t *atm_i, double *chisq_hist, int *niter,
                           orc_fwd_fn fwd, void *user);

/* Error analysis of optimal_estimation (src/retrieval.c:668-726) at the
   last Jacobian: cov, corr, avk n x n; gain n x m; err_total, err_noise,
   err_formod, cont, res n (standard deviations of write_stddev, the
   analyze_avk area and 1 / diagonal per state element).  NULL pointers are
   skipped. */
typedef struct {
  double *cov, *corr, *gain, *avk;
  double *err_total, *err_noise, *err_formod;
  double *cont, *res;
} orc_ana_t;

/* orc_optimal_estimation followed by the error analysis into *ana. */
int orc_optimal_estimation_ana(const orc_ret_t *ret, const ctl_t *ctl,
                               obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                               atm_t *atm_i, double *chisq_hist, int *niter,
                               orc_fwd_fn fwd, void *user,
                               const orc_ana_t *ana);

/* A priori covariance (src/retrieval.c:808-920), n x n row-major. */
int orc_set_cov_apr(const orc_ret_t *ret, const ctl_t *ctl, const atm_t *atm,
                    const int *iqa, const int *ipa, size_t n, double *s_a);

/* Source-function table of the current context: TBLNS x nd (row-major by
   temperature) and the temperature grid (src/jurassic.c:3631-3681). */
int orc_srcfunc_table(const ctl_t *ctl, double *st, double *sr);

/* Ray trace one ray and copy its LOS out (test helper).  Arrays must hold
   NLOS points; q/u are [NLOS][ng].  Returns np or -error. */
int orc_trace_ray(const ctl_t *ctl, const atm_t *atm, obs_t *obs, int ir,
                  double *z, double *p, double *t, double *ds, double *u,
                  double *lon, double *lat, double *sft);

/* tblgen (src/tblgen.c:44-140): emissivity rows of one pressure and
   temperature from an optical-depth spectrum tau[npts] on the regular grid
   nu0..nu1 and a filter function (nuf, filt)[nf].  Writes the rows the
   reference prints, (u, eps), up to maxrow of them, and returns their
   number (-1 for an empty filter range). */
int orc_tblgen(double press, double temp, const double *tau, int npts, double nu0, double nu1,
               const double *nuf, const double *filt, int nf, double *u_out, double *eps_out,
               int maxrow);

/* Drop cached tables (next call re-reads them). */
void orc_reset(void);

const char *orc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
