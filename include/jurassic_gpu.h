/*
  jurassic_gpu.h -- C-ABI of the MI355X forward model (libjurassic_gpu.so).

  The three structs below are the caller-side data model of JURASSIC's
  forward model, laid out field for field as in the reference header
  (src/jurassic.h:488-538 atm_t, :541-675 ctl_t, :734-775 obs_t) with the
  same compile-time capacities (src/jurassic.h:346-444).  A C caller that
  already fills these structs for formod() can pass the same objects to
  formod_gpu().  The capacities must agree between caller and library;
  jur_gpu_abi_check() reports the sizes the library was built with.

  Emissivity tables and filter functions are read from the files named by
  ctl->tblbase exactly as the reference does (src/jurassic.c:4949-5110,
  4907-4946); they are kept resident in HBM in a device context that is
  keyed on the ctl content (no process-wide statics bound to the first ctl).
*/
#ifndef JURASSIC_GPU_H
#define JURASSIC_GPU_H

#include <stddef.h>
#include <stdint.h>

/* ------------------------------------------------------------------
   Capacities (same names and defaults as src/jurassic.h:346-444).
   ------------------------------------------------------------------ */
#ifndef NCL
#define NCL 8
#endif
#ifndef ND
#define ND 128
#endif
#ifndef NG
#define NG 8
#endif
#ifndef NP
#define NP 256
#endif
#ifndef NR
#define NR 256
#endif
#ifndef NSF
#define NSF 8
#endif
#ifndef NW
#define NW 4
#endif
#ifndef LEN
#define LEN 10000
#endif
#ifndef NLOS
#define NLOS 4096
#endif
#ifndef NSHAPE
#define NSHAPE 20000
#endif
#ifndef NFOV
#define NFOV 5
#endif
#ifndef TBLNP
#define TBLNP 41
#endif
#ifndef TBLNT
#define TBLNT 30
#endif
#ifndef TBLNU
#define TBLNU 320
#endif
#ifndef TBLNS
#define TBLNS 1200
#endif

/* ------------------------------------------------------------------
   Data model (field order and types as src/jurassic.h:488-775).
   A caller that already includes the reference's jurassic.h (include guard
   JURASSIC_H) uses its own definitions of these types, which are the same.
   ------------------------------------------------------------------ */
#ifndef JURASSIC_H

/* Atmospheric data (src/jurassic.h:488-538). */
typedef struct {
  int np;                 /* number of levels */
  double time[NP];        /* s since 2000-01-01T00:00Z */
  double z[NP];           /* altitude [km] */
  double lon[NP];         /* [deg] */
  double lat[NP];         /* [deg] */
  double p[NP];           /* pressure [hPa] */
  double t[NP];           /* temperature [K] */
  double q[NG][NP];       /* volume mixing ratio [ppv] */
  double k[NW][NP];       /* extinction [km^-1] */
  double clz;             /* cloud layer height [km] */
  double cldz;            /* cloud layer depth [km] */
  double clk[NCL];        /* cloud layer extinction [km^-1] */
  double sfz;             /* surface height [km] */
  double sfp;             /* surface pressure [hPa] */
  double sft;             /* surface temperature [K] */
  double sfeps[NSF];      /* surface emissivity */
} atm_t;

/* Forward model control parameters (src/jurassic.h:541-675). */
typedef struct {
  int ng;
  char emitter[NG][LEN];
  int nd;
  double nu[ND];
  int nw;
  int window[ND];
  int ncl;
  double clnu[NCL];
  int nsf;
  double sfnu[NSF];
  int sftype;             /* 0 none, 1 emission, 2 downward, 3 solar */
  double sfsza;
  char tblbase[LEN];
  int tblfmt;             /* 1 ASCII, 2 binary */
  double hydz;
  int ctm_co2;
  int ctm_h2o;
  int ctm_n2;
  int ctm_o2;
  int refrac;
  double rayds;
  double raydz;
  char fov[LEN];
  double retp_zmin;
  double retp_zmax;
  double rett_zmin;
  double rett_zmax;
  double retq_zmin[NG];
  double retq_zmax[NG];
  double retk_zmin[NW];
  double retk_zmax[NW];
  int ret_clz;
  int ret_cldz;
  int ret_clk;
  int ret_sfz;
  int ret_sfp;
  int ret_sft;
  int ret_sfeps;
  int write_bbt;
  int write_matrix;
  int formod;             /* 0 CGA, 1 EGA, 2 RFM */
  char rfmbin[LEN];
  char rfmhit[LEN];
  char rfmxsc[NG][LEN];
} ctl_t;

/* Observation geometry and radiance data (src/jurassic.h:734-775). */
typedef struct {
  int nr;
  double time[NR];
  double obsz[NR];
  double obslon[NR];
  double obslat[NR];
  double vpz[NR];
  double vplon[NR];
  double vplat[NR];
  double tpz[NR];
  double tplon[NR];
  double tplat[NR];
  double tau[ND][NR];
  double rad[ND][NR];
} obs_t;

#endif /* JURASSIC_H */

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------
   Status codes and options.
   ------------------------------------------------------------------ */
#define JUR_OK 0
#define JUR_ERR_ARG 1         /* invalid argument or capacity exceeded */
#define JUR_ERR_IO 2          /* a required input file cannot be read */
#define JUR_ERR_LOS 3         /* "Too many LOS points!" */
#define JUR_ERR_OBSERVER 4    /* "Observer below surface!" */
#define JUR_ERR_HIP 5         /* HIP runtime failure */
#define JUR_ERR_UNSUPPORTED 6 /* e.g. FORMOD = 2 (RFM) */
#define JUR_ERR_MEM 7         /* device allocation failed */
#define JUR_ERR_FOV 8         /* "Cannot apply FOV convolution!" */

#define JUR_FP64 0            /* all arithmetic in FP64, reference order */
#define JUR_FP32 1            /* FP32 lookups, FP64 geometry */

#define JUR_JAC_FD 0          /* one-sided finite differences (kernel()) */
#define JUR_JAC_ADJOINT 1     /* reverse sweep, path held fixed */

/* State map of the device Jacobian: for one atmosphere, the column of K that
   each state element occupies, -1 if it is not part of the state.  Level
   quantities iq: 0 pressure, 1 temperature, 2+ig volume mixing ratio of
   emitter ig, 2+NG+iw extinction of window iw; then the scalars in the
   order of atm2x (src/jurassic.c:29-83). */
#define JUR_XMAP_NQ (2 + NG + NW)
#define JUR_XMAP_LEV(iq, ip) ((iq) * NP + (ip))
#define JUR_XMAP_CLZ (JUR_XMAP_NQ * NP)
#define JUR_XMAP_CLDZ (JUR_XMAP_CLZ + 1)
#define JUR_XMAP_CLK(icl) (JUR_XMAP_CLZ + 2 + (icl))
#define JUR_XMAP_SFZ (JUR_XMAP_CLZ + 2 + NCL)
#define JUR_XMAP_SFP (JUR_XMAP_SFZ + 1)
#define JUR_XMAP_SFT (JUR_XMAP_SFZ + 2)
#define JUR_XMAP_SFEPS(isf) (JUR_XMAP_SFZ + 3 + (isf))
#define JUR_XMAP_LEN JUR_XMAP_SFEPS(NSF)

/* Text of the last error of the calling thread ("" if none). */
const char *jur_gpu_last_error(void);

/* Size of each ABI struct as compiled into the library, for a caller-side
   check that both sides were built with the same capacities.
   Returns 0 when all three sizes match the arguments. */
int jur_gpu_abi_check(size_t sizeof_ctl, size_t sizeof_atm, size_t sizeof_obs);

/* ------------------------------------------------------------------
   Drop-in entry points.
   ------------------------------------------------------------------ */

/* Replaces formod() (src/jurassic.h:890-893, src/jurassic.c:3026-3071).
   Same inputs, outputs and error behaviour: on error it prints
   "Error (file, func, line): msg" and calls exit(EXIT_FAILURE), as ERRMSG
   does (src/jurassic.h:237-241).  Arithmetic is FP64. */
void formod_gpu(const ctl_t *ctl, atm_t *atm, obs_t *obs);

/* Status-returning form of formod_gpu with a precision selector
   (JUR_FP64 or JUR_FP32).  Returns JUR_OK or an error code; on error obs
   is left unchanged. */
int jur_gpu_formod(const ctl_t *ctl, atm_t *atm, obs_t *obs, int precision);

/* n independent (atm[i], obs[i]) packages with one ctl in one device pass,
   the shape of the external GPU library's
   jur_unified_formod_multiple_packages (src/formod.c:72-73). */
int jur_gpu_formod_batch(const ctl_t *ctl, atm_t *atm, obs_t *obs, int n,
                         int precision);

/* Replaces kernel() (src/jurassic.h:1008-1012, src/jurassic.c:4004-4099):
   k is the m x n Jacobian, row-major with row stride n (a gsl_matrix with
   tda == size2).  m and n must equal the sizes of obs2y/atm2x.
   Finite differences use the reference step sizes and one formod per
   column, all columns in one device pass. */
int jur_gpu_kernel(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k,
                   size_t m, size_t n, int method, int precision);

/* Replaces write_matrix (src/jurassic.c:5492-5666): writes matrix (row-major,
   rows and columns in measurement space "y" = obs2y order or state space
   "x" = atm2x order, as rowspace/colspace say) to <dirname>/<filename> in
   the reference's text format (12 metadata columns, then the element;
   sort "r" row by row, else column by column) when ctl->write_matrix is
   set.  Returns JUR_OK, JUR_ERR_ARG or JUR_ERR_IO ("Cannot create file!"). */
int jur_write_matrix(const char *dirname, const char *filename, const ctl_t *ctl, const double *matrix,
                     const atm_t *atm, const obs_t *obs, const char *rowspace, const char *colspace,
                     const char *sort);

/* The state map of atm2x (src/jurassic.c:29-83) for ctl and atm: fills
   xmap[JUR_XMAP_LEN] and returns the state size n. */
int jur_state_map(const ctl_t *ctl, const atm_t *atm, int *xmap);

/* Emissivity rows of a lookup table from a monochromatic spectrum: replaces
   the tblgen tool (src/tblgen.c:44-140, main(<press> <temp> <spec>
   <filter>)).  tau[npts] is the optical-depth spectrum on the regular grid
   nu0..nu1 (the values of the RFM file that tblgen reads, :84-102), and
   (nuf, filt)[nf] the filter function (:67-74).  The sums over the spectrum
   for the ~600 column densities 1.122^k run on the GPU.  Writes the rows
   tblgen prints, u_out[r] and eps_out[r] (its third and fourth columns), up
   to maxrow of them, and their number to *nrow (which may exceed maxrow).
   Returns JUR_OK or JUR_ERR_ARG / JUR_ERR_MEM / JUR_ERR_HIP. */
int jur_tblgen(double press, double temp, const double *tau, int npts, double nu0, double nu1,
               const double *nuf, const double *filt, int nf, double *u_out, double *eps_out, int maxrow,
               int *nrow);

/* jur_tblgen for nspec spectra on one grid and filter, e.g. the pressures
   and temperatures of one table (the reference runs tblgen once per (p, T),
   src/tblgen.c:44-140): tau is [nspec][npts]; spectrum s writes its rows
   to u_out/eps_out[s * maxrow + r] and their number to nrow[s]. */
int jur_tblgen_batch(int nspec, const double *press, const double *temp, const double *tau, int npts,
                     double nu0, double nu1, const double *nuf, const double *filt, int nf, double *u_out,
                     double *eps_out, int maxrow, int *nrow);

/* Retrieval settings: the part of the reference's ret_t
   (src/retrieval.c:31-113) that optimal_estimation reads. */
typedef struct {
  int kernel_recomp;      /* KERNEL_RECOMP */
  int conv_itmax;         /* CONV_ITMAX */
  double conv_dmin;       /* CONV_DMIN */
  double err_formod[ND];  /* ERR_FORMOD [%] */
  double err_noise[ND];   /* ERR_NOISE */
  double err_press, err_press_cz, err_press_ch;
  double err_temp, err_temp_cz, err_temp_ch;
  double err_q[NG], err_q_cz[NG], err_q_ch[NG];
  double err_k[NW], err_k_cz[NW], err_k_ch[NW];
  double err_clz, err_cldz, err_clk[NCL];
  double err_sfz, err_sfp, err_sft, err_sfeps[NSF];
} jur_ret_t;

/* Replaces optimal_estimation (src/retrieval.c:448-666; the file output and
   error analysis are the caller's): n independent Levenberg-Marquardt
   retrievals in lock-step, each one exactly the reference iteration, with
   the model evaluations of all retrievals in shared device passes and the
   normal equations (K^T S_eps^-1 K + S_a^-1 (1 + lambda)) solved on the
   GPU by Cholesky.  For retrieval i: obs_meas[i], atm_apr[i] in;
   atm_i[i], obs_i[i] out (the last evaluated state, as the reference),
   chisq_hist[i * (conv_itmax + 1) + it] the cost of iteration it, niter[i].
   method: JUR_JAC_FD (kernel(), one-sided differences) or JUR_JAC_ADJOINT.
   Returns JUR_OK or an error code. */
int jur_gpu_optimal_estimation(const jur_ret_t *ret, const ctl_t *ctl, int n,
                               obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                               atm_t *atm_i, double *chisq_hist, int *niter,
                               int method, int precision);

/* Error analysis of a retrieval (optimal_estimation, src/retrieval.c:
   668-726, with ERR_ANA = 1), at the last Jacobian of the iteration; n is
   the state size (atm2x) and m the number of finite measurements (obs2y).
   Every pointer may be NULL (that output is not produced):
     cov   n x n retrieval covariance (K^T S_eps^-1 K + S_a^-1)^-1
           (matrix_cov_ret.tab)
     corr  n x n correlation matrix (matrix_corr.tab)
     gain  n x m gain matrix cov K^T S_eps^-1 (matrix_gain.tab)
     avk   n x n averaging kernels gain K (matrix_avk.tab)
     err_total, err_noise, err_formod  n standard deviations: sqrt of the
           diagonals of cov, G S_noise G^T, G S_formod G^T (write_stddev)
     cont, res  n: per state element the sum of its averaging-kernel row
           over its own quantity and 1 / its diagonal (analyze_avk)
     atm_*  the same as atmospheres, as the reference writes them to
           atm_err_{total,noise,formod}.tab, atm_cont.tab and atm_res.tab
           (copy_atm with init = 1 of the final state, then the values at
           the state elements). */
typedef struct {
  double *cov, *corr, *gain, *avk;
  double *err_total, *err_noise, *err_formod;
  double *cont, *res;
  atm_t *atm_err_total, *atm_err_noise, *atm_err_formod;
  atm_t *atm_cont, *atm_res;
} jur_ana_t;

/* jur_gpu_optimal_estimation with the error analysis of retrieval i into
   ana[i] (ana may be NULL: no analysis). */
int jur_gpu_optimal_estimation_ana(const jur_ret_t *ret, const ctl_t *ctl,
                                   int n, obs_t *obs_meas, obs_t *obs_i,
                                   atm_t *atm_apr, atm_t *atm_i,
                                   double *chisq_hist, int *niter,
                                   int method, int precision,
                                   const jur_ana_t *ana);

/* ------------------------------------------------------------------
   Request server.  Forward models, Jacobians and retrievals are submitted
   at any time from any thread; worker lanes (threads with their own HIP
   stream) batch everything in flight into shared device passes (continuous
   batching: requests that arrive while a pass runs join the next round of
   the first lane to start one, admitted in arrival order while their
   device memory fits the budget).  No counterpart in the
   reference, whose programs run one job at a time.
   ------------------------------------------------------------------ */
#define JUR_REQ_FORWARD 0     /* formod */
#define JUR_REQ_JACOBIAN 1    /* formod and kernel */
#define JUR_REQ_RETRIEVAL 2   /* optimal_estimation */
#define JUR_ERR_TIMEOUT 9     /* jur_server_wait: not finished in time */

typedef struct jur_server jur_server_t;

typedef struct {
  int method;             /* JUR_JAC_ADJOINT or JUR_JAC_FD */
  int precision;          /* JUR_FP32 or JUR_FP64 */
  size_t max_bytes;       /* device memory for requests in flight (0: 16 GiB) */
  int max_batch;          /* most requests in flight (0: no limit) */
  int lanes;              /* worker lanes, each with its own stream (0: 1) */
  int forward_lane;       /* 1: one more lane that takes the forward
                             requests (forward pass, no Jacobian-mode
                             sweep); the other lanes then take Jacobians
                             and retrievals */
  int path_split;         /* 1: one more lane that takes the requests with
                             no limb ray (nadir and slant views that end on
                             the ground: a few hundred LOS points); the
                             other lanes take those with limb rays (up to
                             some 2000 points), so that the short requests'
                             rounds do not last as long as the longest
                             limb ray's sweeps.  Not with forward_lane. */
} jur_server_opts_t;

typedef struct {
  int kind;               /* JUR_REQ_* */
  const atm_t *atm;       /* the state (retrieval: the a priori) */
  const obs_t *obs;       /* geometry; retrieval: also the measurement.
                             Non-finite rad entries are masked as in formod */
  /* state windows of Jacobian and retrieval requests (RET*_ZMIN/ZMAX,
     src/jurassic.c:4622-4633): the state vector is that of atm2x */
  double retp_zmin, retp_zmax, rett_zmin, rett_zmax;
  double retq_zmin[NG], retq_zmax[NG];
  double retk_zmin[NW], retk_zmax[NW];
} jur_request_t;

typedef struct {
  uint64_t id;
  int kind;
  int status;             /* JUR_OK or the request's error code */
  char message[256];
  double t_submit, t_start, t_done;  /* seconds on a steady clock */
  int n, m;               /* state and measurement sizes */
  int niter;              /* retrieval iterations */
} jur_result_t;

/* A server for one forward model (ctl) and one set of retrieval settings. */
int jur_server_create(const ctl_t *ctl, const jur_ret_t *ret,
                      const jur_server_opts_t *opt, jur_server_t **srv);
void jur_server_destroy(jur_server_t *srv);
/* Copies the request's inputs and queues it; *id names it.  A request needs
   1 <= obs->nr <= NR rays (JUR_ERR_ARG otherwise). */
int jur_server_submit(jur_server_t *srv, const jur_request_t *req,
                      uint64_t *id);
/* Waits until request id is finished (timeout_s < 0: no limit). */
int jur_server_wait(jur_server_t *srv, uint64_t id, double timeout_s,
                    jur_result_t *res);
/* Copies a finished request's outputs and releases it: obs (the model at
   the final state), atm (retrieval: the final state), k (Jacobian: m x n
   row-major, rows of obs2y), chisq_hist (retrieval: niter + 1 costs).
   Any output may be NULL.  Returns the request's status. */
int jur_server_fetch(jur_server_t *srv, uint64_t id, obs_t *obs, atm_t *atm,
                     double *k, double *chisq_hist);
/* Device passes run so far and the most requests in flight at once. */
int jur_server_stats(const jur_server_t *srv, long *rounds, long *max_active);
/* Worker time in seconds, summed over lanes: [0] admission, [1] device
   passes (with the adjoint, every pass), [2] forward-only passes (the
   finite-difference method), [3] dense algebra and bookkeeping; then,
   counted over the whole process since it started, [4] seconds spent in the
   library's device and pinned-host allocations, [5] how many there were and
   [6] the seconds of those that first freed a smaller buffer (a free waits
   for the whole device) (n <= 7). */
int jur_server_profile(const jur_server_t *srv, double *sec, int n);

/* ERRMSG-style wrapper of jur_gpu_kernel(..., JUR_JAC_FD, JUR_FP64). */
void kernel_gpu(ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m,
                size_t n);

/* ------------------------------------------------------------------
   Device-resident batch interface (inputs and outputs in HBM).
   ------------------------------------------------------------------ */
typedef struct jur_ctx jur_ctx_t;

/* Load tables and filter functions for ctl and build the device context. */
int jur_ctx_create(const ctl_t *ctl, jur_ctx_t **ctx);

/* As jur_ctx_create, with a channel list of any length (nd may exceed ND).
   window[] may be NULL (all channels in window 0). */
int jur_ctx_create_wide(const ctl_t *ctl, int nd, const double *nu,
                        const int *window, jur_ctx_t **ctx);

/* Emissivity tables already in device memory: the arrays of the
   reference's tbl_t (src/jurassic.h:778-807), one table per (channel,
   emitter), channel-major as there, with capacities tp, tt, tu:
     np[id][ig]; p[id][ig][tp]; nt[id][ig][tp]; t[id][ig][tp][tt];
     nu[id][ig][tp][tt]; u, eps[id][ig][tp][tt][tu]
   np < 0 marks a missing table (transparent, as read_tbl's warning,
   src/jurassic.c:4980-4983). Values are taken as they are, as read_tbl
   takes a binary table (src/jurassic.c:5049-5083).
   Filter functions: filt_off == NULL reads <tblbase>_<nu>.filt as
   read_shape does; otherwise channel id has the filt_off[id+1] -
   filt_off[id] points filt_nu[filt_off[id] ...], filt_w[...] (host). */
typedef struct {
  int tp, tt, tu;
  const int *np, *nt, *nu;
  const double *p, *t;
  const float *u, *eps;
  const int *filt_off;
  const double *filt_nu, *filt_w;
} jur_dtables_t;

/* As jur_ctx_create_wide, with the tables taken from device memory instead
   of read_tbl's files (src/jurassic.c:4949-5110). The arrays are read during
   the call only. */
int jur_ctx_create_device_tables(const ctl_t *ctl, int nd, const double *nu,
                                 const int *window, const jur_dtables_t *tab,
                                 jur_ctx_t **ctx);

void jur_ctx_destroy(jur_ctx_t *ctx);

/* Number of channels of the context. */
int jur_ctx_nd(const jur_ctx_t *ctx);

/* One forward pass over device-resident rays.
   Rays are processed in blocks of 64 consecutive rays; ordering rays so
   that each block holds similar geometries (e.g. one view angle or tangent
   height over many profiles) lets the lanes of a wavefront share table
   cache lines. */
typedef struct {
  int nprof;                  /* number of atmospheres */
  atm_t *atm;                 /* device: nprof records (p updated if HYDZ>=0) */
  int nray;                   /* number of rays */
  const int *ray_prof;        /* device: atmosphere index of each ray */
  const double *ray_geom;     /* device: [nray][7] time obsz obslon obslat vpz vplon vplat */
  const unsigned char *mask;  /* device or NULL: [nray][nd], nonzero -> NaN radiance */
  double *rad;                /* device out: [nray][nd] radiance or BT */
  double *tau;                /* device out: [nray][nd] transmittance */
  double *tp;                 /* device out or NULL: [nray][3] tpz tplon tplat */
} jur_batch_t;

int jur_ctx_formod_device(jur_ctx_t *ctx, const jur_batch_t *batch,
                          int precision, void *hip_stream);

/* Jacobian of a device batch by the adjoint (path held fixed).
   K row (ray, channel) holds d rad / d x (d BT / d x with WRITE_BBT) over the
   ncol columns of the ray's state map; masked radiances give NaN rows.
   The batch's rad/tau receive the forward model of the same pass.
   Surface altitude and pressure only move the path: their map entries must
   be -1 (jur_gpu_kernel covers them by finite differences). */
typedef struct {
  int nmap;                   /* number of state maps */
  const int *xmap;            /* host: [nmap][JUR_XMAP_LEN] */
  const int *prof_map;        /* device: [nprof] map of each atmosphere (<0: none) */
  int ncol;                   /* row length of K */
  double *k;                  /* device out: [nray][nd][ncol] */
} jur_jac_t;

int jur_ctx_jacobian_device(jur_ctx_t *ctx, const jur_batch_t *batch,
                            const jur_jac_t *jac, int precision,
                            void *hip_stream);

/* Kernel timing of the last jur_ctx_formod_device call on this context, in
   milliseconds, measured with HIP events on the stream the kernels ran on:
   [0] ray tracing, [1] emissivity-table lookups, [2] radiance accumulation
   (continua, source function, surface terms), or for
   jur_ctx_jacobian_device the reverse sweep and the K store.
   n <= 3 entries are written. */
int jur_ctx_last_timing(const jur_ctx_t *ctx, double *ms, int n);

/* Mean number of LOS points per ray of the last pass. */
double jur_ctx_last_mean_np(const jur_ctx_t *ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): jur_ctx_formod_device with every work buffer (line of
   sight, emissivities, per-ray state) taken from a fixed device array of
   JUR_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the
   process ends; the call itself allocates no device memory.  Same inputs,
   outputs and results as jur_ctx_formod_device.  Large batches are processed
   in chunks of ray blocks; a batch whose per-ray buffers, or one block of
   64 rays, do not fit the array is rejected with JUR_ERR_MEM and a message
   (few channels and emitters fit: e.g. 12 channels x 3 emitters, limb).
   The context must belong to the current device.  Calls that overlap on one
   GPU (other threads or streams) share the array and race. */
#define JUR_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int jur_ctx_formod_device_noalloc(jur_ctx_t *ctx, const jur_batch_t *batch,
                                  int precision, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif
