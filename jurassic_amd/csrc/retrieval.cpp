// Batched model evaluations, Jacobians and Levenberg-Marquardt optimal
// estimation (optimal_estimation, src/retrieval.c:448-666, without the file
// output and the error analysis), and the request server built on them.
//
// The engine holds a set of jobs.  A job is one request: a forward model, a
// Jacobian, or a retrieval.  Every retrieval follows the reference loop
// exactly (a priori covariance set_cov_apr :808-920, measurement errors
// set_cov_meas :924-964, cost :340-369, the damped normal equations solved by
// Cholesky, the state clamps and the convergence test).  The engine advances
// all jobs in rounds: each round runs one device pass over the rays of every
// job that needs a forward model (and one pass for those that need the
// Jacobian as well), then the dense algebra of all retrievals at the same
// stage, one launch per kernel (lm_kernels.hip).  Jobs of any size and
// progress share those passes, and jobs can join or leave between rounds,
// which is what the server does: requests that arrive while a round runs
// join the next one (continuous batching), admitted while their device
// memory fits the budget.
#include "jur_host.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace jur {


void launch_lm_gather(const LmDesc *, const int *, int, int, const double *, int, int, hipStream_t);
void launch_lm_normal(const LmDesc *, const int *, int, int, hipStream_t);
void launch_lm_cholinv(const LmDesc *, const int *, int, int *, hipStream_t);
void launch_lm_rhs(const LmDesc *, const int *, int, int, hipStream_t);
void launch_lm_solve(const LmDesc *, const int *, int, const double *, int *, hipStream_t);
void launch_lm_quad(const LmDesc *, const int *, int, double *, hipStream_t);
void launch_lm_stage(const LmDesc *, const int *, int, const size_t *, double *, int, hipStream_t);
void launch_lm_cov_apr(const LmDesc *, const int *, int, const size_t *, const double *, hipStream_t);
void launch_ana(const LmDesc *, const int *, int, int *, hipStream_t, int);

// Device memory for the jobs' matrices and vectors: blocks of 2^k doubles
// kept on free lists and reused, so that admitting a request allocates
// nothing once the pool has warmed up.
struct DevPool {
  std::vector<std::vector<double *>> free_;
  std::vector<double *> all_;
  DevPool() : free_(48) {}
  ~DevPool() {
    for (double *p : all_) (void)hipFree(p);
  }
  static int cls(size_t n) {
    int k = 6;
    while (((size_t)1 << k) < n) k++;
    return k;
  }
  double *get(size_t n, int &c) {
    c = cls(n);
    if (!free_[c].empty()) {
      double *p = free_[c].back();
      free_[c].pop_back();
      return p;
    }
    double *p = nullptr;
    AllocTimer t;
    if (hipMalloc((void **)&p, ((size_t)1 << c) * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    all_.push_back(p);
    return p;
  }
  void put(double *p, int c) {
    if (p) free_[c].push_back(p);
  }
};
int kernel_fd_host(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n, int precision);

#define POW2(x) ((x) * (x))

namespace {

// ST_EVAL_FT: a trial step evaluated with its Jacobian, because the next
// iteration recomputes K (kernel_recomp) and, if the step is accepted, that
// K is the one at the new state.
enum Stage { ST_EVAL_F_OUT, ST_EVAL_FK_OUT, ST_EVAL_FK0, ST_ITER, ST_EVAL_FK, ST_TRIAL, ST_EVAL_F, ST_EVAL_FT, ST_DONE };

struct Job {
  int kind = JUR_REQ_RETRIEVAL;
  ctl_t *sctl = nullptr;      // state windows of this job (ctl fields of atm2x)
  atm_t *atm_apr = nullptr;   // retrieval: a priori; otherwise: the state
  atm_t *atm_i = nullptr;     // retrieval: current state
  obs_t *obs_meas = nullptr;  // geometry (and measurement)
  obs_t *obs_i = nullptr;     // last model evaluation
  double *hist = nullptr;
  int *niter = nullptr;
  double *kout = nullptr;     // Jacobian requests: K (mv x n)
  int n = 0, m = 0, mv = 0, nr = 0;
  std::vector<int> iqa, ipa, xmap;
  std::vector<double> x_a, x_i, y_m, sig, dx, dy, b, step;
  double chisq = 0, chisq_old = 0, lmpar = 0.001;
  int it = 0, it2 = 0;
  bool fk_done = false;
  Stage st = ST_EVAL_FK0;
  int rc = JUR_OK;
  char err[256] = "";
  // device buffers of a retrieval (from the engine's pool):
  // S_a^-1, C, A, K, sig, dy, dx, b, step
  double *dp[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int dc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double *ktr = nullptr;  // K of a trial step (ST_EVAL_FT), m x n
  int ktrc = 0;
  std::vector<double> sa_in;  // packed inputs of S_a (8 n)
  int ray0 = 0;
  bool limb = false;       // a ray of the job crosses a tangent point
  size_t bytes = 0;        // device footprint of the job's own buffers
  size_t admit_bytes = 0;  // what the server reserved for it
  // server bookkeeping
  uint64_t id = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  std::unique_ptr<ctl_t> own_ctl;
  std::unique_ptr<atm_t> own_apr, own_atm_i;
  std::unique_ptr<obs_t> own_meas, own_obs_i;
  std::vector<double> own_hist, own_k;
  int own_niter = 0;

  const jur_ana_t *ana = nullptr;  // error analysis outputs (retrievals)
  double *g = nullptr;             // gain matrix during the analysis
  int gc = 0;
  LmDesc desc() {
    return LmDesc{n, m, ray0, nr, dp[0], dp[1], dp[2], st == ST_EVAL_FT ? ktr : dp[3], dp[4], dp[5], dp[6], dp[7],
                  dp[8], g};
  }
  bool failed() const { return rc != JUR_OK; }
};

// The inputs of set_cov_apr (src/retrieval.c:808-920) per state element,
// packed for k_lm_cov_apr: standard deviation, altitude, geo2cart(0, lon,
// lat), correlation lengths in z and horizontally, quantity index.
int cov_apr_inputs(const jur_ret_t *ret, const ctl_t *ctl, const atm_t *atm, Job &J) {
  const size_t n = J.n;
  std::vector<double> xa(n + 1);
  atm2x_h(ctl, atm, xa.data(), nullptr, nullptr);
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl, nsf = ctl->nsf;
  J.sa_in.assign(8 * n, 0.0);
  double *sg = J.sa_in.data(), *z = sg + n, *px = sg + 2 * n, *py = sg + 3 * n, *pz = sg + 4 * n;
  double *czv = sg + 5 * n, *chv = sg + 6 * n, *qv = sg + 7 * n;
  for (size_t i = 0; i < n; i++) {
    const int q = J.iqa[i];
    double cz = 0, ch = 0;
    if (q == IDXP) { xa[i] = ret->err_press / 100 * xa[i]; cz = ret->err_press_cz; ch = ret->err_press_ch; }
    if (q == IDXT) { xa[i] = ret->err_temp; cz = ret->err_temp_cz; ch = ret->err_temp_ch; }
    for (int ig = 0; ig < ng; ig++)
      if (q == 2 + ig) { xa[i] = ret->err_q[ig] / 100 * xa[i]; cz = ret->err_q_cz[ig]; ch = ret->err_q_ch[ig]; }
    for (int iw = 0; iw < nw; iw++)
      if (q == 2 + ng + iw) { xa[i] = ret->err_k[iw]; cz = ret->err_k_cz[iw]; ch = ret->err_k_ch[iw]; }
    if (q == 2 + ng + nw) xa[i] = ret->err_clz;
    if (q == 3 + ng + nw) xa[i] = ret->err_cldz;
    for (int icl = 0; icl < ncl; icl++)
      if (q == 4 + ng + nw + icl) xa[i] = ret->err_clk[icl];
    if (q == 4 + ng + nw + ncl) xa[i] = ret->err_sfz;
    if (q == 5 + ng + nw + ncl) xa[i] = ret->err_sfp;
    if (q == 6 + ng + nw + ncl) xa[i] = ret->err_sft;
    for (int isf = 0; isf < nsf; isf++)
      if (q == 7 + ng + nw + ncl + isf) xa[i] = ret->err_sfeps[isf];
    if (POW2(xa[i]) <= 0) return set_err(JUR_ERR_ARG, "Check a priori data (zero standard deviation)!");
    const int ip = J.ipa[i];
    const double latr = atm->lat[ip] / 180. * JUR_PI, lonr = atm->lon[ip] / 180. * JUR_PI;
    sg[i] = xa[i];
    z[i] = atm->z[ip];
    px[i] = JUR_RE * cos(latr) * cos(lonr);
    py[i] = JUR_RE * cos(latr) * sin(lonr);
    pz[i] = JUR_RE * sin(latr);
    czv[i] = cz;
    chv[i] = ch;
    qv[i] = q;
  }
  return JUR_OK;
}

// The rows of an observation that a job uses: geometry of its nr rays and
// radiance/transmittance of nd channels (an obs_t holds NR x ND of each).
void copy_obs_rows(obs_t *dst, const obs_t *src, int nd) {
  const int nr = src->nr;
  dst->nr = nr;
  const size_t b = (size_t)nr * sizeof(double);
  memcpy(dst->time, src->time, b);
  memcpy(dst->obsz, src->obsz, b);
  memcpy(dst->obslon, src->obslon, b);
  memcpy(dst->obslat, src->obslat, b);
  memcpy(dst->vpz, src->vpz, b);
  memcpy(dst->vplon, src->vplon, b);
  memcpy(dst->vplat, src->vplat, b);
  memcpy(dst->tpz, src->tpz, b);
  memcpy(dst->tplon, src->tplon, b);
  memcpy(dst->tplat, src->tplat, b);
  for (int id = 0; id < nd; id++) {
    memcpy(dst->rad[id], src->rad[id], b);
    memcpy(dst->tau[id], src->tau[id], b);
  }
}

// The state clamps of optimal_estimation (src/retrieval.c:572-598).
void clamp_atm(const ctl_t *ctl, atm_t *a) {
  for (int ip = 0; ip < a->np; ip++) {
    a->p[ip] = std::min(std::max(a->p[ip], 5e-7), 5e4);
    a->t[ip] = std::min(std::max(a->t[ip], 100.0), 400.0);
    for (int ig = 0; ig < ctl->ng; ig++) a->q[ig][ip] = std::min(std::max(a->q[ig][ip], 0.0), 1.0);
    for (int iw = 0; iw < ctl->nw; iw++) a->k[iw][ip] = std::max(a->k[iw][ip], 0.0);
  }
  a->clz = std::max(a->clz, 0.0);
  a->cldz = std::max(a->cldz, 0.1);
  for (int icl = 0; icl < ctl->ncl; icl++) a->clk[icl] = std::max(a->clk[icl], 0.0);
  a->sfz = std::max(a->sfz, 0.0);
  a->sfp = std::max(a->sfp, 0.0);
  a->sft = std::min(std::max(a->sft, 100.0), 400.0);
  for (int isf = 0; isf < ctl->nsf; isf++) a->sfeps[isf] = std::min(std::max(a->sfeps[isf], 0.0), 1.0);
}

}  // namespace

// ------------------------------------------------------------------
// The engine.
// ------------------------------------------------------------------
struct Engine {
  const jur_ret_t *ret = nullptr;
  const ctl_t *ctl = nullptr;  // forward model
  jur_ctx *c = nullptr;
  int method = JUR_JAC_ADJOINT, precision = JUR_FP64;
  // forward-only engine (a server lane for forward requests): radiances from
  // the forward pass, not from the adjoint pass that keeps retrieval costs
  // consistent with their Jacobians
  bool fwd_pass = false;
  std::vector<Job *> jobs;  // active
  DevBuf<LmDesc> d_desc;
  DevBuf<int> d_list, d_status;
  DevBuf<double> d_lmpar, d_q, d_stage;
  DevBuf<size_t> d_off;
  DevPool pool;
  std::vector<Job *> pending;  // retrievals whose device setup is due
  // S_a^-1 by the content of its inputs: retrievals with the same state
  // layout, a priori and settings share it (typical for a served model)
  struct SaEntry {
    double *p;
    int cls, n;
    uint64_t used;
  };
  std::unordered_map<std::string, SaEntry> sa_cache;
  uint64_t sa_clock = 0;
  static constexpr size_t SA_CACHE_MAX = 64;
  ~Engine() {
    for (auto &e : sa_cache) pool.put(e.second.p, e.second.cls);
  }
  DevBuf<atm_t> p_atm;
  DevBuf<int> p_prof, p_xmap, p_pmap;
  DevBuf<double> p_geom, p_rad, p_tau, p_k;
  DevBuf<unsigned char> p_mask;
  // pinned transfer buffers: the atmospheres of a pass, and the values that
  // cross PCIe each round (radiances, K rows, LM vectors)
  HostBuf<atm_t> h_atm;
  HostBuf<double> h_buf;
  hipStream_t st = 0;
  // seconds spent: [0] admission, [1] Jacobian passes, [2] forward passes,
  // [3] dense algebra and bookkeeping
  double prof[4] = {0, 0, 0, 0};

  // Adds a job (host preparation); on error the job is marked failed and
  // not added.  Retrievals wait in pending for setup().
  int add(Job *J);
  // Device setup of the pending retrievals, batched: pooled buffers, S_a
  // built and inverted on the GPU.  Jobs that fail go to done.
  int setup(std::vector<Job *> &done);
  void release(Job *J) {
    for (int b = 0; b < 9; b++) {
      pool.put(J->dp[b], J->dc[b]);
      J->dp[b] = nullptr;
    }
    if (J->ktr) pool.put(J->ktr, J->ktrc);
    J->ktr = nullptr;
  }
  // One round over all active jobs; finished jobs are moved to done.
  int round(std::vector<Job *> &done);

 private:
  int fail(Job *J, int rc) {
    J->rc = rc;
    snprintf(J->err, sizeof J->err, "%s", g_err);
    J->st = ST_DONE;
    return rc;
  }
  int upload_desc();
  int list(const std::vector<int> &ids) { return ids.empty() ? JUR_OK : d_list.upload(ids.data(), ids.size(), st); }
  int eval(const std::vector<int> &ids, const std::vector<char> &wantk);
  int chisq(const std::vector<int> &ids);
  int stage_get(const std::vector<int> &ids, int which);
  int analyze(const std::vector<int> &ids);
};

int Engine::add(Job *J) {
  const ctl_t *sc = J->sctl;
  const int nd = ctl->nd;
  J->nr = J->obs_meas->nr;
  if (J->nr < 1 || J->nr > NR) return fail(J, set_err(JUR_ERR_ARG, "Too many rays!"));
  if (J->atm_apr->np < 2 || J->atm_apr->np > NP) return fail(J, set_err(JUR_ERR_ARG, "Too many data points!"));
  J->m = J->nr * nd;
  J->mv = (int)obs2y_h(sc, J->obs_meas, nullptr);
  J->n = (int)atm2x_h(sc, J->atm_apr, nullptr, nullptr, nullptr);
  copy_obs_rows(J->obs_i, J->obs_meas, nd);
  if (J->kind == JUR_REQ_FORWARD) {
    J->st = ST_EVAL_F_OUT;
    J->bytes = (size_t)J->m * 16;
    jobs.push_back(J);
    return JUR_OK;
  }
  if (J->mv == 0 || J->n == 0) return fail(J, set_err(JUR_ERR_ARG, "Check problem definition!"));
  J->iqa.resize(J->n);
  J->ipa.resize(J->n);
  J->x_a.resize(J->n);
  atm2x_h(sc, J->atm_apr, J->x_a.data(), J->iqa.data(), J->ipa.data());
  J->x_i = J->x_a;
  J->xmap.resize(JUR_XMAP_LEN);
  state_map_h(sc, J->atm_apr, J->xmap.data());
  if (method == JUR_JAC_ADJOINT && (J->xmap[JUR_XMAP_SFZ] >= 0 || J->xmap[JUR_XMAP_SFP] >= 0))
    return fail(J, set_err(JUR_ERR_UNSUPPORTED, "surface altitude/pressure in the state need JUR_JAC_FD"));
  if (J->kind == JUR_REQ_JACOBIAN) {
    J->st = ST_EVAL_FK_OUT;
    J->bytes = (size_t)J->m * (J->n + 2) * 8;
    jobs.push_back(J);
    return JUR_OK;
  }
  // retrieval: measurement vector and errors over all (ray, channel) rows;
  // masked rows carry sig = 0 (set_cov_meas, src/retrieval.c:924-964)
  J->y_m.assign(J->m, 0.0);
  J->sig.assign(J->m, 0.0);
  for (int ir = 0; ir < J->nr; ir++)
    for (int id = 0; id < nd; id++) {
      const double v = J->obs_meas->rad[id][ir];
      if (!isfinite(v)) continue;
      const int i = ir * nd + id;
      J->y_m[i] = v;
      const double sn = ret->err_noise[id];
      const double sf = fabs(ret->err_formod[id] / 100 * v);
      J->sig[i] = 1 / sqrt(POW2(sn) + POW2(sf));
      if (J->sig[i] <= 0 || !isfinite(J->sig[i]))
        return fail(J, set_err(JUR_ERR_ARG, "Check measurement errors (zero standard deviation)!"));
    }
  J->dx.assign(J->n, 0.0);
  J->dy.assign(J->m, 0.0);
  J->b.assign(J->n, 0.0);
  J->step.assign(J->n, 0.0);
  memcpy(J->atm_i, J->atm_apr, sizeof(atm_t));
  int rc;
  if ((rc = cov_apr_inputs(ret, sc, J->atm_apr, *J))) return fail(J, rc);
  const size_t nn = (size_t)J->n * J->n;
  const size_t sz[9] = {nn, nn, nn, (size_t)J->m * J->n, (size_t)J->m, (size_t)J->m,
                        (size_t)J->n, (size_t)J->n, (size_t)J->n};
  for (int b = 0; b < 9; b++) {
    J->dp[b] = pool.get(sz[b], J->dc[b]);
    if (!J->dp[b]) {
      release(J);
      return fail(J, set_err(JUR_ERR_MEM, "device allocation failed"));
    }
  }
  if (method == JUR_JAC_ADJOINT) {
    J->ktr = pool.get((size_t)J->m * J->n, J->ktrc);
    if (!J->ktr) {
      release(J);
      return fail(J, set_err(JUR_ERR_MEM, "device allocation failed"));
    }
  }
  J->bytes = (3 * nn + (size_t)J->m * J->n * (J->ktr ? 2 : 1) + 2 * J->m + 3 * J->n) * 8;
  J->st = ST_EVAL_FK0;
  pending.push_back(J);
  return JUR_OK;
}

int Engine::setup(std::vector<Job *> &done) {
  if (pending.empty()) return JUR_OK;
  // descriptors of the pending jobs at the end of the active list
  const size_t base = jobs.size();
  jobs.insert(jobs.end(), pending.begin(), pending.end());
  int rc;
  if ((rc = upload_desc())) return rc;
  // S_a^-1 from the cache where its inputs were seen before
  std::vector<std::string> key(pending.size());
  std::vector<int> miss;  // indices into jobs
  std::vector<size_t> miss_k;
  for (size_t k = 0; k < pending.size(); k++) {
    Job &J = *pending[k];
    key[k].assign((const char *)J.sa_in.data(), J.sa_in.size() * sizeof(double));
    auto it = sa_cache.find(key[k]);
    if (it != sa_cache.end() && it->second.n == J.n) {
      HIPCHK(hipMemcpyAsync(J.dp[0], it->second.p, (size_t)J.n * J.n * sizeof(double), hipMemcpyDeviceToDevice,
                            st));
      it->second.used = ++sa_clock;
    } else {
      miss.push_back((int)(base + k));
      miss_k.push_back(k);
    }
  }
  // staging: the S_a inputs of the misses, then sig of every pending job
  std::vector<size_t> off_sa(miss.size()), off_sig(pending.size());
  size_t tot = 0;
  for (size_t q = 0; q < miss.size(); q++) {
    off_sa[q] = tot;
    tot += pending[miss_k[q]]->sa_in.size();
  }
  for (size_t k = 0; k < pending.size(); k++) {
    off_sig[k] = tot;
    tot += pending[k]->m;
  }
  if ((rc = h_buf.reserve(tot))) return rc;
  for (size_t q = 0; q < miss.size(); q++) {
    const Job &J = *pending[miss_k[q]];
    memcpy(&h_buf[off_sa[q]], J.sa_in.data(), J.sa_in.size() * sizeof(double));
  }
  for (size_t k = 0; k < pending.size(); k++)
    memcpy(&h_buf[off_sig[k]], pending[k]->sig.data(), pending[k]->m * sizeof(double));
  if ((rc = d_stage.upload(h_buf.data(), tot, st))) return rc;
  std::vector<char> bad(pending.size(), 0);  // S_a of a miss not positive definite
  if (!miss.empty()) {
    if ((rc = list(miss)) || (rc = d_off.upload(off_sa.data(), off_sa.size(), st))) return rc;
    launch_lm_cov_apr(d_desc.p, d_list.p, (int)miss.size(), d_off.p, d_stage.p, st);
    HIPCHK(hipGetLastError());
    if ((rc = d_status.reserve(jobs.size()))) return rc;
    launch_lm_cholinv(d_desc.p, d_list.p, (int)miss.size(), d_status.p, st);
    HIPCHK(hipGetLastError());
    // only the entries of the misses were written
    std::vector<int> status(jobs.size(), 0);
    HIPCHK(hipMemcpyAsync(status.data(), d_status.p, jobs.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (size_t q = 0; q < miss.size(); q++) bad[miss_k[q]] = status[miss[q]] != 0;
    for (size_t q = 0; q < miss.size(); q++) {
      const size_t k = miss_k[q];
      Job &J = *pending[k];
      if (bad[k] || sa_cache.count(key[k])) continue;
      if (sa_cache.size() >= SA_CACHE_MAX) {
        auto old = sa_cache.begin();
        for (auto it = sa_cache.begin(); it != sa_cache.end(); ++it)
          if (it->second.used < old->second.used) old = it;
        pool.put(old->second.p, old->second.cls);
        sa_cache.erase(old);
      }
      SaEntry e{nullptr, 0, J.n, ++sa_clock};
      e.p = pool.get((size_t)J.n * J.n, e.cls);
      if (!e.p) continue;
      HIPCHK(hipMemcpyAsync(e.p, J.dp[0], (size_t)J.n * J.n * sizeof(double), hipMemcpyDeviceToDevice, st));
      sa_cache.emplace(key[k], e);
    }
  }
  std::vector<int> all(pending.size());
  for (size_t k = 0; k < pending.size(); k++) all[k] = (int)(base + k);
  if ((rc = list(all)) || (rc = d_off.upload(off_sig.data(), off_sig.size(), st))) return rc;
  launch_lm_stage(d_desc.p, d_list.p, (int)all.size(), d_off.p, d_stage.p, 3, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  std::vector<Job *> keep(jobs.begin(), jobs.begin() + base);
  for (size_t k = 0; k < pending.size(); k++) {
    Job *J = pending[k];
    J->sa_in.clear();
    if (bad[k]) {
      release(J);
      fail(J, set_err(JUR_ERR_ARG, "a priori covariance is not positive definite"));
      done.push_back(J);
    } else {
      keep.push_back(J);
    }
  }
  jobs.swap(keep);
  pending.clear();
  return JUR_OK;
}

int Engine::upload_desc() {
  std::vector<LmDesc> desc(jobs.size());
  for (size_t r = 0; r < jobs.size(); r++) desc[r] = jobs[r]->desc();
  return d_desc.upload(desc.data(), std::max<size_t>(desc.size(), 1), st);
}

// One forward evaluation of the jobs ids at their current state, with the
// Jacobian for those with wantk set: fills obs_i, and for retrievals the
// scaled compact K, for Jacobian requests the output K.  With the adjoint
// every evaluation is a Jacobian-mode pass (jobs without wantk have no state
// map and skip the reverse sweep), so a job's forward model goes through the
// same arithmetic whatever else shares its pass.
int Engine::eval(const std::vector<int> &ids, const std::vector<char> &wantk) {
  if (ids.empty()) return JUR_OK;
  bool with_k = false;
  for (char w : wantk) with_k |= w != 0;
  const int nd = ctl->nd;
  int rc;
  if (with_k && method == JUR_JAC_FD) {
    // (the finite-difference path is called with all or none of wantk set)
    // kernel() per job (its n + 1 forward models in one device pass)
    for (int r : ids) {
      Job &J = *jobs[r];
      copy_obs_rows(J.obs_i, J.obs_meas, nd);
      std::vector<double> kh((size_t)J.mv * J.n + 1);
      if ((rc = kernel_fd_host(J.sctl, J.kind == JUR_REQ_RETRIEVAL ? J.atm_i : J.atm_apr, J.obs_i, kh.data(),
                               J.mv, J.n, precision))) {
        fail(&J, rc);
        continue;
      }
      if (J.kind == JUR_REQ_JACOBIAN) {
        memcpy(J.kout, kh.data(), (size_t)J.mv * J.n * sizeof(double));
        continue;
      }
      std::vector<double> kc((size_t)J.m * J.n, 0.0);
      size_t v = 0;
      for (int i = 0; i < J.m; i++)
        if (J.sig[i] != 0) {
          for (int j = 0; j < J.n; j++) kc[(size_t)i * J.n + j] = kh[v * J.n + j] * J.sig[i];
          v++;
        }
      HIPCHK(hipMemcpyAsync(J.dp[3], kc.data(), kc.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    }
    return JUR_OK;
  }
  // one device pass over the rays of all jobs in ids, the Jacobian requests
  // that want K first: their K rows are then one range, copied back at once
  {
    auto jk = [&](size_t q) { return jobs[ids[q]]->kind == JUR_REQ_JACOBIAN && wantk[q] != 0; };
    bool lead = true, other = false;
    for (size_t q = 0; q < ids.size() && lead; q++) {
      if (jk(q) && other) lead = false;
      other |= !jk(q);
    }
    if (!lead) {
      std::vector<int> i2;
      std::vector<char> w2;
      for (int pass = 1; pass >= 0; pass--)
        for (size_t q = 0; q < ids.size(); q++)
          if (jk(q) == (pass == 1)) {
            i2.push_back(ids[q]);
            w2.push_back(wantk[q]);
          }
      return eval(i2, w2);
    }
  }
  int nray = 0, maxn = 1;
  for (int r : ids) {
    jobs[r]->ray0 = nray;
    nray += jobs[r]->nr;
    maxn = std::max(maxn, jobs[r]->n);
  }
  const int np = (int)ids.size();
  if ((rc = h_atm.reserve(np))) return rc;
  atm_t *atms = h_atm.data();
  std::vector<int> prof(nray), xm, pmap(np);
  int nmap = 0;
  const bool jac_pass = with_k || (method == JUR_JAC_ADJOINT && !fwd_pass);
  std::vector<double> geom((size_t)7 * nray);
  std::vector<unsigned char> mask((size_t)nray * nd);
  for (int k = 0; k < np; k++) {
    Job &J = *jobs[ids[k]];
    memcpy(&atms[k], J.kind == JUR_REQ_RETRIEVAL ? J.atm_i : J.atm_apr, sizeof(atm_t));
    pmap[k] = wantk[k] ? nmap++ : -1;
    for (int ir = 0; ir < J.nr; ir++) {
      const int i = J.ray0 + ir;
      prof[i] = k;
      double *g = &geom[(size_t)7 * i];
      g[0] = J.obs_meas->time[ir];
      g[1] = J.obs_meas->obsz[ir];
      g[2] = J.obs_meas->obslon[ir];
      g[3] = J.obs_meas->obslat[ir];
      g[4] = J.obs_meas->vpz[ir];
      g[5] = J.obs_meas->vplon[ir];
      g[6] = J.obs_meas->vplat[ir];
      for (int id = 0; id < nd; id++) mask[(size_t)i * nd + id] = !isfinite(J.obs_meas->rad[id][ir]);
    }
    if (wantk[k]) xm.insert(xm.end(), J.xmap.begin(), J.xmap.end());
  }
  // stream-ordered uploads from buffers that live to the end of this call;
  // the pass synchronizes the stream before it returns
  if ((rc = p_atm.upload_async(atms, np, st))) return rc;
  if ((rc = p_prof.upload_async(prof.data(), nray, st))) return rc;
  if ((rc = p_geom.upload_async(geom.data(), geom.size(), st))) return rc;
  if ((rc = p_mask.upload_async(mask.data(), mask.size(), st))) return rc;
  if ((rc = p_rad.reserve((size_t)nray * nd))) return rc;
  if ((rc = p_tau.reserve((size_t)nray * nd))) return rc;
  BatchView B{np, p_atm.p, nray, p_prof.p, p_geom.p};
  if (!with_k) maxn = 1;
  if (jac_pass) {
    if ((rc = p_pmap.upload_async(pmap.data(), np, st))) return rc;
    if (xm.empty()) xm.assign(JUR_XMAP_LEN, -1);  // no job wants K: one unused map
    if ((rc = p_xmap.upload_async(xm.data(), xm.size(), st))) return rc;
    if ((rc = p_k.reserve((size_t)nray * nd * maxn))) return rc;
    JacArgs ja;
    ja.xmap = p_xmap.p;
    ja.nmap = nmap;
    ja.prof_map = p_pmap.p;
    ja.ncol = maxn;
    ja.k = p_k.p;
    rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st, &ja);
  } else {
    rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st);
  }
  if (rc) {
    // the uploads above read this call's buffers
    (void)hipStreamSynchronize(st);
    // A ray that fails (observer below the surface, too many LOS points)
    // fails the pass: evaluate the jobs one by one so that only its own job
    // fails.
    if (ids.size() == 1) {
      fail(jobs[ids[0]], rc);
      return JUR_OK;
    }
    for (size_t k = 0; k < ids.size(); k++)
      if ((rc = eval(std::vector<int>{ids[k]}, std::vector<char>{wantk[k]}))) return rc;
    return JUR_OK;
  }
  // radiances, transmittances and the K rows of the Jacobian requests, with
  // one wait
  std::vector<size_t> koff(ids.size(), 0);
  size_t ktot = 0;
  if (with_k)
    for (size_t q = 0; q < ids.size(); q++) {
      const Job &J = *jobs[ids[q]];
      if (J.kind != JUR_REQ_JACOBIAN || !wantk[q]) continue;
      koff[q] = ktot;
      ktot += (size_t)J.m * maxn;
    }
  if ((rc = h_buf.reserve((size_t)nray * nd * 2 + ktot))) return rc;
  HIPCHK(hipMemcpyAsync(h_buf.data(), p_rad.p, (size_t)nray * nd * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h_buf.data() + (size_t)nray * nd, p_tau.p, (size_t)nray * nd * sizeof(double),
                        hipMemcpyDeviceToHost, st));
  double *krows = h_buf.data() + (size_t)nray * nd * 2;
  // the Jacobian requests lead the pass (ray0 = koff / (nd maxn)): their rows
  // are the first ktot values of p_k
  if (ktot) HIPCHK(hipMemcpyAsync(krows, p_k.p, ktot * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int r : ids) {
    Job &J = *jobs[r];
    copy_obs_rows(J.obs_i, J.obs_meas, nd);
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const size_t i = (size_t)(J.ray0 + ir) * nd + id;
        J.obs_i->rad[id][ir] = h_buf[i];
        J.obs_i->tau[id][ir] = h_buf[(size_t)nray * nd + i];
      }
  }
  if (!with_k) return JUR_OK;
  // Jacobian requests: their finite rows
  for (size_t q = 0; q < ids.size(); q++) {
    Job &J = *jobs[ids[q]];
    if (J.kind != JUR_REQ_JACOBIAN || !wantk[q]) continue;
    const double *rows = krows + koff[q];
    size_t v = 0;
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++)
        if (isfinite(J.obs_meas->rad[id][ir])) {
          memcpy(J.kout + v * J.n, &rows[((size_t)ir * nd + id) * maxn], J.n * sizeof(double));
          v++;
        }
  }
  // retrievals: scaled compact K
  std::vector<int> lm;
  int maxrows = 0;
  for (size_t q = 0; q < ids.size(); q++)
    if (wantk[q] && jobs[ids[q]]->kind == JUR_REQ_RETRIEVAL) {
      lm.push_back(ids[q]);
      maxrows = std::max(maxrows, jobs[ids[q]]->m);
    }
  if (lm.empty()) return JUR_OK;
  if ((rc = upload_desc())) return rc;
  if ((rc = list(lm))) return rc;
  launch_lm_gather(d_desc.p, d_list.p, (int)lm.size(), maxrows, p_k.p, nd, maxn, st);
  HIPCHK(hipGetLastError());
  return JUR_OK;
}

// dx, dy and the cost of the retrievals ids (cost_function,
// src/retrieval.c:340-369).
int Engine::chisq(const std::vector<int> &ids) {
  std::vector<int> live;
  for (int r : ids)
    if (!jobs[r]->failed()) live.push_back(r);
  if (live.empty()) return JUR_OK;
  const int nd = ctl->nd;
  int rc;
  std::vector<size_t> off(live.size());
  size_t tot = 0;
  for (size_t k = 0; k < live.size(); k++) {
    off[k] = tot;
    tot += jobs[live[k]]->n + jobs[live[k]]->m;
  }
  if ((rc = h_buf.reserve(tot))) return rc;
  for (size_t k = 0; k < live.size(); k++) {
    Job &J = *jobs[live[k]];
    for (int i = 0; i < J.n; i++) J.dx[i] = J.x_i[i] - J.x_a[i];
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const int i = ir * nd + id;
        J.dy[i] = J.sig[i] != 0 ? J.y_m[i] - J.obs_i->rad[id][ir] : 0.0;
      }
    memcpy(&h_buf[off[k]], J.dx.data(), J.n * sizeof(double));
    memcpy(&h_buf[off[k] + J.n], J.dy.data(), J.m * sizeof(double));
  }
  if ((rc = upload_desc())) return rc;
  if ((rc = list(live))) return rc;
  if ((rc = d_stage.upload(h_buf.data(), tot, st))) return rc;
  if ((rc = d_off.upload(off.data(), off.size(), st))) return rc;
  launch_lm_stage(d_desc.p, d_list.p, (int)live.size(), d_off.p, d_stage.p, 0, st);
  HIPCHK(hipGetLastError());
  if ((rc = d_q.reserve(jobs.size()))) return rc;
  launch_lm_quad(d_desc.p, d_list.p, (int)live.size(), d_q.p, st);
  HIPCHK(hipGetLastError());
  std::vector<double> q(jobs.size());
  HIPCHK(hipMemcpyAsync(q.data(), d_q.p, jobs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int r : live) {
    Job &J = *jobs[r];
    double cm = 0;
    for (int i = 0; i < J.m; i++) cm += POW2(J.dy[i] * J.sig[i]);
    J.chisq = (cm + q[r]) / (double)J.mv;
  }
  return JUR_OK;
}

// b (which = 1) or step (which = 2) of the jobs ids to the host.
int Engine::stage_get(const std::vector<int> &ids, int which) {
  if (ids.empty()) return JUR_OK;
  std::vector<size_t> off(ids.size());
  size_t tot = 0;
  for (size_t k = 0; k < ids.size(); k++) {
    off[k] = tot;
    tot += jobs[ids[k]]->n;
  }
  int rc;
  if ((rc = d_stage.reserve(tot))) return rc;
  if ((rc = d_off.upload(off.data(), off.size(), st))) return rc;
  if ((rc = list(ids))) return rc;
  launch_lm_stage(d_desc.p, d_list.p, (int)ids.size(), d_off.p, d_stage.p, which, st);
  HIPCHK(hipGetLastError());
  if ((rc = h_buf.reserve(tot))) return rc;
  HIPCHK(hipMemcpyAsync(h_buf.data(), d_stage.p, tot * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (size_t k = 0; k < ids.size(); k++) {
    Job &J = *jobs[ids[k]];
    memcpy(which == 1 ? J.b.data() : J.step.data(), &h_buf[off[k]], J.n * sizeof(double));
  }
  return JUR_OK;
}

int Engine::round(std::vector<Job *> &done) {
  int rc;
  const int kr = std::max(ret ? ret->kernel_recomp : 1, 1);
  const int itmax = ret ? ret->conv_itmax : 0;
  const size_t nj = jobs.size();
  // ---- model evaluations ----
  std::vector<int> need_f, need_fk;
  for (size_t r = 0; r < nj; r++) {
    const Stage s = jobs[r]->st;
    if (s == ST_EVAL_F || s == ST_EVAL_F_OUT) need_f.push_back((int)r);
    if (s == ST_EVAL_FK0 || s == ST_EVAL_FK || s == ST_EVAL_FK_OUT || s == ST_EVAL_FT) need_fk.push_back((int)r);
  }
  auto tnow = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  const double t0 = tnow();
  double t1;
  if (method == JUR_JAC_FD) {
    if ((rc = eval(need_fk, std::vector<char>(need_fk.size(), 1)))) return rc;
    t1 = tnow();
    if ((rc = eval(need_f, std::vector<char>(need_f.size(), 0)))) return rc;
  } else {
    // one pass for every evaluation of the round
    std::vector<int> all(need_fk);
    all.insert(all.end(), need_f.begin(), need_f.end());
    std::vector<char> wk(all.size(), 0);
    for (size_t q = 0; q < need_fk.size(); q++) wk[q] = 1;
    if ((rc = eval(all, wk))) return rc;
    t1 = tnow();
  }
  const double t2 = tnow();
  prof[fwd_pass ? 2 : 1] += t1 - t0;
  prof[2] += t2 - t1;
  for (int r : need_f)
    if (jobs[r]->st == ST_EVAL_F_OUT) jobs[r]->st = ST_DONE;
  for (int r : need_fk)
    if (jobs[r]->st == ST_EVAL_FK_OUT) jobs[r]->st = ST_DONE;
  // first evaluation of a retrieval: cost of the a priori
  std::vector<int> first;
  for (int r : need_fk)
    if (jobs[r]->st == ST_EVAL_FK0) first.push_back(r);
  if ((rc = chisq(first))) return rc;
  for (int r : first) {
    Job &J = *jobs[r];
    if (J.failed()) continue;
    if (J.hist) J.hist[0] = J.chisq;
    J.it = 1;
    J.fk_done = true;
    J.st = ST_ITER;
    if (itmax < 1) {
      J.it = itmax;
      J.st = ST_DONE;
    }
  }
  for (int r : need_fk)
    if (jobs[r]->st == ST_EVAL_FK) {
      jobs[r]->fk_done = true;
      jobs[r]->st = ST_ITER;
    }
  // trial evaluations: accept or reject the step
  std::vector<int> trials;
  for (int r : need_f)
    if (jobs[r]->st == ST_EVAL_F) trials.push_back(r);
  for (int r : need_fk)
    if (jobs[r]->st == ST_EVAL_FT) trials.push_back(r);
  if ((rc = chisq(trials))) return rc;
  for (int r : trials) {
    Job &J = *jobs[r];
    if (J.failed()) continue;
    const bool with_k = J.st == ST_EVAL_FT;
    bool end_iter, accepted = false;
    if (J.chisq > J.chisq_old) {
      J.lmpar *= 10;
      for (int i = 0; i < J.n; i++) J.x_i[i] -= J.step[i];
      J.it2++;
      end_iter = J.it2 >= 20;
      if (!end_iter) J.st = ST_TRIAL;
    } else {
      J.lmpar /= 10;
      end_iter = true;
      accepted = true;
    }
    if (end_iter) {
      if (J.hist) J.hist[J.it] = J.chisq;
      double disq = 0;
      for (int i = 0; i < J.n; i++) disq += J.step[i] * J.b[i];
      disq /= (double)J.n;
      if ((J.it == 1 || J.it % kr == 0) && disq < ret->conv_dmin) {
        J.st = ST_DONE;
      } else if (++J.it > itmax) {
        J.it = itmax;
        J.st = ST_DONE;
      } else {
        J.fk_done = false;
        J.st = ST_ITER;
        if (accepted && with_k && J.it % kr == 0) {
          // the trial's K is the one at the accepted state
          std::swap(J.dp[3], J.ktr);
          std::swap(J.dc[3], J.ktrc);
          J.fk_done = true;
        }
      }
    }
  }
  // ---- start of outer iterations: Jacobian refresh, normal equations ----
  std::vector<int> need_cov, need_rhs;
  int maxn_cov = 1, maxn_rhs = 1;
  for (size_t r = 0; r < nj; r++) {
    Job &J = *jobs[r];
    if (J.st != ST_ITER) continue;
    if (J.it > 1 && J.it % kr == 0 && !J.fk_done) {
      J.st = ST_EVAL_FK;
      continue;
    }
    J.chisq_old = J.chisq;
    if (J.it == 1 || J.it % kr == 0) {
      need_cov.push_back((int)r);
      maxn_cov = std::max(maxn_cov, J.n);
    }
    need_rhs.push_back((int)r);
    maxn_rhs = std::max(maxn_rhs, J.n);
    J.it2 = 0;
    J.st = ST_TRIAL;
  }
  if (!need_rhs.empty()) {
    if ((rc = upload_desc())) return rc;
    if ((rc = list(need_cov))) return rc;
    launch_lm_normal(d_desc.p, d_list.p, (int)need_cov.size(), maxn_cov, st);
    HIPCHK(hipGetLastError());
    if ((rc = list(need_rhs))) return rc;
    launch_lm_rhs(d_desc.p, d_list.p, (int)need_rhs.size(), maxn_rhs, st);
    HIPCHK(hipGetLastError());
    if ((rc = stage_get(need_rhs, 1))) return rc;
  }
  // ---- trial steps ----
  std::vector<int> trial;
  for (size_t r = 0; r < nj; r++)
    if (jobs[r]->st == ST_TRIAL) trial.push_back((int)r);
  if (!trial.empty()) {
    std::vector<double> lm(nj);
    for (size_t r = 0; r < nj; r++) lm[r] = jobs[r]->lmpar;
    if ((rc = d_lmpar.upload(lm.data(), nj, st))) return rc;
    if ((rc = d_status.reserve(nj))) return rc;
    if ((rc = upload_desc())) return rc;
    if ((rc = list(trial))) return rc;
    launch_lm_solve(d_desc.p, d_list.p, (int)trial.size(), d_lmpar.p, d_status.p, st);
    HIPCHK(hipGetLastError());
    std::vector<int> status(nj);
    HIPCHK(hipMemcpyAsync(status.data(), d_status.p, nj * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((rc = stage_get(trial, 2))) return rc;
    for (int r : trial) {
      Job &J = *jobs[r];
      if (status[r]) {
        fail(&J, set_err(JUR_ERR_ARG, "normal matrix is not positive definite"));
        continue;
      }
      for (int i = 0; i < J.n; i++) J.x_i[i] += J.step[i];
      memcpy(J.atm_i, J.atm_apr, sizeof(atm_t));
      x2atm_h(J.sctl, J.x_i.data(), J.atm_i);
      clamp_atm(J.sctl, J.atm_i);
      // with its Jacobian when the next iteration would recompute K
      J.st = (J.ktr && (J.it + 1) % kr == 0 && J.it + 1 <= itmax) ? ST_EVAL_FT : ST_EVAL_F;
    }
  }
  // ---- error analysis of finished retrievals ----
  std::vector<int> ana;
  for (size_t r = 0; r < nj; r++)
    if (jobs[r]->st == ST_DONE && jobs[r]->ana && !jobs[r]->failed() && jobs[r]->kind == JUR_REQ_RETRIEVAL)
      ana.push_back((int)r);
  if (!ana.empty() && (rc = analyze(ana))) return rc;
  prof[3] += tnow() - t2;
  // ---- finished jobs leave ----
  std::vector<Job *> keep;
  for (Job *J : jobs) {
    if (J->st == ST_DONE) {
      if (J->niter) *J->niter = J->it;
      release(J);
      done.push_back(J);
    } else {
      keep.push_back(J);
    }
  }
  jobs.swap(keep);
  return JUR_OK;
}

// copy_atm(ctl, dst, src, 1) (src/jurassic.c:2921-2972): the geometry of
// src with the retrievable quantities zeroed (surface emissivities 1).
static void copy_atm_init(const ctl_t *ctl, atm_t *dst, const atm_t *src) {
  memcpy(dst, src, sizeof(atm_t));
  for (int ip = 0; ip < dst->np; ip++) {
    dst->p[ip] = 0;
    dst->t[ip] = 0;
    for (int ig = 0; ig < ctl->ng; ig++) dst->q[ig][ip] = 0;
    for (int iw = 0; iw < ctl->nw; iw++) dst->k[iw][ip] = 0;
    dst->clz = 0;
    dst->cldz = 0;
    for (int icl = 0; icl < ctl->ncl; icl++) dst->clk[icl] = 0;
    dst->sfz = 0;
    dst->sfp = 0;
    dst->sft = 0;
    for (int isf = 0; isf < ctl->nsf; isf++) dst->sfeps[isf] = 1;
  }
}

// Pointer to the atm_t value of state element i (quantity iqa, level ipa),
// as analyze_avk addresses it (src/retrieval.c:257-312).
static double *atm_elem(const ctl_t *ctl, atm_t *a, int iq, int ip) {
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl;
  if (iq == IDXP) return &a->p[ip];
  if (iq == IDXT) return &a->t[ip];
  if (iq >= 2 && iq < 2 + ng) return &a->q[iq - 2][ip];
  if (iq >= 2 + ng && iq < 2 + ng + nw) return &a->k[iq - 2 - ng][ip];
  int q = iq - 2 - ng - nw;
  if (q == 0) return &a->clz;
  if (q == 1) return &a->cldz;
  if (q >= 2 && q < 2 + ncl) return &a->clk[q - 2];
  q -= 2 + ncl;
  if (q == 0) return &a->sfz;
  if (q == 1) return &a->sfp;
  if (q == 2) return &a->sft;
  return &a->sfeps[q - 3];
}

// Error analysis (src/retrieval.c:668-726) of finished retrievals at their
// last Jacobian: covariance, gain and averaging kernels on the device, the
// diagonals and the atmosphere mappings on the host.
int Engine::analyze(const std::vector<int> &ids) {
  int rc;
  int maxn = 1;
  for (int r : ids) {
    Job &J = *jobs[r];
    J.g = pool.get((size_t)J.n * J.m, J.gc);
    if (!J.g) return set_err(JUR_ERR_MEM, "device allocation failed");
    maxn = std::max(maxn, J.n);
  }
  if ((rc = upload_desc())) return rc;
  if ((rc = list(ids))) return rc;
  launch_lm_normal(d_desc.p, d_list.p, (int)ids.size(), maxn, st);
  HIPCHK(hipGetLastError());
  if ((rc = d_status.reserve(jobs.size()))) return rc;
  launch_ana(d_desc.p, d_list.p, (int)ids.size(), d_status.p, st, 0);
  HIPCHK(hipGetLastError());
  std::vector<int> status(jobs.size(), 0);
  HIPCHK(hipMemcpyAsync(status.data(), d_status.p, jobs.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int> ok;
  for (int r : ids) {
    if (status[r]) fail(jobs[r], set_err(JUR_ERR_ARG, "retrieval covariance is not positive definite"));
    else ok.push_back(r);
  }
  if (!ok.empty()) {
    if ((rc = list(ok))) return rc;
    launch_ana(d_desc.p, d_list.p, (int)ok.size(), d_status.p, st, 1);
    launch_ana(d_desc.p, d_list.p, (int)ok.size(), d_status.p, st, 2);
    HIPCHK(hipGetLastError());
  }
  std::vector<std::vector<double>> cov(ok.size()), gain(ok.size()), avk(ok.size());
  for (size_t q = 0; q < ok.size(); q++) {
    Job &J = *jobs[ok[q]];
    const size_t nn = (size_t)J.n * J.n;
    cov[q].resize(nn);
    gain[q].resize((size_t)J.n * J.m);
    avk[q].resize(nn);
    HIPCHK(hipMemcpyAsync(cov[q].data(), J.dp[0], nn * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(gain[q].data(), J.g, gain[q].size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(avk[q].data(), J.dp[2], nn * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  for (int r : ids) {
    pool.put(jobs[r]->g, jobs[r]->gc);
    jobs[r]->g = nullptr;
  }
  const int nd = ctl->nd;
  for (size_t q = 0; q < ok.size(); q++) {
    Job &J = *jobs[ok[q]];
    const jur_ana_t &A = *J.ana;
    const int n = J.n;
    const double *C = cov[q].data(), *G = gain[q].data(), *V = avk[q].data();
    // measurement rows of obs2y: the finite rows of the (ray, channel) grid
    std::vector<int> rows;
    for (int i = 0; i < J.m; i++)
      if (J.sig[i] != 0) rows.push_back(i);
    const int mv = (int)rows.size();
    if (A.cov) memcpy(A.cov, C, (size_t)n * n * sizeof(double));
    if (A.avk) memcpy(A.avk, V, (size_t)n * n * sizeof(double));
    if (A.corr)
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
          A.corr[(size_t)i * n + j] = C[(size_t)i * n + j] / sqrt(C[(size_t)i * n + i]) / sqrt(C[(size_t)j * n + j]);
    if (A.gain)
      for (int i = 0; i < n; i++)
        for (int jj = 0; jj < mv; jj++) A.gain[(size_t)i * mv + jj] = G[(size_t)i * J.m + rows[jj]];
    // standard deviations: sqrt(diag(cov)), sqrt(diag(G S G^T)) for the
    // noise and forward-model errors of set_cov_meas (src/retrieval.c:924-964)
    std::vector<double> et(n), en(n), ef(n), cont(n, 0.0), res(n);
    for (int i = 0; i < n; i++) {
      et[i] = sqrt(C[(size_t)i * n + i]);
      double sn = 0, sf = 0;
      for (int jj = 0; jj < mv; jj++) {
        const int row = rows[jj], id = row % nd;
        const double g = G[(size_t)i * J.m + row];
        sn += POW2(g * ret->err_noise[id]);
        sf += POW2(g * fabs(ret->err_formod[id] / 100 * J.y_m[row]));
      }
      en[i] = sqrt(sn);
      ef[i] = sqrt(sf);
    }
    // analyze_avk_quantity: row sums within the quantity, 1 / diagonal
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < n; j++)
        if (J.iqa[j] == J.iqa[i]) cont[i] += V[(size_t)i * n + j];
      res[i] = 1 / V[(size_t)i * n + i];
    }
    if (A.err_total) memcpy(A.err_total, et.data(), n * sizeof(double));
    if (A.err_noise) memcpy(A.err_noise, en.data(), n * sizeof(double));
    if (A.err_formod) memcpy(A.err_formod, ef.data(), n * sizeof(double));
    if (A.cont) memcpy(A.cont, cont.data(), n * sizeof(double));
    if (A.res) memcpy(A.res, res.data(), n * sizeof(double));
    const std::vector<double> *vx[5] = {&et, &en, &ef, &cont, &res};
    atm_t *va[5] = {A.atm_err_total, A.atm_err_noise, A.atm_err_formod, A.atm_cont, A.atm_res};
    for (int k = 0; k < 5; k++) {
      if (!va[k]) continue;
      copy_atm_init(J.sctl, va[k], J.atm_i);
      for (int i = 0; i < n; i++) *atm_elem(J.sctl, va[k], J.iqa[i], J.ipa[i]) = (*vx[k])[i];
    }
  }
  return JUR_OK;
}

int optimal_estimation_batch(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas, obs_t *obs_i,
                             atm_t *atm_apr, atm_t *atm_i, double *chisq_hist, int *niter, int method,
                             int precision, const jur_ana_t *ana) {
  if (!ret || !ctl || n < 0 || (n > 0 && (!obs_meas || !obs_i || !atm_apr || !atm_i)))
    return set_err(JUR_ERR_ARG, "invalid argument");
  if (method != JUR_JAC_FD && method != JUR_JAC_ADJOINT) return set_err(JUR_ERR_ARG, "invalid Jacobian method");
  if (precision != JUR_FP64 && precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (ret->conv_itmax < 0) return set_err(JUR_ERR_ARG, "invalid CONV_ITMAX");
  if (n == 0) return JUR_OK;
  Engine E;
  E.ret = ret;
  E.ctl = ctl;
  E.method = method;
  E.precision = precision;
  int rc = ctx_for(ctl, &E.c);
  if (rc) return rc;
  if (method == JUR_JAC_ADJOINT && (rc = check_jac_ctl(E.c))) return rc;
  if (E.c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "optimal estimation with a FOV is not available");
  std::vector<Job> jobs(n);
  for (int r = 0; r < n; r++) {
    Job &J = jobs[r];
    J.kind = JUR_REQ_RETRIEVAL;
    J.sctl = const_cast<ctl_t *>(ctl);
    J.atm_apr = &atm_apr[r];
    J.atm_i = &atm_i[r];
    J.obs_meas = &obs_meas[r];
    J.obs_i = &obs_i[r];
    J.hist = chisq_hist ? chisq_hist + (size_t)r * (ret->conv_itmax + 1) : nullptr;
    J.niter = niter ? niter + r : nullptr;
    J.ana = ana ? ana + r : nullptr;
    if ((rc = E.add(&J))) return set_err(rc, "retrieval %d: %s", r, J.err);
  }
  std::vector<Job *> done;
  if ((rc = E.setup(done))) return rc;
  while (!E.jobs.empty())
    if ((rc = E.round(done))) return rc;
  for (int r = 0; r < n; r++)
    if (jobs[r].failed()) return set_err(jobs[r].rc, "retrieval %d: %s", r, jobs[r].err);
  return JUR_OK;
}

}  // namespace jur

// ------------------------------------------------------------------
// The request server.
// ------------------------------------------------------------------
// One worker of the server: its own engine, device context (work buffers)
// and HIP stream, so that lanes run their passes concurrently.
struct jur_lane {
  jur::Engine E;
  std::unique_ptr<jur_ctx> ctx;
  hipStream_t st = nullptr;
  std::thread th;
  unsigned kinds = 7u;  // request kinds the lane takes (bit JUR_REQ_*)
  unsigned paths = 3u;  // bit 0: requests without limb rays, bit 1: with
  bool takes(const jur::Job *J) const { return (kinds >> J->kind & 1u) && (paths >> (J->limb ? 1 : 0) & 1u); }
};

// Does ray ir look through a tangent point?  The straight line from the
// observer through the view point (geo2cart, src/jurassic.c:3500-3516)
// comes closest to the Earth's centre ahead of the observer and above the
// ground: a limb view, whose LOS crosses the atmosphere twice.  Nadir and
// slant views reach the ground first.
static bool ray_is_limb(const obs_t *o, int ir) {
  auto cart = [](double z, double lon, double lat, double *x) {
    const double r = z + JUR_RE, lo = lon * M_PI / 180, la = lat * M_PI / 180;
    x[0] = r * cos(la) * cos(lo);
    x[1] = r * cos(la) * sin(lo);
    x[2] = r * sin(la);
  };
  double xo[3], xv[3], d[3];
  cart(o->obsz[ir], o->obslon[ir], o->obslat[ir], xo);
  cart(o->vpz[ir], o->vplon[ir], o->vplat[ir], xv);
  double n = 0;
  for (int i = 0; i < 3; i++) {
    d[i] = xv[i] - xo[i];
    n += d[i] * d[i];
  }
  n = sqrt(n);
  if (!(n > 0)) return false;
  double t = 0;
  for (int i = 0; i < 3; i++) t -= xo[i] * d[i] / n;
  if (!(t > 0)) return false;
  double rm = 0;
  for (int i = 0; i < 3; i++) {
    const double c = xo[i] + t * d[i] / n;
    rm += c * c;
  }
  return sqrt(rm) > JUR_RE;
}

struct jur_server {
  std::vector<std::unique_ptr<jur_lane>> lanes;
  ctl_t ctl;
  jur_ret_t ret;
  int device = 0;
  size_t budget = 0;
  int max_batch = 0;
  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<jur::Job *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<jur::Job>> all;
  uint64_t next_id = 1;
  size_t in_use = 0;
  long in_flight = 0;
  bool stop = false;
  int fatal = JUR_OK;
  char fatal_msg[256] = "";
  // statistics
  std::atomic<long> rounds{0}, max_active{0};

  void loop(jur_lane *L);
};

void jur_server::loop(jur_lane *L) {
  using namespace jur;
  auto now = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  if (hipSetDevice(device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(mu);
    fatal = JUR_ERR_HIP;
    snprintf(fatal_msg, sizeof fatal_msg, "hipSetDevice failed");
    cv_out.notify_all();
    return;
  }
  Engine &E = L->E;
  for (;;) {
    std::vector<Job *> admit;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto queued = [&] {
        for (const Job *J : queue)
          if (L->takes(J)) return true;
        return false;
      };
      cv_in.wait(lk, [&] { return stop || fatal || !E.jobs.empty() || queued(); });
      if (stop || fatal) return;
      // admission: the lane's kinds in arrival order while the device
      // footprint fits (an idle lane always takes one request)
      for (auto it = queue.begin(); it != queue.end();) {
        Job *J = *it;
        if (!L->takes(J)) {
          ++it;
          continue;
        }
        const size_t need = (size_t)J->obs_meas->nr * ctl.nd * 8 * 4 +
                            (J->kind == JUR_REQ_FORWARD ? 0 : (size_t)J->obs_meas->nr * ctl.nd * NP * 8 * 2);
        const bool idle = E.jobs.empty() && admit.empty();
        if (!idle && (in_use + need > budget || (max_batch > 0 && in_flight >= max_batch))) break;
        it = queue.erase(it);
        in_use += need;
        in_flight++;
        J->admit_bytes = need;
        admit.push_back(J);
      }
    }
    const double t0 = now();
    std::vector<Job *> done;
    for (Job *J : admit) {
      J->t_start = t0;
      if (E.add(J)) done.push_back(J);  // failed at admission
    }
    int rc = E.setup(done);
    E.prof[0] += now() - t0;
    if (!rc && !E.jobs.empty()) rc = E.round(done);
    rounds++;
    const double t1 = now();
    std::lock_guard<std::mutex> lk(mu);
    max_active = std::max<long>(max_active, in_flight);
    if (rc) {
      fatal = rc;
      snprintf(fatal_msg, sizeof fatal_msg, "%s", jur::g_err);
      for (Job *J : E.jobs) {
        J->rc = rc;
        snprintf(J->err, sizeof J->err, "%s", jur::g_err);
        done.push_back(J);
      }
      E.jobs.clear();
      for (Job *J : queue) {
        J->rc = rc;
        snprintf(J->err, sizeof J->err, "%s", jur::g_err);
        J->t_done = t1;
      }
      queue.clear();
    }
    for (Job *J : done) {
      J->t_done = t1;
      in_use -= std::min(in_use, J->admit_bytes);
      in_flight--;
      J->st = ST_DONE;
    }
    if (!done.empty() || rc) cv_out.notify_all();
    if (rc) {
      cv_in.notify_all();
      return;  // the device state is unknown: stop serving
    }
  }
}

extern "C" {

int jur_gpu_optimal_estimation(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas, obs_t *obs_i,
                               atm_t *atm_apr, atm_t *atm_i, double *chisq_hist, int *niter, int method,
                               int precision) {
  return jur::optimal_estimation_batch(ret, ctl, n, obs_meas, obs_i, atm_apr, atm_i, chisq_hist, niter, method,
                                       precision, nullptr);
}

int jur_gpu_optimal_estimation_ana(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas, obs_t *obs_i,
                                   atm_t *atm_apr, atm_t *atm_i, double *chisq_hist, int *niter, int method,
                                   int precision, const jur_ana_t *ana) {
  return jur::optimal_estimation_batch(ret, ctl, n, obs_meas, obs_i, atm_apr, atm_i, chisq_hist, niter, method,
                                       precision, ana);
}

int jur_server_create(const ctl_t *ctl, const jur_ret_t *ret, const jur_server_opts_t *opt, jur_server_t **out) {
  using namespace jur;
  if (!ctl || !ret || !out) return set_err(JUR_ERR_ARG, "NULL argument");
  const int method = opt ? opt->method : JUR_JAC_ADJOINT;
  const int precision = opt ? opt->precision : JUR_FP32;
  // forward_lane: lane 0 takes the forward requests, so that forward
  // models do not wait for Jacobian rounds; the others take the rest
  const bool split = opt && opt->forward_lane;
  // path_split: lanes 0 .. lanes - 1 take the requests with limb rays, the
  // last one the others
  const bool psplit = opt && opt->path_split;
  const int nlanes = ((opt && opt->lanes > 0) ? opt->lanes : 1) + (split || psplit ? 1 : 0);
  if (split && psplit) return set_err(JUR_ERR_ARG, "forward_lane and path_split exclude each other");
  if (method != JUR_JAC_FD && method != JUR_JAC_ADJOINT) return set_err(JUR_ERR_ARG, "invalid method");
  if (precision != JUR_FP64 && precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (nlanes > 16) return set_err(JUR_ERR_ARG, "at most 16 lanes");
  if (opt && opt->lanes < 0) return set_err(JUR_ERR_ARG, "negative lane count");
  std::unique_ptr<jur_server> s(new jur_server());
  memcpy(&s->ctl, ctl, sizeof(ctl_t));
  memcpy(&s->ret, ret, sizeof(jur_ret_t));
  s->budget = (opt && opt->max_bytes) ? opt->max_bytes : ((size_t)16 << 30);
  s->max_batch = opt ? opt->max_batch : 0;
  if (hipGetDevice(&s->device) != hipSuccess) return set_err(JUR_ERR_HIP, "hipGetDevice failed");
  int rc;
  for (int l = 0; l < nlanes; l++) {
    std::unique_ptr<jur_lane> L(new jur_lane());
    jur_ctx_t *c = nullptr;
    if ((rc = jur_ctx_create(ctl, &c))) return rc;
    L->ctx.reset(c);
    if (method == JUR_JAC_ADJOINT && (rc = check_jac_ctl(c))) return rc;
    if (c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "the server does not apply a FOV");
    // default priority, non-blocking: lanes overlap on the device
    HIPCHK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
    L->E.ret = &s->ret;
    L->E.ctl = &s->ctl;
    L->E.c = c;
    L->E.method = method;
    L->E.precision = precision;
    L->E.st = L->st;
    if (split) {
      L->kinds = l == 0 ? 1u << JUR_REQ_FORWARD : (1u << JUR_REQ_JACOBIAN | 1u << JUR_REQ_RETRIEVAL);
      L->E.fwd_pass = l == 0;
    }
    if (psplit) L->paths = l == nlanes - 1 ? 1u : 2u;
    s->lanes.push_back(std::move(L));
  }
  jur_server *raw = s.get();
  for (auto &L : raw->lanes) {
    jur_lane *lp = L.get();
    L->th = std::thread([raw, lp] { raw->loop(lp); });
  }
  *out = s.release();
  return JUR_OK;
}

void jur_server_destroy(jur_server_t *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_in.notify_all();
  for (auto &L : s->lanes)
    if (L->th.joinable()) L->th.join();
  for (auto &L : s->lanes) {
    L->E.jobs.clear();
    if (L->st) (void)hipStreamDestroy(L->st);
  }
  delete s;
}

int jur_server_submit(jur_server_t *s, const jur_request_t *req, uint64_t *id) {
  using namespace jur;
  if (!s || !req || !id || !req->atm || !req->obs) return set_err(JUR_ERR_ARG, "NULL argument");
  if (req->kind != JUR_REQ_FORWARD && req->kind != JUR_REQ_JACOBIAN && req->kind != JUR_REQ_RETRIEVAL)
    return set_err(JUR_ERR_ARG, "invalid request kind");
  const double t_submit = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  if (req->obs->nr < 1) return set_err(JUR_ERR_ARG, "request without rays");
  if (req->obs->nr > NR) return set_err(JUR_ERR_ARG, "Too many rays!");
  std::unique_ptr<Job> J(new Job());
  J->kind = req->kind;
  // The job's ctl only defines its state vector (atm2x and friends read the
  // counts and the RET* fields); the forward model is the server's.  The
  // finite-difference Jacobian runs formod on it and needs all of it.
  J->own_ctl.reset(new ctl_t);
  ctl_t *sc = J->own_ctl.get();
  if (s->lanes[0]->E.method == JUR_JAC_FD) {
    memcpy(sc, &s->ctl, sizeof(ctl_t));
  } else {
    sc->ng = s->ctl.ng;
    sc->nd = s->ctl.nd;
    sc->nw = s->ctl.nw;
    sc->ncl = s->ctl.ncl;
    sc->nsf = s->ctl.nsf;
    sc->retp_zmin = sc->retp_zmax = sc->rett_zmin = sc->rett_zmax = -999;
    for (int ig = 0; ig < NG; ig++) sc->retq_zmin[ig] = sc->retq_zmax[ig] = -999;
    for (int iw = 0; iw < NW; iw++) sc->retk_zmin[iw] = sc->retk_zmax[iw] = -999;
    sc->ret_clz = sc->ret_cldz = sc->ret_clk = sc->ret_sfz = sc->ret_sfp = sc->ret_sft = sc->ret_sfeps = 0;
  }
  if (req->kind != JUR_REQ_FORWARD) {
    // state windows of the request
    sc->retp_zmin = req->retp_zmin;
    sc->retp_zmax = req->retp_zmax;
    sc->rett_zmin = req->rett_zmin;
    sc->rett_zmax = req->rett_zmax;
    for (int ig = 0; ig < NG; ig++) {
      sc->retq_zmin[ig] = req->retq_zmin[ig];
      sc->retq_zmax[ig] = req->retq_zmax[ig];
    }
    for (int iw = 0; iw < NW; iw++) {
      sc->retk_zmin[iw] = req->retk_zmin[iw];
      sc->retk_zmax[iw] = req->retk_zmax[iw];
    }
    sc->ret_clz = sc->ret_cldz = sc->ret_clk = sc->ret_sfz = sc->ret_sfp = sc->ret_sft = sc->ret_sfeps = 0;
  }
  J->sctl = sc;
  J->own_apr.reset(new atm_t);
  memcpy(J->own_apr.get(), req->atm, sizeof(atm_t));
  J->own_atm_i.reset(new atm_t);
  J->own_meas.reset(new obs_t);
  copy_obs_rows(J->own_meas.get(), req->obs, s->ctl.nd);
  J->own_obs_i.reset(new obs_t);
  J->atm_apr = J->own_apr.get();
  J->atm_i = J->own_atm_i.get();
  J->obs_meas = J->own_meas.get();
  J->obs_i = J->own_obs_i.get();
  if (req->kind == JUR_REQ_JACOBIAN) {
    const size_t n = atm2x_h(sc, J->atm_apr, nullptr, nullptr, nullptr);
    const size_t m = obs2y_h(sc, J->obs_meas, nullptr);
    J->own_k.assign(std::max<size_t>(n * m, 1), 0.0);
    J->kout = J->own_k.data();
  }
  if (req->kind == JUR_REQ_RETRIEVAL) {
    J->own_hist.assign(s->ret.conv_itmax + 1, 0.0);
    J->hist = J->own_hist.data();
    J->niter = &J->own_niter;
  }
  J->t_submit = t_submit;
  for (int ir = 0; ir < J->obs_meas->nr && !J->limb; ir++) J->limb = ray_is_limb(J->obs_meas, ir);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) return set_err(s->fatal, "server stopped: %s", s->fatal_msg);
    J->id = s->next_id++;
    *id = J->id;
    J->st = ST_EVAL_FK0;
    Job *raw = J.get();
    s->all[J->id] = std::move(J);
    s->queue.push_back(raw);
  }
  // every lane: the one a single wakeup reaches may not take this kind
  s->cv_in.notify_all();
  return JUR_OK;
}

int jur_server_wait(jur_server_t *s, uint64_t id, double timeout_s, jur_result_t *res) {
  using namespace jur;
  if (!s || !res) return set_err(JUR_ERR_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return set_err(JUR_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  Job *J = it->second.get();
  // t_done is written by the worker under the lock when the job is finished
  auto ready = [&] { return J->t_done > 0 || s->fatal; };
  if (timeout_s < 0) s->cv_out.wait(lk, ready);
  else if (!s->cv_out.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return set_err(JUR_ERR_TIMEOUT, "request %llu not finished", (unsigned long long)id);
  if (!(J->t_done > 0)) return set_err(s->fatal, "server stopped: %s", s->fatal_msg);
  memset(res, 0, sizeof *res);
  res->id = id;
  res->kind = J->kind;
  res->status = J->rc;
  snprintf(res->message, sizeof res->message, "%s", J->rc ? J->err : "");
  res->t_submit = J->t_submit;
  res->t_start = J->t_start;
  res->t_done = J->t_done;
  res->n = J->n;
  res->m = J->mv;
  res->niter = J->own_niter;
  return JUR_OK;
}

int jur_server_fetch(jur_server_t *s, uint64_t id, obs_t *obs, atm_t *atm, double *k, double *chisq_hist) {
  using namespace jur;
  if (!s) return set_err(JUR_ERR_ARG, "NULL argument");
  std::unique_ptr<Job> J;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return set_err(JUR_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    if (!(it->second->t_done > 0)) return set_err(JUR_ERR_ARG, "request not finished");
    J = std::move(it->second);
    s->all.erase(it);
  }
  if (obs) copy_obs_rows(obs, J->obs_i, s->ctl.nd);
  if (atm) memcpy(atm, J->kind == JUR_REQ_RETRIEVAL ? J->atm_i : J->atm_apr, sizeof(atm_t));
  if (k && J->kind == JUR_REQ_JACOBIAN) memcpy(k, J->own_k.data(), (size_t)J->n * J->mv * sizeof(double));
  if (chisq_hist && J->kind == JUR_REQ_RETRIEVAL)
    memcpy(chisq_hist, J->own_hist.data(), (J->own_niter + 1) * sizeof(double));
  return J->rc ? set_err(J->rc, "%s", J->err) : JUR_OK;
}

int jur_server_stats(const jur_server_t *s, long *rounds, long *max_active) {
  if (!s) return jur::set_err(JUR_ERR_ARG, "NULL argument");
  if (rounds) *rounds = s->rounds;
  if (max_active) *max_active = s->max_active;
  return JUR_OK;
}

int jur_server_profile(const jur_server_t *s, double *sec, int n) {
  if (!s || !sec) return jur::set_err(JUR_ERR_ARG, "NULL argument");
  for (int i = 0; i < n && i < 4; i++) {
    sec[i] = 0;
    for (auto &L : s->lanes) sec[i] += L->E.prof[i];
  }
  if (n > 4) sec[4] = 1e-9 * (double)jur::g_alloc.ns.load();
  if (n > 5) sec[5] = (double)jur::g_alloc.calls.load();
  if (n > 6) sec[6] = 1e-9 * (double)jur::g_alloc.regrow_ns.load();
  return JUR_OK;
}

}  // extern "C"
