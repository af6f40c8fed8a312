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

struct LmDesc {
  int n, m;
  int ray0, nray;
  double *sainv, *cmat, *a;
  double *k;
  double *sig, *dy;
  double *dx, *b, *step;
};

void launch_lm_gather(const LmDesc *, const int *, int, int, const double *, int, int, hipStream_t);
void launch_lm_normal(const LmDesc *, const int *, int, int, hipStream_t);
void launch_lm_cholinv(const LmDesc *, const int *, int, int *, hipStream_t);
void launch_lm_rhs(const LmDesc *, const int *, int, int, hipStream_t);
void launch_lm_solve(const LmDesc *, const int *, int, const double *, int *, hipStream_t);
void launch_lm_quad(const LmDesc *, const int *, int, double *, hipStream_t);
int kernel_fd_host(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n, int precision);

#define POW2(x) ((x) * (x))

namespace {

enum Stage { ST_EVAL_F_OUT, ST_EVAL_FK_OUT, ST_EVAL_FK0, ST_ITER, ST_EVAL_FK, ST_TRIAL, ST_EVAL_F, ST_DONE };

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
  // device buffers of a retrieval
  DevBuf<double> d_sainv, d_cmat, d_a, d_k, d_sig, d_dy, d_dx, d_b, d_step;
  int ray0 = 0;
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

  LmDesc desc() {
    return LmDesc{n, m, ray0, nr, d_sainv.p, d_cmat.p, d_a.p, d_k.p, d_sig.p, d_dy.p, d_dx.p, d_b.p, d_step.p};
  }
  bool failed() const { return rc != JUR_OK; }
};

// set_cov_apr (src/retrieval.c:808-920).
int cov_apr(const jur_ret_t *ret, const ctl_t *ctl, const atm_t *atm, const Job &J, std::vector<double> &sa) {
  const size_t n = J.n;
  std::vector<double> xa(n + 1);
  atm2x_h(ctl, atm, xa.data(), nullptr, nullptr);
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl, nsf = ctl->nsf;
  for (size_t i = 0; i < n; i++) {
    const int q = J.iqa[i];
    if (q == IDXP) xa[i] = ret->err_press / 100 * xa[i];
    if (q == IDXT) xa[i] = ret->err_temp;
    for (int ig = 0; ig < ng; ig++)
      if (q == 2 + ig) xa[i] = ret->err_q[ig] / 100 * xa[i];
    for (int iw = 0; iw < nw; iw++)
      if (q == 2 + ng + iw) xa[i] = ret->err_k[iw];
    if (q == 2 + ng + nw) xa[i] = ret->err_clz;
    if (q == 3 + ng + nw) xa[i] = ret->err_cldz;
    for (int icl = 0; icl < ncl; icl++)
      if (q == 4 + ng + nw + icl) xa[i] = ret->err_clk[icl];
    if (q == 4 + ng + nw + ncl) xa[i] = ret->err_sfz;
    if (q == 5 + ng + nw + ncl) xa[i] = ret->err_sfp;
    if (q == 6 + ng + nw + ncl) xa[i] = ret->err_sft;
    for (int isf = 0; isf < nsf; isf++)
      if (q == 7 + ng + nw + ncl + isf) xa[i] = ret->err_sfeps[isf];
  }
  for (size_t i = 0; i < n; i++)
    if (POW2(xa[i]) <= 0) return set_err(JUR_ERR_ARG, "Check a priori data (zero standard deviation)!");
  sa.assign(n * n, 0.0);
  for (size_t i = 0; i < n; i++) sa[i * n + i] = POW2(xa[i]);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++)
      if (i != j && J.iqa[i] == J.iqa[j]) {
        const int q = J.iqa[i];
        double cz = 0, ch = 0;
        if (q == IDXP) { cz = ret->err_press_cz; ch = ret->err_press_ch; }
        if (q == IDXT) { cz = ret->err_temp_cz; ch = ret->err_temp_ch; }
        for (int ig = 0; ig < ng; ig++)
          if (q == 2 + ig) { cz = ret->err_q_cz[ig]; ch = ret->err_q_ch[ig]; }
        for (int iw = 0; iw < nw; iw++)
          if (q == 2 + ng + iw) { cz = ret->err_k_cz[iw]; ch = ret->err_k_ch[iw]; }
        if (cz > 0 && ch > 0) {
          double x0[3], x1[3];
          auto geo = [](double lon, double lat, double *x) {
            const double latr = lat / 180. * JUR_PI, lonr = lon / 180. * JUR_PI;
            x[0] = JUR_RE * cos(latr) * cos(lonr);
            x[1] = JUR_RE * cos(latr) * sin(lonr);
            x[2] = JUR_RE * sin(latr);
          };
          geo(atm->lon[J.ipa[i]], atm->lat[J.ipa[i]], x0);
          geo(atm->lon[J.ipa[j]], atm->lat[J.ipa[j]], x1);
          const double dist = sqrt(POW2(x0[0] - x1[0]) + POW2(x0[1] - x1[1]) + POW2(x0[2] - x1[2]));
          const double rho = exp(-dist / ch - fabs(atm->z[J.ipa[i]] - atm->z[J.ipa[j]]) / cz);
          sa[i * n + j] = xa[i] * xa[j] * rho;
        }
      }
  return JUR_OK;
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
  std::vector<Job *> jobs;  // active
  DevBuf<LmDesc> d_desc;
  DevBuf<int> d_list, d_status;
  DevBuf<double> d_lmpar, d_q;
  DevBuf<atm_t> p_atm;
  DevBuf<int> p_prof, p_xmap, p_pmap;
  DevBuf<double> p_geom, p_rad, p_tau, p_k;
  DevBuf<unsigned char> p_mask;
  std::vector<double> h_buf;
  hipStream_t st = 0;

  // Adds a job; on error the job is marked failed (and not added).
  int add(Job *J);
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
  int list(const std::vector<int> &ids) { return ids.empty() ? JUR_OK : d_list.upload(ids.data(), ids.size()); }
  int eval(const std::vector<int> &ids, bool with_k);
  int chisq(const std::vector<int> &ids);
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
  memcpy(J->obs_i, J->obs_meas, sizeof(obs_t));
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
  const size_t nn = (size_t)J->n * J->n;
  int rc;
  if ((rc = J->d_sainv.reserve(nn)) || (rc = J->d_cmat.reserve(nn)) || (rc = J->d_a.reserve(nn)) ||
      (rc = J->d_k.reserve((size_t)J->m * J->n)) || (rc = J->d_sig.upload(J->sig.data(), J->m)) ||
      (rc = J->d_dy.reserve(J->m)) || (rc = J->d_dx.reserve(J->n)) || (rc = J->d_b.reserve(J->n)) ||
      (rc = J->d_step.reserve(J->n)))
    return fail(J, rc);
  J->bytes = (3 * nn + (size_t)J->m * J->n + 2 * J->m + 3 * J->n) * 8;
  // S_a (host) -> S_a^-1 (device)
  std::vector<double> sa;
  if ((rc = cov_apr(ret, sc, J->atm_apr, *J, sa))) return fail(J, rc);
  if (hipMemcpy(J->d_a.p, sa.data(), nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(J, set_err(JUR_ERR_HIP, "hipMemcpy failed"));
  LmDesc d = J->desc();
  DevBuf<LmDesc> dd;
  const int zero = 0;
  DevBuf<int> dl;
  if ((rc = dd.upload(&d, 1)) || (rc = dl.upload(&zero, 1)) || (rc = d_status.reserve(1))) return fail(J, rc);
  launch_lm_cholinv(dd.p, dl.p, 1, d_status.p, st);
  int status = 0;
  if (hipMemcpy(&status, d_status.p, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(J, set_err(JUR_ERR_HIP, "hipMemcpy failed"));
  if (status) return fail(J, set_err(JUR_ERR_ARG, "a priori covariance is not positive definite"));
  J->st = ST_EVAL_FK0;
  jobs.push_back(J);
  return JUR_OK;
}

int Engine::upload_desc() {
  std::vector<LmDesc> desc(jobs.size());
  for (size_t r = 0; r < jobs.size(); r++) desc[r] = jobs[r]->desc();
  return d_desc.upload(desc.data(), std::max<size_t>(desc.size(), 1));
}

// One forward (and, with with_k, Jacobian) evaluation of the jobs ids at
// their current state: fills obs_i, and for retrievals the scaled compact K,
// for Jacobian requests the output K.
int Engine::eval(const std::vector<int> &ids, bool with_k) {
  if (ids.empty()) return JUR_OK;
  const int nd = ctl->nd;
  int rc;
  if (with_k && method == JUR_JAC_FD) {
    // kernel() per job (its n + 1 forward models in one device pass)
    for (int r : ids) {
      Job &J = *jobs[r];
      memcpy(J.obs_i, J.obs_meas, sizeof(obs_t));
      std::vector<double> kh((size_t)J.mv * J.n + 1);
      if ((rc = kernel_fd_host(J.sctl, J.atm_i, J.obs_i, kh.data(), J.mv, J.n, precision))) {
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
      HIPCHK(hipMemcpy(J.d_k.p, kc.data(), kc.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return JUR_OK;
  }
  // one device pass over the rays of all jobs in ids
  int nray = 0, maxn = 1;
  for (int r : ids) {
    jobs[r]->ray0 = nray;
    nray += jobs[r]->nr;
    maxn = std::max(maxn, jobs[r]->n);
  }
  const int np = (int)ids.size();
  std::vector<atm_t> atms(np);
  std::vector<int> prof(nray), xm, pmap(np);
  std::vector<double> geom((size_t)7 * nray);
  std::vector<unsigned char> mask((size_t)nray * nd);
  for (int k = 0; k < np; k++) {
    Job &J = *jobs[ids[k]];
    memcpy(&atms[k], J.kind == JUR_REQ_RETRIEVAL ? J.atm_i : J.atm_apr, sizeof(atm_t));
    pmap[k] = k;
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
    if (with_k) xm.insert(xm.end(), J.xmap.begin(), J.xmap.end());
  }
  if ((rc = p_atm.upload(atms.data(), np))) return rc;
  if ((rc = p_prof.upload(prof.data(), nray))) return rc;
  if ((rc = p_geom.upload(geom.data(), geom.size()))) return rc;
  if ((rc = p_mask.upload(mask.data(), mask.size()))) return rc;
  if ((rc = p_rad.reserve((size_t)nray * nd))) return rc;
  if ((rc = p_tau.reserve((size_t)nray * nd))) return rc;
  BatchView B{np, p_atm.p, nray, p_prof.p, p_geom.p};
  if (with_k) {
    if ((rc = p_pmap.upload(pmap.data(), np))) return rc;
    if ((rc = p_xmap.upload(xm.data(), xm.size()))) return rc;
    if ((rc = p_k.reserve((size_t)nray * nd * maxn))) return rc;
    JacArgs ja;
    ja.xmap = p_xmap.p;
    ja.nmap = np;
    ja.prof_map = p_pmap.p;
    ja.ncol = maxn;
    ja.k = p_k.p;
    rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st, &ja);
  } else {
    rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st);
  }
  if (rc) {
    // A ray that fails (observer below the surface, too many LOS points)
    // fails the pass: evaluate the jobs one by one so that only its own job
    // fails.
    if (ids.size() == 1) {
      fail(jobs[ids[0]], rc);
      return JUR_OK;
    }
    for (int r : ids)
      if ((rc = eval(std::vector<int>{r}, with_k))) return rc;
    return JUR_OK;
  }
  h_buf.resize((size_t)nray * nd * 2);
  HIPCHK(hipMemcpy(h_buf.data(), p_rad.p, (size_t)nray * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(h_buf.data() + (size_t)nray * nd, p_tau.p, (size_t)nray * nd * sizeof(double),
                   hipMemcpyDeviceToHost));
  for (int r : ids) {
    Job &J = *jobs[r];
    memcpy(J.obs_i, J.obs_meas, sizeof(obs_t));
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const size_t i = (size_t)(J.ray0 + ir) * nd + id;
        J.obs_i->rad[id][ir] = h_buf[i];
        J.obs_i->tau[id][ir] = h_buf[(size_t)nray * nd + i];
      }
  }
  if (!with_k) return JUR_OK;
  // Jacobian requests: copy their finite rows out
  std::vector<double> rows;
  for (int r : ids) {
    Job &J = *jobs[r];
    if (J.kind != JUR_REQ_JACOBIAN) continue;
    rows.resize((size_t)J.m * maxn);
    HIPCHK(hipMemcpy(rows.data(), p_k.p + (size_t)J.ray0 * nd * maxn, rows.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
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
  for (int r : ids)
    if (jobs[r]->kind == JUR_REQ_RETRIEVAL) {
      lm.push_back(r);
      maxrows = std::max(maxrows, jobs[r]->m);
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
  for (int r : live) {
    Job &J = *jobs[r];
    for (int i = 0; i < J.n; i++) J.dx[i] = J.x_i[i] - J.x_a[i];
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const int i = ir * nd + id;
        J.dy[i] = J.sig[i] != 0 ? J.y_m[i] - J.obs_i->rad[id][ir] : 0.0;
      }
    HIPCHK(hipMemcpy(J.d_dx.p, J.dx.data(), J.n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(J.d_dy.p, J.dy.data(), J.m * sizeof(double), hipMemcpyHostToDevice));
  }
  if ((rc = upload_desc())) return rc;
  if ((rc = list(live))) return rc;
  if ((rc = d_q.reserve(jobs.size()))) return rc;
  launch_lm_quad(d_desc.p, d_list.p, (int)live.size(), d_q.p, st);
  HIPCHK(hipGetLastError());
  std::vector<double> q(jobs.size());
  HIPCHK(hipMemcpy(q.data(), d_q.p, jobs.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int r : live) {
    Job &J = *jobs[r];
    double cm = 0;
    for (int i = 0; i < J.m; i++) cm += POW2(J.dy[i] * J.sig[i]);
    J.chisq = (cm + q[r]) / (double)J.mv;
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
    if (s == ST_EVAL_FK0 || s == ST_EVAL_FK || s == ST_EVAL_FK_OUT) need_fk.push_back((int)r);
  }
  if ((rc = eval(need_fk, true))) return rc;
  if ((rc = eval(need_f, false))) return rc;
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
  if ((rc = chisq(trials))) return rc;
  for (int r : trials) {
    Job &J = *jobs[r];
    if (J.failed()) continue;
    bool end_iter;
    if (J.chisq > J.chisq_old) {
      J.lmpar *= 10;
      for (int i = 0; i < J.n; i++) J.x_i[i] -= J.step[i];
      J.it2++;
      end_iter = J.it2 >= 20;
      if (!end_iter) J.st = ST_TRIAL;
    } else {
      J.lmpar /= 10;
      end_iter = true;
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
    for (int r : need_rhs)
      HIPCHK(hipMemcpy(jobs[r]->b.data(), jobs[r]->d_b.p, jobs[r]->n * sizeof(double), hipMemcpyDeviceToHost));
  }
  // ---- trial steps ----
  std::vector<int> trial;
  for (size_t r = 0; r < nj; r++)
    if (jobs[r]->st == ST_TRIAL) trial.push_back((int)r);
  if (!trial.empty()) {
    std::vector<double> lm(nj);
    for (size_t r = 0; r < nj; r++) lm[r] = jobs[r]->lmpar;
    if ((rc = d_lmpar.upload(lm.data(), nj))) return rc;
    if ((rc = d_status.reserve(nj))) return rc;
    if ((rc = upload_desc())) return rc;
    if ((rc = list(trial))) return rc;
    launch_lm_solve(d_desc.p, d_list.p, (int)trial.size(), d_lmpar.p, d_status.p, st);
    HIPCHK(hipGetLastError());
    std::vector<int> status(nj);
    HIPCHK(hipMemcpy(status.data(), d_status.p, nj * sizeof(int), hipMemcpyDeviceToHost));
    for (int r : trial) {
      Job &J = *jobs[r];
      if (status[r]) {
        fail(&J, set_err(JUR_ERR_ARG, "normal matrix is not positive definite"));
        continue;
      }
      HIPCHK(hipMemcpy(J.step.data(), J.d_step.p, J.n * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < J.n; i++) J.x_i[i] += J.step[i];
      memcpy(J.atm_i, J.atm_apr, sizeof(atm_t));
      x2atm_h(J.sctl, J.x_i.data(), J.atm_i);
      clamp_atm(J.sctl, J.atm_i);
      J.st = ST_EVAL_F;
    }
  }
  // ---- finished jobs leave ----
  std::vector<Job *> keep;
  for (Job *J : jobs) {
    if (J->st == ST_DONE) {
      if (J->niter) *J->niter = J->it;
      done.push_back(J);
    } else {
      keep.push_back(J);
    }
  }
  jobs.swap(keep);
  return JUR_OK;
}

int optimal_estimation_batch(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas, obs_t *obs_i,
                             atm_t *atm_apr, atm_t *atm_i, double *chisq_hist, int *niter, int method,
                             int precision) {
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
    if ((rc = E.add(&J))) return set_err(rc, "retrieval %d: %s", r, J.err);
  }
  std::vector<Job *> done;
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
struct jur_server {
  jur::Engine E;
  std::unique_ptr<jur_ctx> ctx;
  ctl_t ctl;
  jur_ret_t ret;
  size_t budget = 0;
  int max_batch = 0;
  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<jur::Job *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<jur::Job>> all;
  uint64_t next_id = 1;
  size_t in_use = 0;
  bool stop = false;
  int fatal = JUR_OK;
  char fatal_msg[256] = "";
  std::thread worker;
  // statistics
  std::atomic<long> rounds{0}, max_active{0};

  void loop();
};

void jur_server::loop() {
  using namespace jur;
  if (hipSetDevice(ctx->device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(mu);
    fatal = JUR_ERR_HIP;
    cv_out.notify_all();
    return;
  }
  for (;;) {
    std::vector<Job *> admit;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_in.wait(lk, [&] { return stop || !queue.empty() || !E.jobs.empty(); });
      if (stop) return;
      // admission: in arrival order while the device footprint fits
      while (!queue.empty()) {
        Job *J = queue.front();
        const size_t need = (size_t)J->obs_meas->nr * ctl.nd * 8 * 4 +
                            (J->kind == JUR_REQ_FORWARD ? 0 : (size_t)J->obs_meas->nr * ctl.nd * NP * 8 * 2);
        if (!E.jobs.empty() && (in_use + need > budget ||
                                (max_batch > 0 && (int)(E.jobs.size() + admit.size()) >= max_batch)))
          break;
        queue.pop_front();
        in_use += need;
        J->admit_bytes = need;
        admit.push_back(J);
      }
    }
    const double t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    std::vector<Job *> done;
    for (Job *J : admit) {
      J->t_start = t0;
      if (E.add(J)) done.push_back(J);  // failed at admission
    }
    int rc = E.jobs.empty() ? JUR_OK : E.round(done);
    rounds++;
    max_active = std::max<long>(max_active, (long)E.jobs.size() + (long)done.size());
    const double t1 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    std::lock_guard<std::mutex> lk(mu);
    if (rc) {
      fatal = rc;
      snprintf(fatal_msg, sizeof fatal_msg, "%s", jur::g_err);
      for (Job *J : E.jobs) {
        J->rc = rc;
        snprintf(J->err, sizeof J->err, "%s", jur::g_err);
        done.push_back(J);
      }
      E.jobs.clear();
    }
    for (Job *J : done) {
      // the job's device buffers go back now; results stay on the host
      J->d_sainv.release();
      J->d_cmat.release();
      J->d_a.release();
      J->d_k.release();
      J->d_sig.release();
      J->d_dy.release();
      J->d_dx.release();
      J->d_b.release();
      J->d_step.release();
      J->t_done = t1;
      in_use -= std::min(in_use, J->admit_bytes);
      J->st = ST_DONE;
    }
    if (!done.empty()) cv_out.notify_all();
  }
}

extern "C" {

int jur_gpu_optimal_estimation(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas, obs_t *obs_i,
                               atm_t *atm_apr, atm_t *atm_i, double *chisq_hist, int *niter, int method,
                               int precision) {
  return jur::optimal_estimation_batch(ret, ctl, n, obs_meas, obs_i, atm_apr, atm_i, chisq_hist, niter, method,
                                       precision);
}

int jur_server_create(const ctl_t *ctl, const jur_ret_t *ret, const jur_server_opts_t *opt, jur_server_t **out) {
  using namespace jur;
  if (!ctl || !ret || !out) return set_err(JUR_ERR_ARG, "NULL argument");
  std::unique_ptr<jur_server> s(new jur_server());
  memcpy(&s->ctl, ctl, sizeof(ctl_t));
  memcpy(&s->ret, ret, sizeof(jur_ret_t));
  jur_ctx_t *c = nullptr;
  int rc = jur_ctx_create(ctl, &c);
  if (rc) return rc;
  s->ctx.reset(c);
  s->E.ret = &s->ret;
  s->E.ctl = &s->ctl;
  s->E.c = c;
  s->E.method = opt ? opt->method : JUR_JAC_ADJOINT;
  s->E.precision = opt ? opt->precision : JUR_FP32;
  s->budget = (opt && opt->max_bytes) ? opt->max_bytes : ((size_t)16 << 30);
  s->max_batch = opt ? opt->max_batch : 0;
  if (s->E.method != JUR_JAC_FD && s->E.method != JUR_JAC_ADJOINT) return set_err(JUR_ERR_ARG, "invalid method");
  if (s->E.precision != JUR_FP64 && s->E.precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (s->E.method == JUR_JAC_ADJOINT && (rc = check_jac_ctl(c))) return rc;
  if (c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "the server does not apply a FOV");
  if (hipGetDevice(&c->device) != hipSuccess) return set_err(JUR_ERR_HIP, "hipGetDevice failed");
  jur_server *raw = s.get();
  raw->worker = std::thread([raw] { raw->loop(); });
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
  if (s->worker.joinable()) s->worker.join();
  delete s;
}

int jur_server_submit(jur_server_t *s, const jur_request_t *req, uint64_t *id) {
  using namespace jur;
  if (!s || !req || !id || !req->atm || !req->obs) return set_err(JUR_ERR_ARG, "NULL argument");
  if (req->kind != JUR_REQ_FORWARD && req->kind != JUR_REQ_JACOBIAN && req->kind != JUR_REQ_RETRIEVAL)
    return set_err(JUR_ERR_ARG, "invalid request kind");
  std::unique_ptr<Job> J(new Job());
  J->kind = req->kind;
  J->own_ctl.reset(new ctl_t);
  memcpy(J->own_ctl.get(), &s->ctl, sizeof(ctl_t));
  ctl_t *sc = J->own_ctl.get();
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
  memcpy(J->own_meas.get(), req->obs, sizeof(obs_t));
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
  J->t_submit = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
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
  s->cv_in.notify_one();
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
  if (obs) memcpy(obs, J->obs_i, sizeof(obs_t));
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

}  // extern "C"
