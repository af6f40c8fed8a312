// Levenberg-Marquardt optimal estimation for a batch of retrievals
// (optimal_estimation, src/retrieval.c:448-666, without the file output and
// the error analysis).
//
// Every retrieval follows the reference loop exactly (a priori covariance
// set_cov_apr :808-920, measurement errors set_cov_meas :924-964, cost
// :340-369, the damped normal equations solved by Cholesky, the state
// clamps and the convergence test).  The retrievals advance in lock-step
// rounds: each round runs one device pass over the rays of every retrieval
// that needs a forward model evaluation (and a second pass for those that
// need the Jacobian), then the dense algebra of all retrievals at the same
// stage in one launch per kernel (lm_kernels.hip).  So retrievals of
// different sizes and progress share the device passes.
#include "jur_host.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>

namespace jur {

struct LmDesc {
  int n, m;
  int ray0, nray;
  size_t o_nn, o_k, o_n, o_m;
};

void launch_lm_gather(const LmDesc *, const int *, int, int, const double *, int, int, const double *, double *,
                      hipStream_t);
void launch_lm_normal(const LmDesc *, const int *, int, int, const double *, double *, hipStream_t);
void launch_lm_cholinv(const LmDesc *, const int *, int, double *, double *, int *, hipStream_t);
void launch_lm_rhs(const LmDesc *, const int *, int, int, const double *, const double *, const double *,
                   const double *, const double *, double *, hipStream_t);
void launch_lm_solve(const LmDesc *, const int *, int, const double *, const double *, const double *,
                     const double *, double *, double *, int *, hipStream_t);
void launch_lm_quad(const LmDesc *, const int *, int, const double *, const double *, double *, hipStream_t);
int kernel_fd_host(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n, int precision);

namespace {

enum Stage { ST_EVAL_FK0, ST_ITER, ST_EVAL_FK, ST_TRIAL, ST_EVAL_F, ST_DONE };

struct Job {
  atm_t *atm_apr = nullptr, *atm_i = nullptr;
  obs_t *obs_meas = nullptr, *obs_i = nullptr;
  double *hist = nullptr;
  int *niter = nullptr;
  int n = 0, m = 0, mv = 0, nr = 0;
  std::vector<int> iqa, ipa, xmap;
  std::vector<double> x_a, x_i, y_m, sig, dx, dy, b, step;
  double chisq = 0, chisq_old = 0, lmpar = 0.001;
  int it = 0, it2 = 0;
  bool fk_done = false;  // the Jacobian of this outer iteration is in place
  Stage st = ST_EVAL_FK0;
  LmDesc d{};
};

#define POW2(x) ((x) * (x))

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

// The batch engine.
struct LmBatch {
  const jur_ret_t *ret;
  const ctl_t *ctl;
  jur_ctx *c;
  int method, precision;
  std::vector<Job> jobs;
  // device arenas
  DevBuf<LmDesc> d_desc;
  DevBuf<int> d_list, d_status;
  DevBuf<double> d_sainv, d_cov, d_a, d_k, d_sig, d_dy, d_dx, d_b, d_step, d_lmpar, d_q;
  // pass buffers
  DevBuf<atm_t> p_atm;
  DevBuf<int> p_prof, p_xmap;
  DevBuf<double> p_geom, p_rad, p_tau, p_k;
  DevBuf<unsigned char> p_mask;
  std::vector<double> h_rad;
  hipStream_t st = 0;

  int setup();
  int run();
  int eval(const std::vector<int> &ids, bool with_k);
  int chisq(const std::vector<int> &ids);
  int launch_list(const std::vector<int> &ids);
};

int LmBatch::launch_list(const std::vector<int> &ids) {
  if (ids.empty()) return JUR_OK;
  int rc = d_list.upload(ids.data(), ids.size());
  return rc;
}

int LmBatch::setup() {
  const int nd = ctl->nd;
  size_t o_nn = 0, o_k = 0, o_n = 0, o_m = 0;
  std::vector<LmDesc> desc(jobs.size());
  for (size_t r = 0; r < jobs.size(); r++) {
    Job &J = jobs[r];
    J.nr = J.obs_meas->nr;
    if (J.nr < 1 || J.nr > NR) return set_err(JUR_ERR_ARG, "retrieval %zu: Too many rays!", r);
    if (J.atm_apr->np < 2 || J.atm_apr->np > NP) return set_err(JUR_ERR_ARG, "retrieval %zu: Too many data points!", r);
    J.m = J.nr * nd;
    J.mv = (int)obs2y_h(ctl, J.obs_meas, nullptr);
    J.n = (int)atm2x_h(ctl, J.atm_apr, nullptr, nullptr, nullptr);
    if (J.mv == 0 || J.n == 0) return set_err(JUR_ERR_ARG, "retrieval %zu: Check problem definition!", r);
    J.iqa.resize(J.n);
    J.ipa.resize(J.n);
    J.x_a.resize(J.n);
    atm2x_h(ctl, J.atm_apr, J.x_a.data(), J.iqa.data(), J.ipa.data());
    J.x_i = J.x_a;
    J.xmap.resize(JUR_XMAP_LEN);
    state_map_h(ctl, J.atm_apr, J.xmap.data());
    if (method == JUR_JAC_ADJOINT && (J.xmap[JUR_XMAP_SFZ] >= 0 || J.xmap[JUR_XMAP_SFP] >= 0))
      return set_err(JUR_ERR_UNSUPPORTED, "retrieval of surface altitude/pressure needs JUR_JAC_FD");
    // measurement vector and errors over all (ray, channel) rows; masked
    // rows carry sig = 0 (set_cov_meas, src/retrieval.c:924-964)
    J.y_m.assign(J.m, 0.0);
    J.sig.assign(J.m, 0.0);
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const double v = J.obs_meas->rad[id][ir];
        if (!isfinite(v)) continue;
        const int i = ir * nd + id;
        J.y_m[i] = v;
        const double sn = ret->err_noise[id];
        const double sf = fabs(ret->err_formod[id] / 100 * v);
        J.sig[i] = 1 / sqrt(POW2(sn) + POW2(sf));
        if (J.sig[i] <= 0 || !isfinite(J.sig[i]))
          return set_err(JUR_ERR_ARG, "retrieval %zu: Check measurement errors (zero standard deviation)!", r);
      }
    J.dx.assign(J.n, 0.0);
    J.dy.assign(J.m, 0.0);
    J.b.assign(J.n, 0.0);
    J.step.assign(J.n, 0.0);
    memcpy(J.atm_i, J.atm_apr, sizeof(atm_t));
    memcpy(J.obs_i, J.obs_meas, sizeof(obs_t));
    J.d = LmDesc{J.n, J.m, 0, J.nr, o_nn, o_k, o_n, o_m};
    desc[r] = J.d;
    o_nn += (size_t)J.n * J.n;
    o_k += (size_t)J.m * J.n;
    o_n += J.n;
    o_m += J.m;
  }
  int rc;
  const size_t nj = jobs.size();
  if ((rc = d_desc.upload(desc.data(), nj))) return rc;
  if ((rc = d_status.reserve(nj))) return rc;
  if ((rc = d_sainv.reserve(o_nn))) return rc;
  if ((rc = d_cov.reserve(o_nn))) return rc;
  if ((rc = d_a.reserve(o_nn))) return rc;
  if ((rc = d_k.reserve(o_k))) return rc;
  if ((rc = d_sig.reserve(o_m))) return rc;
  if ((rc = d_dy.reserve(o_m))) return rc;
  if ((rc = d_dx.reserve(o_n))) return rc;
  if ((rc = d_b.reserve(o_n))) return rc;
  if ((rc = d_step.reserve(o_n))) return rc;
  if ((rc = d_lmpar.reserve(nj))) return rc;
  if ((rc = d_q.reserve(nj))) return rc;
  // S_a (host) -> S_a^-1 (device), measurement errors
  std::vector<double> sa, sig_all(o_m);
  std::vector<int> ids(nj);
  for (size_t r = 0; r < nj; r++) {
    Job &J = jobs[r];
    if ((rc = cov_apr(ret, ctl, J.atm_apr, J, sa))) return rc;
    HIPCHK(hipMemcpy(d_a.p + J.d.o_nn, sa.data(), sa.size() * sizeof(double), hipMemcpyHostToDevice));
    memcpy(&sig_all[J.d.o_m], J.sig.data(), J.m * sizeof(double));
    ids[r] = (int)r;
  }
  HIPCHK(hipMemcpy(d_sig.p, sig_all.data(), o_m * sizeof(double), hipMemcpyHostToDevice));
  if ((rc = launch_list(ids))) return rc;
  launch_lm_cholinv(d_desc.p, d_list.p, (int)nj, d_a.p, d_sainv.p, d_status.p, st);
  HIPCHK(hipGetLastError());
  std::vector<int> status(nj);
  HIPCHK(hipMemcpy(status.data(), d_status.p, nj * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t r = 0; r < nj; r++)
    if (status[r]) return set_err(JUR_ERR_ARG, "retrieval %zu: a priori covariance is not positive definite", r);
  return JUR_OK;
}

// One forward (and, with with_k, Jacobian) evaluation of the retrievals ids
// at their current atm_i: fills obs_i, y (into dy's source), and the
// compact scaled K in the device arena.
int LmBatch::eval(const std::vector<int> &ids, bool with_k) {
  if (ids.empty()) return JUR_OK;
  const int nd = ctl->nd;
  int rc;
  if (with_k && method == JUR_JAC_FD) {
    // kernel() per retrieval (n + 1 forward models in one device pass each)
    for (int r : ids) {
      Job &J = jobs[r];
      memcpy(J.obs_i, J.obs_meas, sizeof(obs_t));
      std::vector<double> kh((size_t)J.mv * J.n + 1), kc((size_t)J.m * J.n, 0.0);
      if ((rc = kernel_fd_host(ctl, J.atm_i, J.obs_i, kh.data(), J.mv, J.n, precision))) return rc;
      size_t v = 0;
      for (int i = 0; i < J.m; i++)
        if (J.sig[i] != 0) {
          for (int j = 0; j < J.n; j++) kc[(size_t)i * J.n + j] = kh[v * J.n + j] * J.sig[i];
          v++;
        }
      HIPCHK(hipMemcpy(d_k.p + J.d.o_k, kc.data(), kc.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return JUR_OK;
  }
  // one device pass over the rays of all retrievals in ids
  int nray = 0, maxn = 1;
  for (int r : ids) {
    jobs[r].d.ray0 = nray;
    nray += jobs[r].nr;
    maxn = std::max(maxn, jobs[r].n);
  }
  const int np = (int)ids.size();
  std::vector<atm_t> atms(np);
  std::vector<int> prof(nray), xm;
  std::vector<double> geom((size_t)7 * nray);
  std::vector<unsigned char> mask((size_t)nray * nd);
  for (int k = 0; k < np; k++) {
    Job &J = jobs[ids[k]];
    memcpy(&atms[k], J.atm_i, sizeof(atm_t));
    for (int ir = 0; ir < J.nr; ir++) {
      const int i = J.d.ray0 + ir;
      prof[i] = k;
      double *g = &geom[(size_t)7 * i];
      g[0] = J.obs_meas->time[ir];
      g[1] = J.obs_meas->obsz[ir];
      g[2] = J.obs_meas->obslon[ir];
      g[3] = J.obs_meas->obslat[ir];
      g[4] = J.obs_meas->vpz[ir];
      g[5] = J.obs_meas->vplon[ir];
      g[6] = J.obs_meas->vplat[ir];
      for (int id = 0; id < nd; id++) mask[(size_t)i * nd + id] = J.sig[ir * nd + id] == 0;
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
    std::vector<int> pmap(np);
    for (int k = 0; k < np; k++) pmap[k] = k;
    DevBuf<int> d_pmap;
    if ((rc = d_pmap.upload(pmap.data(), np))) return rc;
    if ((rc = p_xmap.upload(xm.data(), xm.size()))) return rc;
    if ((rc = p_k.reserve((size_t)nray * nd * maxn))) return rc;
    JacArgs ja;
    ja.xmap = p_xmap.p;
    ja.nmap = np;
    ja.prof_map = d_pmap.p;
    ja.ncol = maxn;
    ja.k = p_k.p;
    if ((rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st, &ja))) return rc;
    // rays of each retrieval are contiguous: gather its rows
    std::vector<LmDesc> desc(jobs.size());
    for (size_t r = 0; r < jobs.size(); r++) desc[r] = jobs[r].d;
    if ((rc = d_desc.upload(desc.data(), desc.size()))) return rc;
    if ((rc = launch_list(ids))) return rc;
    int maxrows = 0;
    for (int r : ids) maxrows = std::max(maxrows, jobs[r].m);
    launch_lm_gather(d_desc.p, d_list.p, np, maxrows, p_k.p, nd, maxn, d_sig.p, d_k.p, st);
    HIPCHK(hipGetLastError());
  } else {
    if ((rc = forward_pass(c, B, precision, p_mask.p, p_rad.p, p_tau.p, nullptr, nullptr, st))) return rc;
  }
  h_rad.resize((size_t)nray * nd * 2);
  HIPCHK(hipMemcpy(h_rad.data(), p_rad.p, (size_t)nray * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(h_rad.data() + (size_t)nray * nd, p_tau.p, (size_t)nray * nd * sizeof(double),
                   hipMemcpyDeviceToHost));
  for (int r : ids) {
    Job &J = jobs[r];
    memcpy(J.obs_i, J.obs_meas, sizeof(obs_t));
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const size_t i = (size_t)(J.d.ray0 + ir) * nd + id;
        J.obs_i->rad[id][ir] = h_rad[i];
        J.obs_i->tau[id][ir] = h_rad[(size_t)nray * nd + i];
      }
  }
  return JUR_OK;
}

// dx, dy and the cost of the retrievals ids at their current x_i / obs_i
// (cost_function, src/retrieval.c:340-369).
int LmBatch::chisq(const std::vector<int> &ids) {
  if (ids.empty()) return JUR_OK;
  const int nd = ctl->nd;
  int rc;
  for (int r : ids) {
    Job &J = jobs[r];
    for (int i = 0; i < J.n; i++) J.dx[i] = J.x_i[i] - J.x_a[i];
    for (int ir = 0; ir < J.nr; ir++)
      for (int id = 0; id < nd; id++) {
        const int i = ir * nd + id;
        J.dy[i] = J.sig[i] != 0 ? J.y_m[i] - J.obs_i->rad[id][ir] : 0.0;
      }
    HIPCHK(hipMemcpy(d_dx.p + J.d.o_n, J.dx.data(), J.n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_dy.p + J.d.o_m, J.dy.data(), J.m * sizeof(double), hipMemcpyHostToDevice));
  }
  if ((rc = launch_list(ids))) return rc;
  launch_lm_quad(d_desc.p, d_list.p, (int)ids.size(), d_sainv.p, d_dx.p, d_q.p, st);
  HIPCHK(hipGetLastError());
  std::vector<double> q(jobs.size());
  HIPCHK(hipMemcpy(q.data(), d_q.p, jobs.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int r : ids) {
    Job &J = jobs[r];
    double cm = 0;
    for (int i = 0; i < J.m; i++) cm += POW2(J.dy[i] * J.sig[i]);
    J.chisq = (cm + q[r]) / (double)J.mv;
  }
  return JUR_OK;
}

int LmBatch::run() {
  int rc = setup();
  if (rc) return rc;
  const int kr = std::max(ret->kernel_recomp, 1);
  const size_t nj = jobs.size();
  std::vector<double> lm(nj);
  for (;;) {
    // ---- model evaluations ----
    std::vector<int> need_f, need_fk;
    for (size_t r = 0; r < nj; r++) {
      if (jobs[r].st == ST_EVAL_F) need_f.push_back((int)r);
      if (jobs[r].st == ST_EVAL_FK0 || jobs[r].st == ST_EVAL_FK) need_fk.push_back((int)r);
    }
    if ((rc = eval(need_fk, true))) return rc;
    if ((rc = eval(need_f, false))) return rc;
    // first evaluation: cost of the a priori
    std::vector<int> first;
    for (int r : need_fk)
      if (jobs[r].st == ST_EVAL_FK0) first.push_back(r);
    if ((rc = chisq(first))) return rc;
    for (int r : first) {
      Job &J = jobs[r];
      if (J.hist) J.hist[0] = J.chisq;
      J.it = 1;
      J.fk_done = true;
      J.st = ST_ITER;
      if (ret->conv_itmax < 1) {
        J.it = ret->conv_itmax;
        J.st = ST_DONE;
      }
    }
    for (int r : need_fk)
      if (jobs[r].st == ST_EVAL_FK) {
        jobs[r].fk_done = true;
        jobs[r].st = ST_ITER;
      }
    // trial evaluations: accept or reject the step
    if ((rc = chisq(need_f))) return rc;
    for (int r : need_f) {
      Job &J = jobs[r];
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
        } else if (++J.it > ret->conv_itmax) {
          J.it = ret->conv_itmax;
          J.st = ST_DONE;
        } else {
          J.fk_done = false;
          J.st = ST_ITER;
        }
      }
    }
    // ---- start of outer iterations: Jacobian refresh, normal equations ----
    std::vector<int> need_cov, need_rhs;
    for (size_t r = 0; r < nj; r++) {
      Job &J = jobs[r];
      if (J.st != ST_ITER) continue;
      if (J.it > 1 && J.it % kr == 0 && !J.fk_done) {
        J.st = ST_EVAL_FK;
        continue;
      }
      J.chisq_old = J.chisq;
      if (J.it == 1 || J.it % kr == 0) need_cov.push_back((int)r);
      need_rhs.push_back((int)r);
      J.it2 = 0;
      J.st = ST_TRIAL;
    }
    if (!need_cov.empty()) {
      int maxn = 1;
      for (int r : need_cov) maxn = std::max(maxn, jobs[r].n);
      if ((rc = launch_list(need_cov))) return rc;
      launch_lm_normal(d_desc.p, d_list.p, (int)need_cov.size(), maxn, d_k.p, d_cov.p, st);
      HIPCHK(hipGetLastError());
    }
    if (!need_rhs.empty()) {
      int maxn = 1;
      for (int r : need_rhs) maxn = std::max(maxn, jobs[r].n);
      if ((rc = launch_list(need_rhs))) return rc;
      launch_lm_rhs(d_desc.p, d_list.p, (int)need_rhs.size(), maxn, d_k.p, d_sig.p, d_dy.p, d_sainv.p, d_dx.p,
                    d_b.p, st);
      HIPCHK(hipGetLastError());
      std::vector<double> bb(d_b.n);
      HIPCHK(hipMemcpy(bb.data(), d_b.p, d_b.n * sizeof(double), hipMemcpyDeviceToHost));
      for (int r : need_rhs) memcpy(jobs[r].b.data(), &bb[jobs[r].d.o_n], jobs[r].n * sizeof(double));
    }
    // ---- trial steps ----
    std::vector<int> trial;
    for (size_t r = 0; r < nj; r++)
      if (jobs[r].st == ST_TRIAL) trial.push_back((int)r);
    if (!trial.empty()) {
      for (size_t r = 0; r < nj; r++) lm[r] = jobs[r].lmpar;
      if ((rc = d_lmpar.upload(lm.data(), nj))) return rc;
      if ((rc = launch_list(trial))) return rc;
      launch_lm_solve(d_desc.p, d_list.p, (int)trial.size(), d_sainv.p, d_cov.p, d_lmpar.p, d_b.p, d_a.p, d_step.p,
                      d_status.p, st);
      HIPCHK(hipGetLastError());
      std::vector<double> s(d_step.n);
      std::vector<int> status(nj);
      HIPCHK(hipMemcpy(s.data(), d_step.p, d_step.n * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(status.data(), d_status.p, nj * sizeof(int), hipMemcpyDeviceToHost));
      for (int r : trial) {
        Job &J = jobs[r];
        if (status[r]) return set_err(JUR_ERR_ARG, "retrieval %d: normal matrix is not positive definite", r);
        memcpy(J.step.data(), &s[J.d.o_n], J.n * sizeof(double));
        for (int i = 0; i < J.n; i++) J.x_i[i] += J.step[i];
        memcpy(J.atm_i, J.atm_apr, sizeof(atm_t));
        x2atm_h(ctl, J.x_i.data(), J.atm_i);
        clamp_atm(ctl, J.atm_i);
        J.st = ST_EVAL_F;
      }
    }
    bool all_done = true;
    for (size_t r = 0; r < nj; r++) all_done &= jobs[r].st == ST_DONE;
    if (all_done) break;
  }
  for (size_t r = 0; r < nj; r++)
    if (jobs[r].niter) *jobs[r].niter = jobs[r].it;
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
  LmBatch L;
  L.ret = ret;
  L.ctl = ctl;
  L.method = method;
  L.precision = precision;
  int rc = ctx_for(ctl, &L.c);
  if (rc) return rc;
  if (method == JUR_JAC_ADJOINT && (rc = check_jac_ctl(L.c))) return rc;
  if (L.c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "optimal estimation with a FOV is not available");
  L.jobs.resize(n);
  for (int r = 0; r < n; r++) {
    Job &J = L.jobs[r];
    J.atm_apr = &atm_apr[r];
    J.atm_i = &atm_i[r];
    J.obs_meas = &obs_meas[r];
    J.obs_i = &obs_i[r];
    J.hist = chisq_hist ? chisq_hist + (size_t)r * (ret->conv_itmax + 1) : nullptr;
    J.niter = niter ? niter + r : nullptr;
  }
  return L.run();
}

}  // namespace jur

extern "C" int jur_gpu_optimal_estimation(const jur_ret_t *ret, const ctl_t *ctl, int n, obs_t *obs_meas,
                                          obs_t *obs_i, atm_t *atm_apr, atm_t *atm_i, double *chisq_hist,
                                          int *niter, int method, int precision) {
  return jur::optimal_estimation_batch(ret, ctl, n, obs_meas, obs_i, atm_apr, atm_i, chisq_hist, niter, method,
                                       precision);
}
