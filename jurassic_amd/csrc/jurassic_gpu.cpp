// Host side of libjurassic_gpu.so: table and filter ingest, the device
// context, ray batching and the C-ABI entry points declared in
// include/jurassic_gpu.h.
//
// The context replaces the reference's lazily initialised statics
// (formod_pencil, src/jurassic.c:3202-3217; formod_continua :3081;
// formod_fov :3125-3141), which bind to the first ctl seen: here a context is
// created per distinct ctl content and cached.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ctm_coeffs.h"
#include "jur_host.hpp"

#pragma clang fp contract(off)

namespace jur {

// kernel launchers (raytrace.hip, radiance_f64.hip, radiance_f32.hip, aux_kernels.hip)
size_t ray_tail_bytes();
void launch_raytrace(const ModelView &, const BatchView &, const LosView &, RayInfo *, void *, int, int,
                     hipStream_t);
void launch_hydrostatic(atm_t *, int, double, int, hipStream_t);
void launch_lookup_f64(const ModelView &, const BatchView &, const LosView &, const EpsView &,
                       const RayInfo *, int, int, bool, hipStream_t);
void launch_adjoint(const ModelView &, const BatchView &, const LosView &, const EpsView &,
                    const RayInfo *, const void *, int, int, int, const JacView &, bool, double *, double *,
                    hipStream_t);
void launch_jac_store(const double *, int, int, int, int, int, const double *, const ChanInfo *, int,
                      const unsigned char *, double *, hipStream_t);
void launch_lookup_f32(const ModelView &, const BatchView &, const LosView &, const EpsView &,
                       const RayInfo *, int, int, hipStream_t);
void launch_radiance_f64(const ModelView &, const BatchView &, const LosView &, const EpsView &,
                         const RayInfo *, int, int, double *, double *, hipStream_t);
void launch_radiance_f32(const ModelView &, const BatchView &, const LosView &, const EpsView &,
                         const RayInfo *, int, int, double *, double *, hipStream_t);
void launch_srcfunc(const double *, const double *, const double *, const int *, const int *,
                    const double *, int, double *, hipStream_t);
void launch_zrange(const atm_t *, int, double *, hipStream_t);
void launch_reduce_info(const RayInfo *, int, unsigned long long *, hipStream_t);
void launch_fov(const double *, const double *, double *, double *, int, int, const double *,
                const int *, const int *, const double *, const double *, int, int *, hipStream_t);
void launch_finalize(double *, int, int, const ChanInfo *, int, const unsigned char *,
                     hipStream_t);
void launch_block_np(const RayInfo *, int, int, int, int *, hipStream_t);
void launch_zero(void *, size_t, hipStream_t);
void launch_pack_tables(int, int, int, int, int, const int *, const int *, const int *, int, int, int,
                        const float *, const float *, float4 *, int *, hipStream_t);

// ------------------------------------------------------------------
// Errors.
// ------------------------------------------------------------------
thread_local char g_err[1024] = "";

int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static int setup_from_ctl(const ctl_t *ctl, int nd, const double *nu, const int *window,
                          Setup &s) {
  if (!ctl) return set_err(JUR_ERR_ARG, "ctl is NULL");
  if (ctl->ng < 0 || ctl->ng > NG) return set_err(JUR_ERR_ARG, "Set 0 <= NG <= MAX!");
  if (ctl->nw < 0 || ctl->nw > NW) return set_err(JUR_ERR_ARG, "Set 0 <= NW <= MAX!");
  if (ctl->ncl < 0 || ctl->ncl > NCL || ctl->ncl == 1) return set_err(JUR_ERR_ARG, "Set NCL > 1!");
  if (ctl->nsf < 0 || ctl->nsf > NSF || ctl->nsf == 1) return set_err(JUR_ERR_ARG, "Set NSF > 1!");
  if (nd < 0) return set_err(JUR_ERR_ARG, "Set 0 <= ND <= MAX!");
  if (ctl->formod == 2) return set_err(JUR_ERR_UNSUPPORTED, "FORMOD = 2 (RFM) is not supported");
  s.ng = ctl->ng;
  s.nd = nd;
  s.nw = ctl->nw;
  s.ncl = ctl->ncl;
  s.nsf = ctl->nsf;
  s.sftype = ctl->sftype;
  s.tblfmt = ctl->tblfmt;
  s.formod = ctl->formod;
  s.ctm_co2 = ctl->ctm_co2;
  s.ctm_h2o = ctl->ctm_h2o;
  s.ctm_n2 = ctl->ctm_n2;
  s.ctm_o2 = ctl->ctm_o2;
  s.refrac = ctl->refrac;
  s.write_bbt = ctl->write_bbt;
  s.sfsza = ctl->sfsza;
  s.hydz = ctl->hydz;
  s.rayds = ctl->rayds;
  s.raydz = ctl->raydz;
  for (int ig = 0; ig < ctl->ng; ig++) s.emitter.push_back(ctl->emitter[ig]);
  s.nu.assign(nu, nu + nd);
  s.window.resize(nd);
  for (int id = 0; id < nd; id++) s.window[id] = window ? window[id] : 0;
  s.clnu.assign(ctl->clnu, ctl->clnu + NCL);
  s.sfnu.assign(ctl->sfnu, ctl->sfnu + NSF);
  s.tblbase = ctl->tblbase;
  s.fov = ctl->fov;
  return JUR_OK;
}

// ------------------------------------------------------------------
// Table ingest: read_tbl (src/jurassic.c:4949-5110), one (channel, gas).
// ------------------------------------------------------------------
struct HostTable {
  int np = -1;
  std::vector<double> p;
  std::vector<int> nt;
  std::vector<std::vector<double>> t;
  std::vector<std::vector<int>> nu;
  std::vector<std::vector<std::vector<float>>> u, eps;
};

static int read_table(const Setup &s, int id, int ig, HostTable &tb) {
  char fn[2 * LEN + 64];
  snprintf(fn, sizeof fn, "%s_%.4f_%s.%s", s.tblbase.c_str(), s.nu[id], s.emitter[ig].c_str(),
           s.tblfmt == 1 ? "tab" : "bin");
  tb = HostTable();
  FILE *in = fopen(fn, "r");
  if (!in) {
    // the reference warns and continues: a missing table yields eps = 0
    return JUR_OK;
  }
  if (s.tblfmt == 1) {
    char line[LEN];
    double eps, press, temp, u;
    double eps_old = -999, press_old = -999, temp_old = -999, u_old = -999;
    int np = -1;
    std::vector<int> nt;
    std::vector<std::vector<int>> nuc;
    while (fgets(line, LEN, in)) {
      if (sscanf(line, "%lg %lg %lg %lg", &press, &temp, &u, &eps) != 4) continue;
      if (u < 0 || u > 1e30 || eps < 0 || eps > 1) continue;
      if (press != press_old) {
        press_old = press;
        if ((++np) >= TBLNP) { fclose(in); return set_err(JUR_ERR_ARG, "Too many pressure levels!"); }
        tb.p.push_back(press);
        nt.push_back(-1);
        tb.t.emplace_back();
        nuc.emplace_back();
        tb.u.emplace_back();
        tb.eps.emplace_back();
      }
      if (temp != temp_old) {
        temp_old = temp;
        if ((++nt[np]) >= TBLNT) { fclose(in); return set_err(JUR_ERR_ARG, "Too many temperatures!"); }
        tb.t[np].push_back(temp);
        nuc[np].push_back(-1);
        tb.u[np].emplace_back(TBLNU, 0.f);
        tb.eps[np].emplace_back(TBLNU, 0.f);
      }
      const int it = nt[np];
      if ((eps > eps_old && u > u_old) || nuc[np][it] < 0) {
        eps_old = eps;
        u_old = u;
        if ((++nuc[np][it]) >= TBLNU) { fclose(in); return set_err(JUR_ERR_ARG, "Too many column densities!"); }
      }
      tb.p[np] = press;
      tb.t[np][it] = temp;
      tb.u[np][it][nuc[np][it]] = (float)u;
      tb.eps[np][it][nuc[np][it]] = (float)eps;
    }
    tb.np = np + 1;
    tb.nt.resize(tb.np);
    tb.nu.resize(tb.np);
    for (int ip = 0; ip < tb.np; ip++) {
      tb.nt[ip] = nt[ip] + 1;
      tb.nu[ip].resize(tb.nt[ip]);
      for (int it = 0; it < tb.nt[ip]; it++) {
        tb.nu[ip][it] = nuc[ip][it] + 1;
        tb.u[ip][it].resize(tb.nu[ip][it]);
        tb.eps[ip][it].resize(tb.nu[ip][it]);
      }
    }
  } else if (s.tblfmt == 2) {
#define RD(ptr, sz, cnt)                                                                 \
  if (fread(ptr, sz, (size_t)(cnt), in) != (size_t)(cnt)) {                               \
    fclose(in);                                                                          \
    return set_err(JUR_ERR_IO, "Error while reading %s!", fn);                           \
  }
    RD(&tb.np, sizeof(int), 1);
    if (tb.np > TBLNP) { fclose(in); return set_err(JUR_ERR_ARG, "Too many pressure levels!"); }
    if (tb.np < 0) tb.np = 0;
    tb.p.resize(tb.np);
    RD(tb.p.data(), sizeof(double), tb.np);
    tb.nt.resize(tb.np);
    tb.t.resize(tb.np);
    tb.nu.resize(tb.np);
    tb.u.resize(tb.np);
    tb.eps.resize(tb.np);
    for (int ip = 0; ip < tb.np; ip++) {
      RD(&tb.nt[ip], sizeof(int), 1);
      if (tb.nt[ip] > TBLNT || tb.nt[ip] < 0) { fclose(in); return set_err(JUR_ERR_ARG, "Too many temperatures!"); }
      tb.t[ip].resize(tb.nt[ip]);
      RD(tb.t[ip].data(), sizeof(double), tb.nt[ip]);
      tb.nu[ip].resize(tb.nt[ip]);
      tb.u[ip].resize(tb.nt[ip]);
      tb.eps[ip].resize(tb.nt[ip]);
      for (int it = 0; it < tb.nt[ip]; it++) {
        RD(&tb.nu[ip][it], sizeof(int), 1);
        const int n = tb.nu[ip][it];
        if (n > TBLNU || n < 0) { fclose(in); return set_err(JUR_ERR_ARG, "Too many column densities!"); }
        tb.u[ip][it].resize(n);
        tb.eps[ip][it].resize(n);
        RD(tb.u[ip][it].data(), sizeof(float), n);
        RD(tb.eps[ip][it].data(), sizeof(float), n);
      }
    }
#undef RD
  } else {
    fclose(in);
    return set_err(JUR_ERR_ARG, "Unknown look-up table format!");
  }
  fclose(in);
  return JUR_OK;
}

// read_shape (src/jurassic.c:4907-4946).
static int read_shape(const char *fn, std::vector<double> &x, std::vector<double> &y) {
  FILE *in = fopen(fn, "r");
  if (!in) return set_err(JUR_ERR_IO, "Cannot open file %s!", fn);
  char line[LEN];
  double a, b;
  x.clear();
  y.clear();
  while (fgets(line, LEN, in))
    if (sscanf(line, "%lg %lg", &a, &b) == 2) {
      x.push_back(a);
      y.push_back(b);
      if ((int)x.size() > NSHAPE) { fclose(in); return set_err(JUR_ERR_ARG, "Too many data points!"); }
    }
  fclose(in);
  if (x.size() < 2) return set_err(JUR_ERR_IO, "Could not read any data from %s!", fn);
  return JUR_OK;
}

// locate_irr (src/jurassic.c:4103-4129) on the host, for per-channel indices.
static int locate_irr_h(const double *xx, int n, double x) {
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

}  // namespace jur


namespace jur {

// Shared grids: every table with np >= 30 has the same p grid, T-row sizes
// and T grids as the first such table (one cell search per segment then
// serves all of them).
static void detect_shared(jur_ctx *c, const std::vector<int> &h_np, const std::vector<int> &h_nt,
                          const std::vector<double> &h_p, const std::vector<double> &h_t) {
  const size_t ntab = (size_t)c->s.nd * c->s.ng;
  const int TP = c->TP, TT = c->TT;
  c->shared = 0;
  c->ref_tix = 0;
  int ref = -1;
  for (size_t k = 0; k < ntab; k++)
    if (h_np[k] >= 30) { ref = (int)k; break; }
  if (ref < 0) return;
  bool same = true;
  for (size_t k = 0; k < ntab && same; k++) {
    if (h_np[k] < 30) continue;
    if (h_np[k] != h_np[ref]) { same = false; break; }
    for (int ip = 0; ip < TP && same; ip++) {
      if (h_p[k * TP + ip] != h_p[ref * TP + ip] || h_nt[k * TP + ip] != h_nt[ref * TP + ip]) same = false;
      for (int it = 0; it < TT && same; it++)
        if (h_t[(k * TP + ip) * TT + it] != h_t[(ref * TP + ip) * TT + it]) same = false;
    }
  }
  c->shared = same ? 1 : 0;
  c->ref_tix = ref;
}

// Direction of a table's pressure grid: +1 strictly ascending, -1 strictly
// descending, 0 otherwise.
static int p_direction(const double *p, int np) {
  int asc = 1, desc = 1;
  for (int ip = 0; ip + 1 < np; ip++) {
    if (!(p[ip + 1] > p[ip])) asc = 0;
    if (!(p[ip + 1] < p[ip])) desc = 0;
  }
  return (np >= 2 && asc) ? 1 : ((np >= 2 && desc) ? -1 : 0);
}

static int build_tables(jur_ctx *c) {
  const Setup &s = c->s;
  const int nd = s.nd, ng = s.ng;
  std::vector<HostTable> tabs((size_t)nd * ng);
  int TP = 2, TT = 2, TU = 2;
  for (int ig = 0; ig < ng; ig++)
    for (int id = 0; id < nd; id++) {
      HostTable &tb = tabs[(size_t)ig * nd + id];
      int rc = read_table(s, id, ig, tb);
      if (rc) return rc;
      TP = std::max(TP, tb.np);
      for (int ip = 0; ip < tb.np; ip++) {
        TT = std::max(TT, tb.nt[ip]);
        for (int it = 0; it < tb.nt[ip]; it++) TU = std::max(TU, tb.nu[ip][it]);
      }
    }
  c->TP = TP;
  c->TT = TT;
  c->TU = TU;
  const size_t ntab = (size_t)nd * ng;
  std::vector<int> h_np(ntab + 1, -1), h_nt(ntab * TP + 1, 0), h_nu(ntab * TP * TT + 1, 0),
      h_mono(ntab + 1, 1), h_pdir(ntab + 1, 0);
  std::vector<double> h_p(ntab * TP + 1, 0.0), h_t(ntab * TP * TT + 1, 0.0);
  std::vector<float4> h_rec(ntab * TP * TT * TU + 1, make_float4(0, 0, 0, 0));
  c->all_mono = 1;
  for (size_t k = 0; k < ntab; k++) {
    const HostTable &tb = tabs[k];
    h_np[k] = tb.np;
    int mono = 1;
    for (int ip = 0; ip < tb.np; ip++) {
      h_p[k * TP + ip] = tb.p[ip];
      h_nt[k * TP + ip] = tb.nt[ip];
      for (int it = 0; it < tb.nt[ip]; it++) {
        h_t[(k * TP + ip) * TT + it] = tb.t[ip][it];
        const int n = tb.nu[ip][it];
        h_nu[(k * TP + ip) * TT + it] = n;
        float4 *col = &h_rec[((k * TP + ip) * TT + it) * TU];
        for (int i = 0; i < n; i++) {
          const int j = (i + 1 < n) ? i + 1 : i;
          col[i] = make_float4(tb.u[ip][it][i], tb.eps[ip][it][i], tb.u[ip][it][j], tb.eps[ip][it][j]);
          if (i + 1 < n && (!(tb.u[ip][it][j] >= tb.u[ip][it][i]) ||
                            !(tb.eps[ip][it][j] >= tb.eps[ip][it][i])))
            mono = 0;
        }
      }
    }
    h_mono[k] = mono;
    if (tb.np >= 30 && !mono) c->all_mono = 0;
    h_pdir[k] = p_direction(tb.p.data(), tb.np);
  }
  int rc;
  if ((rc = c->d_np.upload(h_np.data(), h_np.size()))) return rc;
  if ((rc = c->d_nt.upload(h_nt.data(), h_nt.size()))) return rc;
  if ((rc = c->d_nu.upload(h_nu.data(), h_nu.size()))) return rc;
  if ((rc = c->d_mono.upload(h_mono.data(), h_mono.size()))) return rc;
  if ((rc = c->d_pdir.upload(h_pdir.data(), h_pdir.size()))) return rc;
  if ((rc = c->d_p.upload(h_p.data(), h_p.size()))) return rc;
  if ((rc = c->d_t.upload(h_t.data(), h_t.size()))) return rc;
  if ((rc = c->d_rec.upload(h_rec.data(), h_rec.size()))) return rc;
  c->table_bytes = h_rec.size() * sizeof(float4);
  detect_shared(c, h_np, h_nt, h_p, h_t);
  return JUR_OK;
}

// Tables from device memory in the reference's tbl_t order ([channel]
// [emitter]...): the counts and grids come to the host (the checks above
// need them), the bracket records are packed on the device.
static int build_tables_device(jur_ctx *c, const jur_dtables_t *d) {
  const Setup &s = c->s;
  const int nd = s.nd, ng = s.ng;
  const size_t ntab = (size_t)nd * ng;
  const int tp = d->tp, tt = d->tt, tu = d->tu;
  if (ntab > 0 && (tp < 1 || tt < 1 || tu < 1 || !d->np || !d->nt || !d->nu || !d->p || !d->t || !d->u ||
                   !d->eps))
    return set_err(JUR_ERR_ARG, "invalid device tables");
  std::vector<int> i_np(ntab + 1), i_nt(ntab * tp + 1), i_nu(ntab * tp * tt + 1);
  std::vector<double> i_p(ntab * tp + 1), i_t(ntab * tp * tt + 1);
  if (ntab > 0) {
    HIPCHK(hipMemcpy(i_np.data(), d->np, ntab * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(i_nt.data(), d->nt, ntab * tp * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(i_nu.data(), d->nu, ntab * tp * tt * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(i_p.data(), d->p, ntab * tp * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(i_t.data(), d->t, ntab * tp * tt * sizeof(double), hipMemcpyDeviceToHost));
  }
  // sizes, checked against the capacities as read_tbl checks them
  int TP = 2, TT = 2, TU = 2;
  for (size_t q = 0; q < ntab; q++) {
    const int np = i_np[q];
    if (np > tp) return set_err(JUR_ERR_ARG, "Too many pressure levels!");
    for (int ip = 0; ip < np; ip++) {
      const int nt = i_nt[q * tp + ip];
      if (nt > tt || nt < 0) return set_err(JUR_ERR_ARG, "Too many temperatures!");
      for (int it = 0; it < nt; it++) {
        const int n = i_nu[(q * tp + ip) * tt + it];
        if (n > tu || n < 0) return set_err(JUR_ERR_ARG, "Too many column densities!");
        TU = std::max(TU, n);
      }
      TT = std::max(TT, nt);
    }
    TP = std::max(TP, np);
  }
  c->TP = TP;
  c->TT = TT;
  c->TU = TU;
  // the library's order: table k = ig * nd + id
  std::vector<int> h_np(ntab + 1, -1), h_nt(ntab * TP + 1, 0), h_nu(ntab * TP * TT + 1, 0),
      h_mono(ntab + 1, 1), h_pdir(ntab + 1, 0);
  std::vector<double> h_p(ntab * TP + 1, 0.0), h_t(ntab * TP * TT + 1, 0.0);
  for (int ig = 0; ig < ng; ig++)
    for (int id = 0; id < nd; id++) {
      const size_t k = (size_t)ig * nd + id, q = (size_t)id * ng + ig;
      const int np = i_np[q];
      h_np[k] = np;
      for (int ip = 0; ip < np; ip++) {
        h_p[k * TP + ip] = i_p[q * tp + ip];
        const int nt = i_nt[q * tp + ip];
        h_nt[k * TP + ip] = nt;
        for (int it = 0; it < nt; it++) {
          h_t[(k * TP + ip) * TT + it] = i_t[(q * tp + ip) * tt + it];
          h_nu[(k * TP + ip) * TT + it] = i_nu[(q * tp + ip) * tt + it];
        }
      }
      h_pdir[k] = p_direction(&h_p[k * TP], std::max(np, 0));
    }
  int rc;
  if ((rc = c->d_np.upload(h_np.data(), h_np.size()))) return rc;
  if ((rc = c->d_nt.upload(h_nt.data(), h_nt.size()))) return rc;
  if ((rc = c->d_nu.upload(h_nu.data(), h_nu.size()))) return rc;
  if ((rc = c->d_mono.upload(h_mono.data(), h_mono.size()))) return rc;
  if ((rc = c->d_pdir.upload(h_pdir.data(), h_pdir.size()))) return rc;
  if ((rc = c->d_p.upload(h_p.data(), h_p.size()))) return rc;
  if ((rc = c->d_t.upload(h_t.data(), h_t.size()))) return rc;
  const size_t nrec = ntab * TP * TT * TU + 1;
  if ((rc = c->d_rec.reserve(nrec))) return rc;
  c->table_bytes = nrec * sizeof(float4);
  launch_pack_tables(nd, ng, TP, TT, TU, c->d_np.p, c->d_nt.p, c->d_nu.p, tp, tt, tu, d->u, d->eps,
                     c->d_rec.p, c->d_mono.p, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(h_mono.data(), c->d_mono.p, ntab * sizeof(int), hipMemcpyDeviceToHost));
  c->all_mono = 1;
  for (size_t k = 0; k < ntab; k++)
    if (h_np[k] >= 30 && !h_mono[k]) c->all_mono = 0;
  detect_shared(c, h_np, h_nt, h_p, h_t);
  return JUR_OK;
}

static int build_srcfunc(jur_ctx *c, const jur_dtables_t *mem = nullptr) {
  const Setup &s = c->s;
  const int nd = s.nd;
  const bool in_mem = mem && mem->filt_off;
  if (in_mem && nd > 0 && (!mem->filt_nu || !mem->filt_w))
    return set_err(JUR_ERR_ARG, "invalid filter functions");
  std::vector<double> st(TBLNS);
  for (int it = 0; it < TBLNS; it++) {
    // LIN(0.0, TMIN, TBLNS - 1.0, TMAX, (double) it)  (src/jurassic.c:3662)
    st[it] = JUR_TMIN + (JUR_TMAX - JUR_TMIN) / ((TBLNS - 1.0) - 0.0) * ((double)it - 0.0);
  }
  std::vector<double> fnu, f, dnu(nd + 1);
  std::vector<int> off(nd + 1), nf(nd + 1);
  for (int id = 0; id < nd; id++) {
    std::vector<double> x, y;
    if (in_mem) {
      const int a = mem->filt_off[id], b = mem->filt_off[id + 1];
      if (b - a < 2 || b - a > NSHAPE) return set_err(JUR_ERR_ARG, "invalid filter function of channel %d", id);
      x.assign(mem->filt_nu + a, mem->filt_nu + b);
      y.assign(mem->filt_w + a, mem->filt_w + b);
    } else {
      char fn[2 * LEN + 32];
      snprintf(fn, sizeof fn, "%s_%.4f.filt", s.tblbase.c_str(), s.nu[id]);
      int rc = read_shape(fn, x, y);
      if (rc) return rc;
    }
    double d = 1.0;
    for (size_t i = 1; i < x.size(); i++) d = std::min(d, x[i] - x[i - 1]);
    off[id] = (int)fnu.size();
    nf[id] = (int)x.size();
    dnu[id] = d;
    fnu.insert(fnu.end(), x.begin(), x.end());
    f.insert(f.end(), y.begin(), y.end());
  }
  DevBuf<double> d_fnu, d_f, d_dnu;
  DevBuf<int> d_off, d_nf;
  int rc;
  if ((rc = c->d_st.upload(st.data(), st.size()))) return rc;
  if ((rc = c->d_sr.reserve((size_t)TBLNS * std::max(nd, 1)))) return rc;
  if (nd > 0) {
    if ((rc = d_fnu.upload(fnu.data(), fnu.size()))) return rc;
    if ((rc = d_f.upload(f.data(), f.size()))) return rc;
    if ((rc = d_dnu.upload(dnu.data(), dnu.size()))) return rc;
    if ((rc = d_off.upload(off.data(), off.size()))) return rc;
    if ((rc = d_nf.upload(nf.data(), nf.size()))) return rc;
    launch_srcfunc(c->d_st.p, d_fnu.p, d_f.p, d_off.p, d_nf.p, d_dnu.p, nd, c->d_sr.p, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  }
  return JUR_OK;
}

static int build_model(jur_ctx *c) {
  const Setup &s = c->s;
  std::vector<ChanInfo> ch(std::max(s.nd, 1));
  for (int id = 0; id < s.nd; id++) {
    ChanInfo &ci = ch[id];
    ci.nu = s.nu[id];
    ci.window = s.window[id];
    ci.icl = s.ncl > 0 ? locate_irr_h(s.clnu.data(), s.ncl, s.nu[id]) : 0;
    ci.isf = s.nsf > 0 ? locate_irr_h(s.sfnu.data(), s.nsf, s.nu[id]) : 0;
    ci.pad = 0;
    const double nu = s.nu[id];
    ci.planck_sun = JUR_C1 * (nu * nu * nu) / expm1(JUR_C2 * nu / JUR_TSUN);
  }
  int rc;
  if ((rc = c->d_chan.upload(ch.data(), ch.size()))) return rc;
  if ((rc = c->d_clnu.upload(s.clnu.data(), s.clnu.size()))) return rc;
  if ((rc = c->d_sfnu.upload(s.sfnu.data(), s.sfnu.size()))) return rc;
  // continuum coefficient tables, concatenated
  const double *tabs[] = {JUR_CTM_CO2_296, JUR_CTM_CO2_260, JUR_CTM_CO2_230, JUR_CTM_H2O_296,
                          JUR_CTM_H2O_260, JUR_CTM_H2O_FRN, JUR_CTM_H2O_XFCREV, JUR_CTM_N2_B,
                          JUR_CTM_N2_BETA, JUR_CTM_N2_NU, JUR_CTM_O2_B, JUR_CTM_O2_BETA,
                          JUR_CTM_O2_NU};
  const int lens[] = {2001, 2001, 2001, 2001, 2001, 2001, 15, 98, 98, 98, 90, 90, 90};
  std::vector<double> all;
  std::vector<size_t> offs;
  for (int i = 0; i < 13; i++) {
    offs.push_back(all.size());
    all.insert(all.end(), tabs[i], tabs[i] + lens[i]);
  }
  if ((rc = c->d_ctm.upload(all.data(), all.size()))) return rc;
  const double *b = c->d_ctm.p;
  c->ctm = CtmView{b + offs[0], b + offs[1], b + offs[2], b + offs[3], b + offs[4],
                   b + offs[5], b + offs[6], b + offs[7], b + offs[8], b + offs[9],
                   b + offs[10], b + offs[11], b + offs[12]};
  if (s.fov.size() && s.fov[0] != '-') {
    if ((rc = read_shape(s.fov.c_str(), c->fov_dz, c->fov_w))) return rc;
    if ((rc = c->d_fov_dz.upload(c->fov_dz.data(), c->fov_dz.size()))) return rc;
    if ((rc = c->d_fov_w.upload(c->fov_w.data(), c->fov_w.size()))) return rc;
  }
  for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
  return JUR_OK;
}

static int find_emitter(const Setup &s, const char *name) {
  for (int ig = 0; ig < s.ng; ig++)
    if (strcasecmp(s.emitter[ig].c_str(), name) == 0) return ig;
  return -1;
}

static ModelView model_view(const jur_ctx *c) {
  const Setup &s = c->s;
  ModelView M{};
  M.nd = s.nd;
  M.ng = s.ng;
  M.nw = s.nw;
  M.ncl = s.ncl;
  M.nsf = s.nsf;
  M.sftype = s.sftype;
  M.sfsza = s.sfsza;
  M.formod = s.formod;
  M.refrac = s.refrac;
  M.rayds = s.rayds;
  M.raydz = s.raydz;
  M.ig_co2 = s.ctm_co2 ? find_emitter(s, "CO2") : -1;
  M.ig_h2o = s.ctm_h2o ? find_emitter(s, "H2O") : -1;
  M.ctm_n2 = s.ctm_n2;
  M.ctm_o2 = s.ctm_o2;
  M.write_bbt = s.write_bbt;
  M.chan = c->d_chan.p;
  M.clnu = c->d_clnu.p;
  M.sfnu = c->d_sfnu.p;
  M.st = c->d_st.p;
  M.sr = c->d_sr.p;
  M.ctm = c->ctm;
  TabView &T = M.tab;
  T.nd = s.nd;
  T.ng = s.ng;
  T.TP = c->TP;
  T.TT = c->TT;
  T.TU = c->TU;
  T.np = c->d_np.p;
  T.nt = c->d_nt.p;
  T.nu = c->d_nu.p;
  T.p = c->d_p.p;
  T.t = c->d_t.p;
  T.rec = c->d_rec.p;
  T.mono = c->d_mono.p;
  T.pdir = c->d_pdir.p;
  T.shared = c->shared;
  T.ref_tix = c->ref_tix;
  T.all_mono = c->all_mono;
  return M;
}

static int ctx_build(jur_ctx *c) {
  int rc;
  HIPCHK(hipGetDevice(&c->device));
  if ((rc = build_tables(c))) return rc;
  if ((rc = build_model(c))) return rc;
  if ((rc = build_srcfunc(c))) return rc;
  return JUR_OK;
}

// ------------------------------------------------------------------
// One forward pass over a device batch.
// ------------------------------------------------------------------
// One pass over a device batch: the forward model, or with jac the forward
// model and its Jacobian (adjoint sweep).
int forward_pass(jur_ctx *c, const BatchView &B, int precision, const unsigned char *d_mask, double *d_rad,
                 double *d_tau, double *d_tp, const FovPkg *fov, hipStream_t st, const JacArgs *jac,
                 PassMem *mem) {
  const Setup &s = c->s;
  ModelView M = model_view(c);
  PassMem own;
  PassMem &pm = mem ? *mem : own;
  int rc;
  c->last_ms[0] = c->last_ms[1] = c->last_ms[2] = 0;
  if (B.nray <= 0) return JUR_OK;

  if (s.hydz >= 0) {
    launch_hydrostatic(B.atm, B.nprof, s.hydz, find_emitter(s, "H2O"), st);
    HIPCHK(hipGetLastError());
  }

  // LOS capacity from the altitude range of the atmospheres: vertical steps
  // of RAYDZ down and up plus horizontal steps of RAYDS along the longest
  // chord, with margin; rays that still exceed it are traced again with the
  // reference limit NLOS.
  double *d_zr;
  if ((rc = pm.get(c->d_zr, (size_t)2 * B.nprof, &d_zr))) return rc;
  launch_zrange(B.atm, B.nprof, d_zr, st);
  std::vector<double> zr((size_t)2 * B.nprof);
  HIPCHK(hipMemcpyAsync(zr.data(), d_zr, zr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double dz = 0, rmax = 0, rmin = 1e30;
  for (int i = 0; i < B.nprof; i++) {
    dz = std::max(dz, zr[2 * i + 1] - zr[2 * i]);
    rmax = std::max(rmax, JUR_RE + zr[2 * i + 1]);
    rmin = std::min(rmin, JUR_RE + zr[2 * i]);
  }
  double est = 16;
  if (s.raydz > 0) est += 2.5 * dz / s.raydz;
  if (s.rayds > 0) est += 1.25 * 2 * sqrt(std::max(rmax * rmax - rmin * rmin, 0.0)) / s.rayds;
  if (!(s.raydz > 0) && !(s.rayds > 0)) est = NLOS;
  int cap = (int)std::min((double)NLOS, est);

  const int nblk = (B.nray + WAVE - 1) / WAVE;
  RayInfo *d_info;
  char *d_tail;
  unsigned long long *d_red;
  if ((rc = pm.get(c->d_info, B.nray, &d_info))) return rc;
  if ((rc = pm.get(c->d_tail, (size_t)B.nray * ray_tail_bytes(), &d_tail))) return rc;
  if ((rc = pm.get(c->d_red, 3, &d_red))) return rc;
  const size_t mark0 = pm.used;

  for (int attempt = 0; attempt < 2; attempt++) {
    pm.used = mark0;
    LosView L{};
    L.cap = cap;
    L.f_k = F_Q0 + s.ng;
    L.f_cl = s.ncl > 0 ? L.f_k + s.nw : -1;
    L.f_u = L.f_k + s.nw + (s.ncl > 0 ? 1 : 0);
    L.nfield = L.f_u + s.ng;
    // with refraction the Jacobian follows the path's dependence on p and T:
    // the march keeps each point's position and direction
    const bool dpath = jac && s.refrac;
    L.f_xe = dpath ? L.nfield : -1;
    if (dpath) L.nfield += 6;
    const size_t esz = (precision == JUR_FP32) ? sizeof(float) : sizeof(double);
    // D_i is needed only where the surface reflects (sfeps < 1 needs NSF > 0)
    const int nscr0 = (s.sftype >= 2 && s.nsf > 0) ? 2 : 1;
    const int nscr = nscr0 + (dpath ? 2 : 0);
    // FP32 EGA Jacobian over monotone tables: the lookups store each step's
    // partials (16 B per point, channel and emitter) for the reverse chains
    const bool part = jac && precision == JUR_FP32 && M.formod != 0 && M.tab.all_mono;
    // Memory per ray block: the LOS at full capacity, and per LOS point of
    // the block's longest ray the emissivities (and for the Jacobian the
    // forward-sweep scratch), plus the Jacobian's K tile.
    // the FP32 forward pass also keeps the fields rounded to float (what its
    // radiance kernel reads)
    const bool losf = precision == JUR_FP32 && !jac;
    const size_t los_blk = (size_t)L.nfield * L.cap * WAVE * (sizeof(double) + (losf ? sizeof(float) : 0));
    const size_t pt_blk = (size_t)s.nd * s.ng * WAVE * (esz + (part ? sizeof(float4) : 0)) +
                          (jac ? (size_t)s.nd * nscr * WAVE * sizeof(double) : 0) +
                          (dpath ? (size_t)JUR_NPCO * WAVE * sizeof(double) : 0);
    const size_t tile_blk = jac ? (size_t)s.nd * jac->ncol * WAVE * sizeof(double) : 0;
    size_t budget = c->los_budget;
    if (pm.base) {
      // the scratch array after the per-ray buffers, less the per-block
      // offset arrays and the alignment of the slices
      const size_t small = (size_t)nblk * (sizeof(int) + 2 * sizeof(size_t)) + 16 * 256;
      budget = pm.cap - pm.used > small ? pm.cap - pm.used - small : 0;
    } else if (budget == 0) {
      // half of the HBM free now (plus what these buffers already hold), at
      // most 96 GiB.  A pass of another kind divides the budget otherwise
      // (LOS fields, emissivities, partials, scratch): its buffers are
      // released first, so that capacities kept from the last kind do not
      // hold memory the new division counts as free.
      const int kind = precision * 4 + (jac ? 2 : 0) + (dpath ? 1 : 0);
      if (kind != c->pass_kind) {
        c->d_los.release();
        c->d_losf.release();
        c->d_eps.release();
        c->d_part.release();
        c->d_kt.release();
        c->d_scr.release();
        c->d_pco.release();
        c->pass_kind = kind;
      }
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      fr += c->d_los.n * sizeof(double) + c->d_losf.n * sizeof(float) + c->d_eps.n + c->d_part.n * sizeof(float4) +
            (c->d_kt.n + c->d_scr.n + c->d_pco.n) * sizeof(double);
      budget = std::min(fr / 2, (size_t)96 << 30);
    }
    // Rays are traced in chunks whose LOS takes at most half the budget; the
    // lookups and the radiance (or adjoint) pass then run over sub-chunks of
    // a chunk sized by the blocks' actual LOS lengths.
    int chunk = (int)std::max<size_t>(1, std::min<size_t>(nblk, (budget / 2) / (los_blk + 1)));
    size_t rest = budget > los_blk * chunk ? budget - los_blk * chunk : 0;
    double *d_los;
    float *d_losf = nullptr;
    int *d_npb;
    size_t *d_eoff, *d_soff = nullptr;
    if ((rc = pm.get(c->d_npb, chunk, &d_npb))) return rc;
    if ((rc = pm.get(c->d_eoff, chunk, &d_eoff))) return rc;
    if (jac && (rc = pm.get(c->d_soff, chunk, &d_soff))) return rc;
    if ((rc = pm.get(c->d_los, (size_t)L.nfield * L.cap * WAVE * chunk, &d_los))) return rc;
    if (losf && (rc = pm.get(c->d_losf, (size_t)L.nfield * L.cap * WAVE * chunk, &d_losf))) return rc;
    if (pm.base) rest = pm.cap - pm.used;
    const size_t mark1 = pm.used;
    L.los = d_los;
    L.losf = d_losf;
    std::vector<int> npb(chunk);
    std::vector<size_t> eoff(chunk), soff(chunk);
    float ms[3] = {0, 0, 0};
    for (int b0 = 0; b0 < nblk; b0 += chunk) {
      const int nb = std::min(chunk, nblk - b0);
      HIPCHK(hipEventRecord(c->ev[0], st));
      launch_raytrace(M, B, L, d_info, d_tail, b0, nb, st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c->ev[1], st));
      launch_block_np(d_info, B.nray, b0, nb, d_npb, st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(npb.data(), d_npb, nb * sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      {
        float a = 0;
        HIPCHK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        ms[0] += a;
      }
      for (int s0 = 0; s0 < nb;) {
        // sub-chunk [s0, s1): emissivities packed by each block's longest LOS
        int s1 = s0;
        size_t tot = 0, stot = 0, mem = 0;
        while (s1 < nb) {
          // in scratch mode each slice is rounded up to 256 bytes
          const size_t need = (size_t)npb[s1] * pt_blk + tile_blk + (pm.base ? 6 * 256 : 0);
          if (s1 > s0 && mem + need > rest) break;
          eoff[s1 - s0] = tot;
          tot += (size_t)s.nd * s.ng * npb[s1] * WAVE;
          soff[s1 - s0] = stot;
          stot += (size_t)s.nd * npb[s1] * nscr * WAVE;
          mem += need;
          s1++;
        }
        const int ns = s1 - s0;
        pm.used = mark1;
        char *d_eps;
        float4 *d_part = nullptr;
        if ((rc = pm.get(c->d_eps, tot * esz + 1, &d_eps))) return rc;
        if (part && (rc = pm.get(c->d_part, tot + 1, &d_part))) return rc;
        HIPCHK(hipMemcpyAsync(d_eoff, eoff.data(), ns * sizeof(size_t), hipMemcpyHostToDevice, st));
        LosView Ls = L;
        Ls.los = L.los + (size_t)s0 * L.cap * L.nfield * WAVE;
        if (Ls.losf) Ls.losf = L.losf + (size_t)s0 * L.cap * L.nfield * WAVE;
        EpsView E{d_eps, d_eoff, d_npb + s0, d_part};
        const int bs = b0 + s0;
        HIPCHK(hipEventRecord(c->ev[1], st));
        if (precision == JUR_FP32) launch_lookup_f32(M, B, Ls, E, d_info, bs, ns, st);
        else launch_lookup_f64(M, B, Ls, E, d_info, bs, ns, jac != nullptr, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], st));
        if (jac) {
          double *d_scr, *d_kt, *d_pco = nullptr;
          int *d_nef;
          if ((rc = pm.get(c->d_scr, stot + 1, &d_scr))) return rc;
          if (dpath && (rc = pm.get(c->d_pco, stot / ((size_t)s.nd * nscr) * JUR_NPCO + 1, &d_pco))) return rc;
          if ((rc = pm.get(c->d_kt, (size_t)ns * s.nd * jac->ncol * WAVE + 1, &d_kt))) return rc;
          if ((rc = pm.get(c->d_nef, (size_t)ns * s.nd * WAVE + 1, &d_nef))) return rc;
          HIPCHK(hipMemcpyAsync(d_soff, soff.data(), ns * sizeof(size_t), hipMemcpyHostToDevice, st));
          launch_zero(d_kt, (size_t)ns * s.nd * jac->ncol * WAVE, st);
          JacView J{jac->xmap, jac->nmap, jac->prof_map, jac->ncol, d_kt, d_scr, d_soff, nscr, d_nef,
                    dpath ? nscr0 : -1, d_pco};
          const int maxnp = *std::max_element(npb.begin() + s0, npb.begin() + s1);
          launch_adjoint(M, B, Ls, E, d_info, d_tail, bs, ns, maxnp, J, precision == JUR_FP32, d_rad, d_tau, st);
          HIPCHK(hipGetLastError());
          launch_jac_store(d_kt, s.nd, jac->ncol, bs, ns, B.nray, d_rad, c->d_chan.p, s.write_bbt, d_mask,
                           jac->k, st);
        } else if (precision == JUR_FP32) {
          launch_radiance_f32(M, B, Ls, E, d_info, bs, ns, d_rad, d_tau, st);
        } else {
          launch_radiance_f64(M, B, Ls, E, d_info, bs, ns, d_rad, d_tau, st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[3], st));
        HIPCHK(hipEventSynchronize(c->ev[3]));
        for (int k = 1; k < 3; k++) {
          float a = 0;
          HIPCHK(hipEventElapsedTime(&a, c->ev[k], c->ev[k + 1]));
          ms[k] += a;
        }
        s0 = s1;
      }
    }
    launch_reduce_info(d_info, B.nray, d_red, st);
    HIPCHK(hipGetLastError());
    unsigned long long red[3];
    HIPCHK(hipMemcpyAsync(red, d_red, sizeof red, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < 3; k++) c->last_ms[k] = ms[k];
    c->last_mean_np = (double)red[1] / B.nray;
    if (red[0] == 100 && cap < NLOS) {
      cap = NLOS;
      continue;
    }
    if (red[0] == JUR_ERR_LOS) return set_err(JUR_ERR_LOS, "Too many LOS points!");
    if (red[0] == JUR_ERR_OBSERVER) return set_err(JUR_ERR_OBSERVER, "Observer below surface!");
    if (red[0] != 0) return set_err(JUR_ERR_ARG, "ray tracing failed (status %llu)", red[0]);
    break;
  }
  pm.used = mark0;

  // tangent points
  if (d_tp) {
    // gather tpz/tplon/tplat from RayInfo with a strided device copy
    HIPCHK(hipMemcpy2DAsync(d_tp, 3 * sizeof(double), (const char *)d_info + offsetof(RayInfo, tpz),
                            sizeof(RayInfo), 3 * sizeof(double), B.nray, hipMemcpyDeviceToDevice, st));
  }

  // field-of-view convolution (before the brightness temperature)
  if (fov && fov->lo && c->fov_dz.size()) {
    double *d_rad0, *d_tau0;
    int *d_err;
    if ((rc = pm.get(c->d_rad0, (size_t)B.nray * s.nd, &d_rad0))) return rc;
    if ((rc = pm.get(c->d_tau0, (size_t)B.nray * s.nd, &d_tau0))) return rc;
    if ((rc = pm.get(c->d_err, 1, &d_err))) return rc;
    HIPCHK(hipMemcpyAsync(d_rad0, d_rad, (size_t)B.nray * s.nd * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d_tau0, d_tau, (size_t)B.nray * s.nd * sizeof(double), hipMemcpyDeviceToDevice, st));
    int zero = 0;
    HIPCHK(hipMemcpyAsync(d_err, &zero, sizeof zero, hipMemcpyHostToDevice, st));
    launch_fov(d_rad0, d_tau0, d_rad, d_tau, B.nray, s.nd, B.ray_geom, fov->lo, fov->hi, c->d_fov_dz.p,
               c->d_fov_w.p, (int)c->fov_dz.size(), d_err, st);
    HIPCHK(hipGetLastError());
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (err) return set_err(JUR_ERR_FOV, "Cannot apply FOV convolution!");
  }
  launch_finalize(d_rad, B.nray, s.nd, c->d_chan.p, s.write_bbt, d_mask, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return JUR_OK;
}

// ------------------------------------------------------------------
// Scratch array of the no-allocation entry point
// (jur_ctx_formod_device_noalloc).  Each GPU holds its own copy at its own
// address; host code uses that address, looked up once per device, never
// the array's name.
// ------------------------------------------------------------------
}  // namespace jur

static __device__ __align__(256) unsigned char scratch[JUR_NOALLOC_SCRATCH_BYTES];

namespace jur {

static std::mutex g_scratch_mu;
static char *g_scratch_dev[256];

// Device address of the scratch array on device dev (the current device).
static int noalloc_scratch(int dev, char **out) {
  if (dev < 0 || dev >= 256) return set_err(JUR_ERR_ARG, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(JUR_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_dev[dev] = (char *)p;
  }
  *out = g_scratch_dev[dev];
  return JUR_OK;
}

// ------------------------------------------------------------------
// Context cache for the ctl_t entry points.
// ------------------------------------------------------------------
static std::mutex g_mu;
static std::map<std::string, std::unique_ptr<jur_ctx>> g_cache;

int ctx_for(const ctl_t *ctl, jur_ctx **out) {
  Setup s;
  if (ctl->nd < 0 || ctl->nd > ND) return set_err(JUR_ERR_ARG, "Set 0 <= ND <= MAX!");
  int rc = setup_from_ctl(ctl, ctl->nd, ctl->nu, ctl->window, s);
  if (rc) return rc;
  const std::string k = s.key();
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(k);
  if (it != g_cache.end()) {
    *out = it->second.get();
    return JUR_OK;
  }
  std::unique_ptr<jur_ctx> c(new jur_ctx());
  c->s = s;
  if ((rc = ctx_build(c.get()))) return rc;
  *out = c.get();
  g_cache[k] = std::move(c);
  return JUR_OK;
}

// Host staging of obs_t/atm_t packages into a device batch.
struct Staging {
  DevBuf<atm_t> atm;
  DevBuf<int> prof, lo, hi;
  DevBuf<double> geom, rad, tau, tp;
  DevBuf<unsigned char> mask;
};

static int formod_packages(const ctl_t *ctl, atm_t *atm, obs_t *obs, int n, int precision) {
  jur_ctx *c;
  int rc = ctx_for(ctl, &c);
  if (rc) return rc;
  const int nd = ctl->nd;
  int nray = 0;
  for (int k = 0; k < n; k++) {
    if (obs[k].nr < 0 || obs[k].nr > NR) return set_err(JUR_ERR_ARG, "Too many rays!");
    if (atm[k].np < 2 || atm[k].np > NP) return set_err(JUR_ERR_ARG, "Too many data points!");
    nray += obs[k].nr;
  }
  // Ray order: ray index major, package minor, so that a block of 64 rays
  // holds the same geometry slot of many packages.
  std::vector<int> prof(nray), lo(nray), hi(nray), pkg_of(nray), ray_of(nray);
  std::vector<double> geom((size_t)7 * nray);
  std::vector<unsigned char> mask((size_t)nray * std::max(nd, 1));
  int maxnr = 0;
  for (int k = 0; k < n; k++) maxnr = std::max(maxnr, obs[k].nr);
  std::vector<int> pos((size_t)n * std::max(maxnr, 1), -1);
  int q = 0;
  for (int ir = 0; ir < maxnr; ir++)
    for (int k = 0; k < n; k++)
      if (ir < obs[k].nr) {
        pos[(size_t)k * maxnr + ir] = q;
        pkg_of[q] = k;
        ray_of[q] = ir;
        q++;
      }
  for (int i = 0; i < nray; i++) {
    const int k = pkg_of[i], ir = ray_of[i];
    const obs_t &o = obs[k];
    prof[i] = k;
    double *g = &geom[(size_t)7 * i];
    g[0] = o.time[ir];
    g[1] = o.obsz[ir];
    g[2] = o.obslon[ir];
    g[3] = o.obslat[ir];
    g[4] = o.vpz[ir];
    g[5] = o.vplon[ir];
    g[6] = o.vplat[ir];
    for (int id = 0; id < nd; id++) mask[(size_t)i * nd + id] = !isfinite(o.rad[id][ir]);
  }
  const bool use_fov = c->fov_dz.size() > 0;
  Staging S;
  if ((rc = S.atm.upload(atm, n))) return rc;
  if ((rc = S.prof.upload(prof.data(), nray))) return rc;
  if ((rc = S.geom.upload(geom.data(), geom.size()))) return rc;
  if ((rc = S.mask.upload(mask.data(), (size_t)nray * nd))) return rc;
  if ((rc = S.rad.reserve((size_t)nray * nd))) return rc;
  if ((rc = S.tau.reserve((size_t)nray * nd))) return rc;
  if ((rc = S.tp.reserve((size_t)nray * 3))) return rc;
  FovPkg fp;
  BatchView B{n, S.atm.p, nray, S.prof.p, S.geom.p};
  if (!use_fov) {
    rc = forward_pass(c, B, precision, S.mask.p, S.rad.p, S.tau.p, S.tp.p, nullptr, 0);
  } else {
    // FOV neighbours are rays ir-NFOV..ir+NFOV of the same package
    // (src/jurassic.c:3148-3160): package-major order keeps them contiguous
    std::vector<int> prof2(nray), lo2(nray), hi2(nray);
    std::vector<double> geom2((size_t)7 * nray);
    std::vector<unsigned char> mask2((size_t)nray * std::max(nd, 1));
    int i = 0;
    for (int k = 0; k < n; k++) {
      const int start = i;
      for (int ir = 0; ir < obs[k].nr; ir++, i++) {
        const int src = pos[(size_t)k * maxnr + ir];
        prof2[i] = k;
        lo2[i] = start;
        hi2[i] = start + obs[k].nr;
        memcpy(&geom2[(size_t)7 * i], &geom[(size_t)7 * src], 7 * sizeof(double));
        for (int id = 0; id < nd; id++) mask2[(size_t)i * nd + id] = mask[(size_t)src * nd + id];
        pos[(size_t)k * maxnr + ir] = i;
      }
    }
    if ((rc = S.prof.upload(prof2.data(), nray))) return rc;
    if ((rc = S.geom.upload(geom2.data(), geom2.size()))) return rc;
    if ((rc = S.mask.upload(mask2.data(), (size_t)nray * nd))) return rc;
    if ((rc = S.lo.upload(lo2.data(), nray))) return rc;
    if ((rc = S.hi.upload(hi2.data(), nray))) return rc;
    fp.lo = S.lo.p;
    fp.hi = S.hi.p;
    rc = forward_pass(c, B, precision, S.mask.p, S.rad.p, S.tau.p, S.tp.p, &fp, 0);
  }
  if (rc) return rc;
  // one spare element each: a batch without rays copies nothing, from and to
  // valid addresses
  std::vector<double> rad((size_t)nray * nd + 1), tau((size_t)nray * nd + 1), tp((size_t)nray * 3 + 1);
  HIPCHK(hipMemcpy(rad.data(), S.rad.p, (size_t)nray * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tau.data(), S.tau.p, (size_t)nray * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tp.data(), S.tp.p, (size_t)nray * 3 * sizeof(double), hipMemcpyDeviceToHost));
  if (ctl->hydz >= 0) {
    // hydrostatic() rewrote the pressure profiles in place
    std::vector<atm_t> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), S.atm.p, n * sizeof(atm_t), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) memcpy(atm[k].p, tmp[k].p, sizeof atm[k].p);
  }
  for (int k = 0; k < n; k++)
    for (int ir = 0; ir < obs[k].nr; ir++) {
      const int i = pos[(size_t)k * maxnr + ir];
      for (int id = 0; id < nd; id++) {
        obs[k].rad[id][ir] = rad[(size_t)i * nd + id];
        obs[k].tau[id][ir] = tau[(size_t)i * nd + id];
      }
      obs[k].tpz[ir] = tp[(size_t)3 * i];
      obs[k].tplon[ir] = tp[(size_t)3 * i + 1];
      obs[k].tplat[ir] = tp[(size_t)3 * i + 2];
    }
  return JUR_OK;
}

// ------------------------------------------------------------------
// State and measurement vectors (atm2x, x2atm, obs2y: src/jurassic.c:29-104,
// 5896-5946, 4174-4197) and the finite-difference Jacobian
// (kernel, src/jurassic.c:4004-4099).
// ------------------------------------------------------------------

size_t atm2x_h(const ctl_t *ctl, const atm_t *atm, double *x, int *iqa, int *ipa) {
  size_t n = 0;
  auto add = [&](double v, int q, int ip) {
    if (x) x[n] = v;
    if (iqa) iqa[n] = q;
    if (ipa) ipa[n] = ip;
    n++;
  };
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl;
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->retp_zmin && atm->z[ip] <= ctl->retp_zmax) add(atm->p[ip], IDXP, ip);
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->rett_zmin && atm->z[ip] <= ctl->rett_zmax) add(atm->t[ip], IDXT, ip);
  for (int ig = 0; ig < ng; ig++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retq_zmin[ig] && atm->z[ip] <= ctl->retq_zmax[ig])
        add(atm->q[ig][ip], 2 + ig, ip);
  for (int iw = 0; iw < nw; iw++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retk_zmin[iw] && atm->z[ip] <= ctl->retk_zmax[iw])
        add(atm->k[iw][ip], 2 + ng + iw, ip);
  if (ctl->ret_clz) add(atm->clz, 2 + ng + nw, 0);
  if (ctl->ret_cldz) add(atm->cldz, 3 + ng + nw, 0);
  if (ctl->ret_clk)
    for (int icl = 0; icl < ncl; icl++) add(atm->clk[icl], 4 + ng + nw + icl, 0);
  if (ctl->ret_sfz) add(atm->sfz, 4 + ng + nw + ncl, 0);
  if (ctl->ret_sfp) add(atm->sfp, 5 + ng + nw + ncl, 0);
  if (ctl->ret_sft) add(atm->sft, 6 + ng + nw + ncl, 0);
  if (ctl->ret_sfeps)
    for (int isf = 0; isf < ctl->nsf; isf++) add(atm->sfeps[isf], 7 + ng + nw + ncl + isf, 0);
  return n;
}

void x2atm_h(const ctl_t *ctl, const double *x, atm_t *atm) {
  size_t n = 0;
  const int ng = ctl->ng, nw = ctl->nw;
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->retp_zmin && atm->z[ip] <= ctl->retp_zmax) atm->p[ip] = x[n++];
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->z[ip] >= ctl->rett_zmin && atm->z[ip] <= ctl->rett_zmax) atm->t[ip] = x[n++];
  for (int ig = 0; ig < ng; ig++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retq_zmin[ig] && atm->z[ip] <= ctl->retq_zmax[ig]) atm->q[ig][ip] = x[n++];
  for (int iw = 0; iw < nw; iw++)
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= ctl->retk_zmin[iw] && atm->z[ip] <= ctl->retk_zmax[iw]) atm->k[iw][ip] = x[n++];
  if (ctl->ret_clz) atm->clz = x[n++];
  if (ctl->ret_cldz) atm->cldz = x[n++];
  if (ctl->ret_clk)
    for (int icl = 0; icl < ctl->ncl; icl++) atm->clk[icl] = x[n++];
  if (ctl->ret_sfz) atm->sfz = x[n++];
  if (ctl->ret_sfp) atm->sfp = x[n++];
  if (ctl->ret_sft) atm->sft = x[n++];
  if (ctl->ret_sfeps)
    for (int isf = 0; isf < ctl->nsf; isf++) atm->sfeps[isf] = x[n++];
}

size_t obs2y_h(const ctl_t *ctl, const obs_t *obs, double *y) {
  size_t m = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (isfinite(obs->rad[id][ir])) {
        if (y) y[m] = obs->rad[id][ir];
        m++;
      }
  return m;
}

// Perturbation size of kernel() (src/jurassic.c:4046-4068).
int fd_step(const ctl_t *ctl, int iq, double x0, double *h) {
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl, nsf = ctl->nsf;
  if (iq == IDXP) *h = std::max(fabs(0.01 * x0), 1e-7);
  else if (iq == IDXT) *h = 1.0;
  else if (iq >= 2 && iq < 2 + ng) *h = std::max(fabs(0.01 * x0), 1e-15);
  else if (iq >= 2 + ng && iq < 2 + ng + nw) *h = 1e-4;
  else if (iq == 2 + ng + nw || iq == 3 + ng + nw) *h = 1.0;
  else if (iq >= 4 + ng + nw && iq < 4 + ng + nw + ncl) *h = 1e-4;
  else if (iq == 4 + ng + nw + ncl) *h = 0.1;
  else if (iq == 5 + ng + nw + ncl) *h = 10.0;
  else if (iq == 6 + ng + nw + ncl) *h = 1.0;
  else if (iq >= 7 + ng + nw + ncl && iq < 7 + ng + nw + ncl + nsf) *h = 1e-2;
  else return set_err(JUR_ERR_ARG, "Cannot set perturbation size!");
  return JUR_OK;
}

// Forward model of n atmospheres that share one observation geometry, in
// one device pass.  Rays are ordered ray-major, atmosphere-minor, so that
// each block of 64 rays is one geometry over 64 atmospheres.
// rad/tau are [n][nr][nd] (host).
static int formod_shared_geometry(const ctl_t *ctl, const atm_t *atms, int n, const obs_t *obs,
                                  int precision, double *rad, double *tau) {
  jur_ctx *c;
  int rc = ctx_for(ctl, &c);
  if (rc) return rc;
  const int nd = ctl->nd, nr = obs->nr;
  const int nray = n * nr;
  std::vector<int> prof(nray);
  std::vector<double> geom((size_t)7 * nray);
  std::vector<unsigned char> mask((size_t)nray * std::max(nd, 1));
  for (int ir = 0; ir < nr; ir++)
    for (int k = 0; k < n; k++) {
      const int i = ir * n + k;
      prof[i] = k;
      double *g = &geom[(size_t)7 * i];
      g[0] = obs->time[ir];
      g[1] = obs->obsz[ir];
      g[2] = obs->obslon[ir];
      g[3] = obs->obslat[ir];
      g[4] = obs->vpz[ir];
      g[5] = obs->vplon[ir];
      g[6] = obs->vplat[ir];
      for (int id = 0; id < nd; id++) mask[(size_t)i * nd + id] = !isfinite(obs->rad[id][ir]);
    }
  if (c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "FOV with a shared-geometry batch");
  Staging S;
  if ((rc = S.atm.upload(atms, n))) return rc;
  if ((rc = S.prof.upload(prof.data(), nray))) return rc;
  if ((rc = S.geom.upload(geom.data(), geom.size()))) return rc;
  if ((rc = S.mask.upload(mask.data(), (size_t)nray * nd))) return rc;
  if ((rc = S.rad.reserve((size_t)nray * nd))) return rc;
  if ((rc = S.tau.reserve((size_t)nray * nd))) return rc;
  BatchView B{n, S.atm.p, nray, S.prof.p, S.geom.p};
  if ((rc = forward_pass(c, B, precision, S.mask.p, S.rad.p, S.tau.p, nullptr, nullptr, 0))) return rc;
  std::vector<double> r((size_t)nray * nd), t((size_t)nray * nd);
  HIPCHK(hipMemcpy(r.data(), S.rad.p, r.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(t.data(), S.tau.p, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int ir = 0; ir < nr; ir++)
    for (int k = 0; k < n; k++)
      for (int id = 0; id < nd; id++) {
        const size_t src = (size_t)(ir * n + k) * nd + id, dst = ((size_t)k * nr + ir) * nd + id;
        rad[dst] = r[src];
        tau[dst] = t[src];
      }
  return JUR_OK;
}

static int kernel_fd(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n,
                     int precision) {
  int rc = formod_packages(ctl, atm, obs, 1, precision);
  if (rc) return rc;
  std::vector<double> x0(n + 1), y0(m + 1);
  std::vector<int> iqa(n + 1);
  if (atm2x_h(ctl, atm, x0.data(), iqa.data(), nullptr) != n || obs2y_h(ctl, obs, y0.data()) != m)
    return set_err(JUR_ERR_ARG, "Jacobian size does not match atm2x/obs2y");
  std::vector<double> h(n + 1);
  std::vector<atm_t> atms(std::max<size_t>(n, 1));
  std::vector<double> x1(n + 1);
  for (size_t j = 0; j < n; j++) {
    if ((rc = fd_step(ctl, iqa[j], x0[j], &h[j]))) return rc;
    memcpy(&atms[j], atm, sizeof(atm_t));
    memcpy(x1.data(), x0.data(), n * sizeof(double));
    x1[j] = x1[j] + h[j];
    x2atm_h(ctl, x1.data(), &atms[j]);
  }
  const int nd = ctl->nd, nr = obs->nr;
  std::vector<double> rad((size_t)n * nr * nd + 1), tau((size_t)n * nr * nd + 1);
  if (n > 0 && (rc = formod_shared_geometry(ctl, atms.data(), (int)n, obs, precision, rad.data(), tau.data())))
    return rc;
  if (m * n) memset(k, 0, m * n * sizeof(double));
  for (size_t j = 0; j < n; j++) {
    size_t i = 0;
    for (int ir = 0; ir < nr; ir++)
      for (int id = 0; id < nd; id++)
        if (isfinite(obs->rad[id][ir])) {
          const double y1 = rad[((size_t)j * nr + ir) * nd + id];
          k[i * n + j] = (y1 - y0[i]) / h[j];
          i++;
        }
  }
  return JUR_OK;
}

int kernel_fd_host(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n, int precision) {
  return kernel_fd(ctl, atm, obs, k, m, n, precision);
}

// State map of atm2x: the column of each state element (JUR_XMAP_*).
size_t state_map_h(const ctl_t *ctl, const atm_t *atm, int *xmap) {
  const size_t n = atm2x_h(ctl, atm, nullptr, nullptr, nullptr);
  std::vector<int> iqa(n + 1), ipa(n + 1);
  atm2x_h(ctl, atm, nullptr, iqa.data(), ipa.data());
  for (int i = 0; i < JUR_XMAP_LEN; i++) xmap[i] = -1;
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl;
  for (size_t j = 0; j < n; j++) {
    const int q = iqa[j], ip = ipa[j];
    int slot;
    if (q < 2 + ng) slot = JUR_XMAP_LEV(q, ip);                              // p, T, q[ig]
    else if (q < 2 + ng + nw) slot = JUR_XMAP_LEV(2 + NG + (q - 2 - ng), ip);  // k[iw]
    else if (q == 2 + ng + nw) slot = JUR_XMAP_CLZ;
    else if (q == 3 + ng + nw) slot = JUR_XMAP_CLDZ;
    else if (q < 4 + ng + nw + ncl) slot = JUR_XMAP_CLK(q - 4 - ng - nw);
    else if (q == 4 + ng + nw + ncl) slot = JUR_XMAP_SFZ;
    else if (q == 5 + ng + nw + ncl) slot = JUR_XMAP_SFP;
    else if (q == 6 + ng + nw + ncl) slot = JUR_XMAP_SFT;
    else slot = JUR_XMAP_SFEPS(q - 7 - ng - nw - ncl);
    xmap[slot] = (int)j;
  }
  return n;
}

int check_jac_ctl(const jur_ctx *c) {
  if (c->s.hydz >= 0)
    return set_err(JUR_ERR_UNSUPPORTED, "adjoint Jacobian with HYDZ >= 0 (hydrostatic p(T)) is not available");
  if (c->fov_dz.size()) return set_err(JUR_ERR_UNSUPPORTED, "adjoint Jacobian with a FOV is not available");
  return JUR_OK;
}

// kernel() by the adjoint: one pass gives y and K; the path-only columns
// (sfz, sfp) are finite differences as in kernel().
static int kernel_adjoint(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n,
                          int precision) {
  jur_ctx *c;
  int rc = ctx_for(ctl, &c);
  if (rc) return rc;
  if ((rc = check_jac_ctl(c))) return rc;
  if (obs->nr < 0 || obs->nr > NR) return set_err(JUR_ERR_ARG, "Too many rays!");
  if (atm->np < 2 || atm->np > NP) return set_err(JUR_ERR_ARG, "Too many data points!");
  std::vector<int> xmap(JUR_XMAP_LEN);
  if (state_map_h(ctl, atm, xmap.data()) != n)
    return set_err(JUR_ERR_ARG, "Jacobian size does not match atm2x/obs2y");
  const int nd = ctl->nd, nr = obs->nr;
  std::vector<int> fdcol;  // path-only columns
  for (int slot : {JUR_XMAP_SFZ, JUR_XMAP_SFP})
    if (xmap[slot] >= 0) {
      fdcol.push_back(xmap[slot]);
      xmap[slot] = -1;
    }
  size_t m0 = 0;
  for (int ir = 0; ir < nr; ir++)
    for (int id = 0; id < nd; id++) m0 += isfinite(obs->rad[id][ir]) ? 1 : 0;
  if (m0 != m) return set_err(JUR_ERR_ARG, "Jacobian size does not match atm2x/obs2y");

  const int ncol = (int)std::max<size_t>(n, 1);
  std::vector<double> geom((size_t)7 * std::max(nr, 1));
  std::vector<unsigned char> mask((size_t)std::max(nr, 1) * std::max(nd, 1));
  std::vector<int> prof(std::max(nr, 1), 0);
  for (int ir = 0; ir < nr; ir++) {
    double *g = &geom[(size_t)7 * ir];
    g[0] = obs->time[ir];
    g[1] = obs->obsz[ir];
    g[2] = obs->obslon[ir];
    g[3] = obs->obslat[ir];
    g[4] = obs->vpz[ir];
    g[5] = obs->vplon[ir];
    g[6] = obs->vplat[ir];
    for (int id = 0; id < nd; id++) mask[(size_t)ir * nd + id] = !isfinite(obs->rad[id][ir]);
  }
  Staging S;
  DevBuf<int> pm, xm;
  DevBuf<double> kd;
  const int zero = 0;
  if ((rc = S.atm.upload(atm, 1))) return rc;
  if ((rc = S.prof.upload(prof.data(), std::max(nr, 1)))) return rc;
  if ((rc = S.geom.upload(geom.data(), geom.size()))) return rc;
  if ((rc = S.mask.upload(mask.data(), mask.size()))) return rc;
  if ((rc = S.rad.reserve((size_t)std::max(nr, 1) * std::max(nd, 1)))) return rc;
  if ((rc = S.tau.reserve((size_t)std::max(nr, 1) * std::max(nd, 1)))) return rc;
  if ((rc = S.tp.reserve((size_t)std::max(nr, 1) * 3))) return rc;
  if ((rc = pm.upload(&zero, 1))) return rc;
  if ((rc = xm.upload(xmap.data(), xmap.size()))) return rc;
  if ((rc = kd.reserve((size_t)std::max(nr, 1) * std::max(nd, 1) * ncol))) return rc;
  JacArgs ja;
  ja.xmap = xm.p;
  ja.nmap = 1;
  ja.prof_map = pm.p;
  ja.ncol = ncol;
  ja.k = kd.p;
  BatchView B{1, S.atm.p, nr, S.prof.p, S.geom.p};
  if ((rc = forward_pass(c, B, precision, S.mask.p, S.rad.p, S.tau.p, S.tp.p, nullptr, 0, &ja))) return rc;
  std::vector<double> rad((size_t)nr * nd + 1), tau((size_t)nr * nd + 1), tp((size_t)nr * 3 + 1);
  std::vector<double> kh((size_t)nr * nd * ncol + 1);
  HIPCHK(hipMemcpy(rad.data(), S.rad.p, (size_t)nr * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tau.data(), S.tau.p, (size_t)nr * nd * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tp.data(), S.tp.p, (size_t)nr * 3 * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(kh.data(), kd.p, (size_t)nr * nd * ncol * sizeof(double), hipMemcpyDeviceToHost));
  for (int ir = 0; ir < nr; ir++) {
    for (int id = 0; id < nd; id++) {
      obs->rad[id][ir] = rad[(size_t)ir * nd + id];
      obs->tau[id][ir] = tau[(size_t)ir * nd + id];
    }
    obs->tpz[ir] = tp[(size_t)3 * ir];
    obs->tplon[ir] = tp[(size_t)3 * ir + 1];
    obs->tplat[ir] = tp[(size_t)3 * ir + 2];
  }
  size_t i = 0;
  for (int ir = 0; ir < nr; ir++)
    for (int id = 0; id < nd; id++)
      if (!mask[(size_t)ir * nd + id]) {
        memcpy(&k[i * n], &kh[((size_t)ir * nd + id) * ncol], n * sizeof(double));
        i++;
      }
  if (fdcol.empty()) return JUR_OK;

  // finite differences of the path-only columns (src/jurassic.c:4040-4090)
  std::vector<double> x0(n + 1), x1(n + 1), h(fdcol.size());
  std::vector<int> iqa(n + 1);
  atm2x_h(ctl, atm, x0.data(), iqa.data(), nullptr);
  const int nf = (int)fdcol.size();
  std::vector<atm_t> atms(nf + 1);
  memcpy(&atms[0], atm, sizeof(atm_t));
  for (int f = 0; f < nf; f++) {
    const int j = fdcol[f];
    if ((rc = fd_step(ctl, iqa[j], x0[j], &h[f]))) return rc;
    memcpy(&atms[f + 1], atm, sizeof(atm_t));
    memcpy(x1.data(), x0.data(), n * sizeof(double));
    x1[j] = x1[j] + h[f];
    x2atm_h(ctl, x1.data(), &atms[f + 1]);
  }
  std::vector<double> fr((size_t)(nf + 1) * nr * nd + 1), ft((size_t)(nf + 1) * nr * nd + 1);
  if ((rc = formod_shared_geometry(ctl, atms.data(), nf + 1, obs, precision, fr.data(), ft.data()))) return rc;
  for (int f = 0; f < nf; f++) {
    const int j = fdcol[f];
    size_t r = 0;
    for (int ir = 0; ir < nr; ir++)
      for (int id = 0; id < nd; id++)
        if (!mask[(size_t)ir * nd + id]) {
          const double y0 = fr[(size_t)ir * nd + id];
          const double y1 = fr[((size_t)(f + 1) * nr + ir) * nd + id];
          k[r * n + j] = (y1 - y0) / h[f];
          r++;
        }
  }
  return JUR_OK;
}

}  // namespace jur

// ------------------------------------------------------------------
// C-ABI.
// ------------------------------------------------------------------
using namespace jur;

extern "C" {

const char *jur_gpu_last_error(void) { return g_err; }

int jur_gpu_abi_check(size_t sc, size_t sa, size_t so) {
  return (sc == sizeof(ctl_t) && sa == sizeof(atm_t) && so == sizeof(obs_t)) ? 0 : 1;
}

int jur_gpu_formod(const ctl_t *ctl, atm_t *atm, obs_t *obs, int precision) {
  if (!ctl || !atm || !obs) return set_err(JUR_ERR_ARG, "NULL argument");
  return formod_packages(ctl, atm, obs, 1, precision);
}

int jur_gpu_formod_batch(const ctl_t *ctl, atm_t *atm, obs_t *obs, int n, int precision) {
  if (!ctl || (n > 0 && (!atm || !obs)) || n < 0) return set_err(JUR_ERR_ARG, "invalid argument");
  if (n == 0) return JUR_OK;
  return formod_packages(ctl, atm, obs, n, precision);
}

void formod_gpu(const ctl_t *ctl, atm_t *atm, obs_t *obs) {
  const int rc = jur_gpu_formod(ctl, atm, obs, JUR_FP64);
  if (rc) {
    printf("\nError (%s, %s, l%d): %s\n", __FILE__, __func__, __LINE__, g_err);
    exit(EXIT_FAILURE);
  }
}

int jur_gpu_kernel(const ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n,
                   int method, int precision) {
  if (!ctl || !atm || !obs || (!k && m * n)) return set_err(JUR_ERR_ARG, "NULL argument");
  if (precision != JUR_FP64 && precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (method == JUR_JAC_FD) return kernel_fd(ctl, atm, obs, k, m, n, precision);
  if (method == JUR_JAC_ADJOINT) return kernel_adjoint(ctl, atm, obs, k, m, n, precision);
  return set_err(JUR_ERR_UNSUPPORTED, "Jacobian method %d is not available", method);
}

void kernel_gpu(ctl_t *ctl, atm_t *atm, obs_t *obs, double *k, size_t m, size_t n) {
  const int rc = jur_gpu_kernel(ctl, atm, obs, k, m, n, JUR_JAC_FD, JUR_FP64);
  if (rc) {
    printf("\nError (%s, %s, l%d): %s\n", __FILE__, __func__, __LINE__, g_err);
    exit(EXIT_FAILURE);
  }
}

int jur_ctx_create(const ctl_t *ctl, jur_ctx_t **out) {
  if (!ctl || !out) return set_err(JUR_ERR_ARG, "NULL argument");
  if (ctl->nd < 0 || ctl->nd > ND) return set_err(JUR_ERR_ARG, "Set 0 <= ND <= MAX!");
  return jur_ctx_create_wide(ctl, ctl->nd, ctl->nu, ctl->window, out);
}

int jur_ctx_create_wide(const ctl_t *ctl, int nd, const double *nu, const int *window,
                        jur_ctx_t **out) {
  if (!ctl || !out || (nd > 0 && !nu)) return set_err(JUR_ERR_ARG, "NULL argument");
  std::unique_ptr<jur_ctx> c(new jur_ctx());
  int rc = setup_from_ctl(ctl, nd, nu, window, c->s);
  if (rc) return rc;
  if ((rc = ctx_build(c.get()))) return rc;
  *out = c.release();
  return JUR_OK;
}

int jur_ctx_create_device_tables(const ctl_t *ctl, int nd, const double *nu, const int *window,
                                 const jur_dtables_t *tab, jur_ctx_t **out) {
  if (!ctl || !out || !tab || (nd > 0 && !nu)) return set_err(JUR_ERR_ARG, "NULL argument");
  std::unique_ptr<jur_ctx> c(new jur_ctx());
  int rc = setup_from_ctl(ctl, nd, nu, window, c->s);
  if (rc) return rc;
  HIPCHK(hipGetDevice(&c->device));
  if ((rc = build_tables_device(c.get(), tab))) return rc;
  if ((rc = build_model(c.get()))) return rc;
  if ((rc = build_srcfunc(c.get(), tab))) return rc;
  *out = c.release();
  return JUR_OK;
}

void jur_ctx_destroy(jur_ctx_t *ctx) { delete ctx; }

int jur_ctx_nd(const jur_ctx_t *ctx) { return ctx ? ctx->s.nd : -1; }

int jur_ctx_formod_device(jur_ctx_t *ctx, const jur_batch_t *b, int precision, void *stream) {
  if (!ctx || !b) return set_err(JUR_ERR_ARG, "NULL argument");
  if (b->nray < 0 || b->nprof < 0 || (b->nray > 0 && (!b->atm || !b->ray_prof || !b->ray_geom ||
                                                      !b->rad || !b->tau)))
    return set_err(JUR_ERR_ARG, "invalid batch");
  BatchView B{b->nprof, b->atm, b->nray, b->ray_prof, b->ray_geom};
  return forward_pass(ctx, B, precision, b->mask, b->rad, b->tau, b->tp, nullptr,
                      (hipStream_t)stream);
}

int jur_ctx_formod_device_noalloc(jur_ctx_t *ctx, const jur_batch_t *b, int precision, void *stream) {
  if (!ctx || !b) return set_err(JUR_ERR_ARG, "NULL argument");
  if (precision != JUR_FP64 && precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (b->nray < 0 || b->nprof < 0 || (b->nray > 0 && (!b->atm || !b->ray_prof || !b->ray_geom ||
                                                      !b->rad || !b->tau)))
    return set_err(JUR_ERR_ARG, "invalid batch");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev != ctx->device)
    return set_err(JUR_ERR_ARG, "the context belongs to device %d, the current device is %d", ctx->device, dev);
  PassMem pm;
  int rc = noalloc_scratch(dev, &pm.base);
  if (rc) return rc;
  pm.cap = JUR_NOALLOC_SCRATCH_BYTES;
  BatchView B{b->nprof, b->atm, b->nray, b->ray_prof, b->ray_geom};
  return forward_pass(ctx, B, precision, b->mask, b->rad, b->tau, b->tp, nullptr, (hipStream_t)stream, nullptr,
                      &pm);
}

// Name of state quantity idx (idx2name, src/jurassic.c:3586-3627).
static void quantity_name(const ctl_t *ctl, int idx, char *q, size_t len) {
  const int ng = ctl->ng, nw = ctl->nw, ncl = ctl->ncl;
  q[0] = 0;
  if (idx == IDXP) snprintf(q, len, "PRESSURE");
  else if (idx == IDXT) snprintf(q, len, "TEMPERATURE");
  else if (idx < 2 + ng) snprintf(q, len, "%s", ctl->emitter[idx - 2]);
  else if (idx < 2 + ng + nw) snprintf(q, len, "EXTINCT_WINDOW_%d", idx - 2 - ng);
  else if (idx == 2 + ng + nw) snprintf(q, len, "CLOUD_HEIGHT");
  else if (idx == 3 + ng + nw) snprintf(q, len, "CLOUD_DEPTH");
  else if (idx < 4 + ng + nw + ncl) snprintf(q, len, "CLOUD_EXTINCT_%.4f", ctl->clnu[idx - 4 - ng - nw]);
  else if (idx == 4 + ng + nw + ncl) snprintf(q, len, "SURFACE_HEIGHT");
  else if (idx == 5 + ng + nw + ncl) snprintf(q, len, "SURFACE_PRESSURE");
  else if (idx == 6 + ng + nw + ncl) snprintf(q, len, "SURFACE_TEMPERATURE");
  else if (idx < 7 + ng + nw + ncl + ctl->nsf)
    snprintf(q, len, "SURFACE_EMISSIVITY_%.4f", ctl->sfnu[idx - 7 - ng - nw - ncl]);
}

int jur_write_matrix(const char *dirname, const char *filename, const ctl_t *ctl, const double *matrix,
                     const atm_t *atm, const obs_t *obs, const char *rowspace, const char *colspace,
                     const char *sort) {
  if (!filename || !ctl || !matrix || !atm || !obs || !rowspace || !colspace || !sort)
    return set_err(JUR_ERR_ARG, "NULL argument");
  if (!ctl->write_matrix) return JUR_OK;
  if (obs->nr < 0 || obs->nr > NR || ctl->nd < 0 || ctl->nd > ND || atm->np < 0 || atm->np > NP)
    return set_err(JUR_ERR_ARG, "Dimension out of range!");
  // measurement space: rows of obs2y (ray-major, finite radiances);
  // state space: elements of atm2x
  std::vector<int> yd, yr, xq(JUR_XMAP_LEN), xp(JUR_XMAP_LEN);
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (isfinite(obs->rad[id][ir])) {
        yd.push_back(id);
        yr.push_back(ir);
      }
  const size_t nx = atm2x_h(ctl, atm, nullptr, xq.data(), xp.data());
  const bool ry = rowspace[0] == 'y', cy = colspace[0] == 'y';
  const size_t nr = ry ? yd.size() : nx, nc = cy ? yd.size() : nx;
  std::string file = dirname ? std::string(dirname) + "/" + filename : std::string(filename);
  FILE *out = fopen(file.c_str(), "w");
  if (!out) return set_err(JUR_ERR_IO, "Cannot create file!");
  static const char *yhead[6] = {"index (measurement space)", "channel wavenumber [cm^-1]",
                                 "time (seconds since 2000-01-01T00:00Z)", "view point altitude [km]",
                                 "view point longitude [deg]", "view point latitude [deg]"};
  static const char *xhead[6] = {"index (state space)", "name of quantity",
                                 "time (seconds since 2000-01-01T00:00Z)", "altitude [km]", "longitude [deg]",
                                 "latitude [deg]"};
  for (int k = 0; k < 6; k++) fprintf(out, "# $%d = Row: %s\n", k + 1, ry ? yhead[k] : xhead[k]);
  for (int k = 0; k < 6; k++) fprintf(out, "# $%d = Col: %s\n", k + 7, cy ? yhead[k] : xhead[k]);
  fprintf(out, "# $13 = Matrix element\n\n");
  char q[LEN];
  auto meta = [&](bool y, size_t i, const char *sep) {
    if (y) {
      const int id = yd[i], ir = yr[i];
      fprintf(out, "%s%d %.4f %.2f %g %g %g", sep, (int)i, ctl->nu[id], obs->time[ir], obs->vpz[ir],
              obs->vplon[ir], obs->vplat[ir]);
    } else {
      const int ip = xp[i];
      quantity_name(ctl, xq[i], q, sizeof q);
      fprintf(out, "%s%d %s %.2f %g %g %g", sep, (int)i, q, atm->time[ip], atm->z[ip], atm->lon[ip], atm->lat[ip]);
    }
  };
  size_t i = 0, j = 0;
  while (i < nr && j < nc) {
    meta(ry, i, "");
    meta(cy, j, " ");
    fprintf(out, " %g\n", matrix[i * nc + j]);
    if (sort[0] == 'r') {
      if (++j >= nc) {
        j = 0;
        i++;
        fprintf(out, "\n");
      }
    } else {
      if (++i >= nr) {
        i = 0;
        j++;
        fprintf(out, "\n");
      }
    }
  }
  const bool bad = ferror(out) != 0;
  if (fclose(out) != 0 || bad) return set_err(JUR_ERR_IO, "Cannot write file!");
  return JUR_OK;
}

int jur_state_map(const ctl_t *ctl, const atm_t *atm, int *xmap) {
  if (!ctl || !atm || !xmap) return set_err(JUR_ERR_ARG, "NULL argument"), -1;
  return (int)state_map_h(ctl, atm, xmap);
}

int jur_ctx_jacobian_device(jur_ctx_t *ctx, const jur_batch_t *b, const jur_jac_t *jac, int precision,
                            void *stream) {
  if (!ctx || !b || !jac) return set_err(JUR_ERR_ARG, "NULL argument");
  if (precision != JUR_FP64 && precision != JUR_FP32) return set_err(JUR_ERR_ARG, "invalid precision");
  if (b->nray < 0 || b->nprof < 0 || (b->nray > 0 && (!b->atm || !b->ray_prof || !b->ray_geom ||
                                                      !b->rad || !b->tau)))
    return set_err(JUR_ERR_ARG, "invalid batch");
  if (jac->nmap < 0 || jac->ncol < 1 || (jac->nmap > 0 && !jac->xmap) || (b->nprof > 0 && !jac->prof_map) ||
      (b->nray > 0 && !jac->k))
    return set_err(JUR_ERR_ARG, "invalid Jacobian request");
  int rc = check_jac_ctl(ctx);
  if (rc) return rc;
  for (int im = 0; im < jac->nmap; im++) {
    const int *x = jac->xmap + (size_t)im * JUR_XMAP_LEN;
    for (int i = 0; i < JUR_XMAP_LEN; i++)
      if (x[i] < -1 || x[i] >= jac->ncol) return set_err(JUR_ERR_ARG, "state map %d: column out of range", im);
    if (x[JUR_XMAP_SFZ] >= 0 || x[JUR_XMAP_SFP] >= 0)
      return set_err(JUR_ERR_UNSUPPORTED, "surface altitude/pressure columns need finite differences");
  }
  if ((rc = ctx->d_xmap.upload(jac->xmap, (size_t)jac->nmap * JUR_XMAP_LEN))) return rc;
  JacArgs ja;
  ja.xmap = ctx->d_xmap.p;
  ja.nmap = jac->nmap;
  ja.prof_map = jac->prof_map;
  ja.ncol = jac->ncol;
  ja.k = jac->k;
  BatchView B{b->nprof, b->atm, b->nray, b->ray_prof, b->ray_geom};
  return forward_pass(ctx, B, precision, b->mask, b->rad, b->tau, b->tp, nullptr, (hipStream_t)stream, &ja);
}

int jur_ctx_last_timing(const jur_ctx_t *ctx, double *ms, int n) {
  if (!ctx || !ms) return set_err(JUR_ERR_ARG, "NULL argument");
  for (int i = 0; i < n && i < 3; i++) ms[i] = ctx->last_ms[i];
  return JUR_OK;
}

double jur_ctx_last_mean_np(const jur_ctx_t *ctx) { return ctx ? ctx->last_mean_np : 0; }

}  // extern "C"
