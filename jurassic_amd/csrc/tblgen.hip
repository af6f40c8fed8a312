// Emissivity tables from monochromatic spectra: the device part of tblgen
// (src/tblgen.c:44-140).  For one pressure and temperature, tblgen turns an
// optical-depth spectrum into cross sections a_i (optical depth over the
// column density of 1 km at 1 ppmv) and evaluates, for the column densities
// u_k = 1.122^k (k = 0, 1, ... while u_k <= 1e30),
//   eps_k = 1 - sum_i f_i exp(-a_i u_k) / sum_i f_i
// over the spectrum points inside the filter function, f_i the filter
// interpolated at the point.  The rows it prints are the eps_k in
// [1e-6, 0.999999] that exceed the previous one, up to the first above
// 0.999999.
//
// Here the host selects the points and their filter weights in the
// reference's order (one pass over the grid, shared by all spectra of a
// batch on that grid), the device evaluates the ~600 sums over up to 10^7
// points per spectrum (one exp per point and column density: the FP64 work
// of the tool), and the host applies the row selection.
//
// k_tblgen_part: one workgroup per (chunk of TG_CH points, spectrum), the
// chunk's cross sections and weights staged in LDS; thread t sums the
// chunk's terms in point order for k = t, t + 256, ...
// k_tblgen_sum: the chunk sums of each (spectrum, k) added in chunk order.
#include "jur_device.hpp"
#include "jur_host.hpp"

#include <algorithm>
#include <vector>

namespace jur {

constexpr int TG_T = 256;
constexpr int TG_CH = 2048;

__global__ __launch_bounds__(TG_T) void k_tblgen_part(const double *tau, const int *pts, const double *f, int n,
                                                      int npts, const double *uref, const double *u, int nu,
                                                      double *part) {
  __shared__ double sa[TG_CH], sf[TG_CH];
  const int s = blockIdx.y;
  const int c0 = blockIdx.x * TG_CH;
  const int m = n - c0 < TG_CH ? n - c0 : TG_CH;
  const double *ts = tau + (size_t)s * npts;
  const double ur = uref[s];
  for (int i = threadIdx.x; i < m; i += TG_T) {
    sa[i] = ts[pts[c0 + i]] / ur;  // src/tblgen.c:94, :99
    sf[i] = f[c0 + i];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nu; k += TG_T) {
    const double uk = u[k];
    double acc = 0;
    for (int i = 0; i < m; i++) acc += sf[i] * exp(-sa[i] * uk);
    part[((size_t)s * gridDim.x + blockIdx.x) * nu + k] = acc;
  }
}

__global__ __launch_bounds__(TG_T) void k_tblgen_sum(const double *part, int nchunk, int nu, double *out) {
  const int s = blockIdx.y;
  const int k = blockIdx.x * TG_T + threadIdx.x;
  if (k >= nu) return;
  double acc = 0;
  for (int c = 0; c < nchunk; c++) acc += part[((size_t)s * nchunk + c) * nu + k];
  out[(size_t)s * nu + k] = acc;
}

// Filter weights of the points tblgen integrates (src/tblgen.c:114-125),
// with its filter bracket carried in and out (idx); returns fsum in the
// reference's order.
static double tblgen_weights(double nu0, double dnu, int i0, int npts, const double *nuf, const double *filt,
                             int nf, int &idx, std::vector<int> &pts, std::vector<double> &w) {
  double fsum = 0;
  pts.clear();
  w.clear();
  for (int i = i0 < 0 ? 0 : i0; i < npts; i++) {
    const double nu = nu0 + dnu * (double)i;
    if (nu < nuf[0]) continue;
    if (nu > nuf[nf - 1]) break;
    if (nu < nuf[idx] || nu > nuf[idx + 1]) idx = locate_irr(nuf, nf, nu);
    const double f = filt[idx] + (filt[idx + 1] - filt[idx]) / (nuf[idx + 1] - nuf[idx]) * (nu - nuf[idx]);
    fsum += f;
    pts.push_back(i);
    w.push_back(f);
  }
  return fsum;
}

}  // namespace jur

using namespace jur;

extern "C" int jur_tblgen_batch(int nspec, const double *press, const double *temp, const double *tau, int npts,
                                double nu0, double nu1, const double *nuf, const double *filt, int nf,
                                double *u_out, double *eps_out, int maxrow, int *nrow) {
  if (nspec < 0 || (nspec > 0 && (!press || !temp || !tau || !nrow)) || !nuf || !filt ||
      (maxrow > 0 && nspec > 0 && (!u_out || !eps_out)))
    return set_err(JUR_ERR_ARG, "jur_tblgen: null pointer or negative count");
  if (npts < 2 || nf < 2) return set_err(JUR_ERR_ARG, "jur_tblgen: need npts >= 2 and nf >= 2");
  if (nspec == 0) return JUR_OK;
  // src/tblgen.c:106-108 (the grid and the filter are common to the spectra)
  const double dnu = (nu1 - nu0) / ((double)npts - 1.0);
  const double r0 = (nuf[0] - nu0) / (nu1 - nu0) * (double)npts;
  const int i0 = (int)r0;
  // The bracket carried from point to point also carries from one column
  // density to the next (a static in the reference): the first column
  // density starts from bracket 0, the others from where the previous one
  // ended.  The weights differ only where a point lies on a filter node.
  int idx = 0;
  std::vector<int> pts, pts2;
  std::vector<double> w1, w2;
  const double fsum1 = tblgen_weights(nu0, dnu, i0, npts, nuf, filt, nf, idx, pts, w1);
  const double fsum2 = tblgen_weights(nu0, dnu, i0, npts, nuf, filt, nf, idx, pts2, w2);
  const bool two = fsum2 != fsum1 || w2 != w1;
  const int n = (int)pts.size();
  std::vector<double> uk;
  for (double u = 1.0; u <= 1e30; u *= 1.122) uk.push_back(u);  // src/tblgen.c:111
  const int nu = (int)uk.size();
  std::vector<double> S((size_t)nspec * nu, 0.0), S2(two ? (size_t)nspec * nu : 0, 0.0);
  if (n > 0) {
    // src/tblgen.c:64: the column density of 1 km at 1 ppmv
    std::vector<double> uref(nspec);
    for (int s = 0; s < nspec; s++) uref[s] = 1e-6 * press[s] * 100 / (1.380658e-23 * temp[s]) * 1000 / 1e4;
    const int nchunk = (n + TG_CH - 1) / TG_CH;
    // spectra in groups of at most 2^27 spectrum values (1 GiB) and 65535
    // (the grid's y extent)
    const int gmax = (int)std::max<size_t>(1, std::min<size_t>(65535, ((size_t)1 << 27) / (size_t)npts));
    DevBuf<double> d_tau, d_f1, d_f2, d_u, d_uref, d_part, d_s;
    DevBuf<int> d_pts;
    int rc;
    if ((rc = d_pts.upload(pts.data(), n))) return rc;
    if ((rc = d_f1.upload(w1.data(), n))) return rc;
    if (two && (rc = d_f2.upload(w2.data(), n))) return rc;
    if ((rc = d_u.upload(uk.data(), nu))) return rc;
    if ((rc = d_uref.upload(uref.data(), nspec))) return rc;
    for (int s0 = 0; s0 < nspec; s0 += gmax) {
      const int g = std::min(gmax, nspec - s0);
      if ((rc = d_tau.upload(tau + (size_t)s0 * npts, (size_t)g * npts))) return rc;
      if ((rc = d_part.reserve((size_t)g * nchunk * nu))) return rc;
      if ((rc = d_s.reserve((size_t)g * nu))) return rc;
      const dim3 gp(nchunk, g), gs((nu + TG_T - 1) / TG_T, g);
      for (int pass = 0; pass < (two ? 2 : 1); pass++) {
        hipLaunchKernelGGL(k_tblgen_part, gp, dim3(TG_T), 0, 0, d_tau.p, d_pts.p, pass ? d_f2.p : d_f1.p, n, npts,
                           d_uref.p + s0, d_u.p, nu, d_part.p);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_tblgen_sum, gs, dim3(TG_T), 0, 0, d_part.p, nchunk, nu, d_s.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy((pass ? S2 : S).data() + (size_t)s0 * nu, d_s.p, (size_t)g * nu * sizeof(double),
                         hipMemcpyDeviceToHost));
      }
    }
  }
  // row selection and termination per spectrum (src/tblgen.c:129-138)
  for (int s = 0; s < nspec; s++) {
    double epsold = 0;
    int nr = 0;
    for (int k = 0; k < nu; k++) {
      const bool first = k == 0 || !two;
      const double e = 1 - (first ? S[(size_t)s * nu + k] : S2[(size_t)s * nu + k]) / (first ? fsum1 : fsum2);
      if (e >= 1e-6 && e <= 0.999999 && e > epsold) {
        if (nr < maxrow) {
          u_out[(size_t)s * maxrow + nr] = uk[k];
          eps_out[(size_t)s * maxrow + nr] = e;
        }
        nr++;
      }
      epsold = e;
      if (e > 0.999999) break;
    }
    nrow[s] = nr;
  }
  return JUR_OK;
}

extern "C" int jur_tblgen(double press, double temp, const double *tau, int npts, double nu0, double nu1,
                          const double *nuf, const double *filt, int nf, double *u_out, double *eps_out,
                          int maxrow, int *nrow) {
  if (!tau || !nrow) return set_err(JUR_ERR_ARG, "jur_tblgen: null pointer");
  return jur_tblgen_batch(1, &press, &temp, tau, npts, nu0, nu1, nuf, filt, nf, u_out, eps_out, maxrow, nrow);
}
