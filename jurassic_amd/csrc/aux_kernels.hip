// Small device kernels around the radiance pass: the source-function table
// (init_srcfunc, src/jurassic.c:3631-3681), per-atmosphere altitude ranges,
// status/LOS-length reductions, field-of-view convolution
// (formod_fov, src/jurassic.c:3121-3192) and the final brightness
// temperature / NaN-mask step of formod (src/jurassic.c:3057-3067).
#include "jur_device.hpp"

namespace jur {

// One thread per (temperature, channel): the filter-weighted Planck mean with
// the reference's sampling loop fnu += dnu from the first to the last
// filter point.  sr is [nd][TBLNS].
__global__ __launch_bounds__(256) void k_srcfunc(const double *st, const double *fnu_all,
                                                 const double *f_all, const int *off, const int *nf,
                                                 const double *dnu, int nd, double *sr) {
  JUR_FP_EXACT
  const int gid = blockIdx.x * 256 + threadIdx.x;
  if (gid >= nd * TBLNS) return;
  const int id = gid / TBLNS, it = gid - id * TBLNS;
  const double *nu = fnu_all + off[id];
  const double *f = f_all + off[id];
  const int n = nf[id];
  const double T = st[it];
  double fsum = 0, acc = 0;
  for (double fnu = nu[0]; fnu <= nu[n - 1]; fnu += dnu[id]) {
    const int i = locate_irr(nu, n, fnu);
    const double ff = lin(nu[i], f[i], nu[i + 1], f[i + 1], fnu);
    fsum += ff;
    acc += ff * planck(T, fnu);
  }
  sr[(size_t)id * TBLNS + it] = acc / fsum;
}

// Altitude range of each atmosphere (for the LOS capacity bound).
__global__ void k_zrange(const atm_t *atm, int nprof, double *zr) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= nprof) return;
  const atm_t *a = &atm[ia];
  double lo = a->z[0], hi = a->z[0];
  for (int i = 1; i < a->np; i++) {
    lo = fmin(lo, a->z[i]);
    hi = fmax(hi, a->z[i]);
  }
  zr[2 * ia] = lo;
  zr[2 * ia + 1] = hi;
}

// Reduce ray statuses and LOS lengths: out[0] = max status, out[1] = sum np,
// out[2] = max np.  out must be zeroed (by a fill kernel) before the launch.
__global__ void k_reduce_info(const RayInfo *info, int nray, unsigned long long *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int st = 0, np = 0;
  if (i < nray) {
    st = info[i].status;
    np = info[i].np;
  }
  // wave-level reduction, then one atomic per wave
  for (int o = 32; o > 0; o >>= 1) {
    st = max(st, __shfl_xor(st, o));
    np = np + __shfl_xor(np, o);
  }
  int npm = (i < nray) ? info[i].np : 0;
  for (int o = 32; o > 0; o >>= 1) npm = max(npm, __shfl_xor(npm, o));
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], (unsigned long long)st);
    atomicAdd(&out[1], (unsigned long long)np);
    atomicMax(&out[2], (unsigned long long)npm);
  }
}

// Longest LOS of each block of WAVE rays (the per-chain stride of the
// emissivity buffer).  One thread per block.
__global__ void k_block_np(const RayInfo *info, int nray, int blk0, int nb, int *npb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int m = 0;
  for (int l = 0; l < WAVE; l++) {
    const int ir = (blk0 + b) * WAVE + l;
    if (ir < nray && info[ir].status == JUR_OK) m = max(m, info[ir].np);
  }
  npb[b] = m;
}

__global__ void k_fill_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Zero n 8-byte words (grid-stride, vector stores).
__global__ void k_zero(unsigned long long *p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// Bracket records of tables given in the reference's tbl_t order (one
// table per [channel][emitter], capacities tp, tt, tu) in the library's order
// (table k = ig * nd + id, capacities TP, TT, TU, counts np/nt/nu already in
// that order): rec[i] = (u_i, eps_i, u_{i+1}, eps_{i+1}), the last record of a
// column repeating its knot, zeros past the column.  A table with a column
// that decreases in u or eps gets mono[k] = 0 (every writer stores 0).
__global__ void k_pack_tables(int nd, int ng, int TP, int TT, int TU, const int *np, const int *nt,
                              const int *nu, int tp, int tt, int tu, const float *u_in, const float *eps_in,
                              float4 *rec, int *mono) {
  const size_t total = (size_t)nd * ng * TP * TT * TU;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += stride) {
    size_t q = r;
    const int i = (int)(q % TU);
    q /= TU;
    const int it = (int)(q % TT);
    q /= TT;
    const int ip = (int)(q % TP);
    const int k = (int)(q / TP);
    const int ig = k / nd, id = k - ig * nd;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ip < np[k] && it < nt[(size_t)k * TP + ip]) {
      const int n = nu[((size_t)k * TP + ip) * TT + it];
      if (i < n) {
        const size_t src = ((((size_t)id * ng + ig) * tp + ip) * tt + it) * tu;
        const int j = i + 1 < n ? i + 1 : i;
        const float u0 = u_in[src + i], e0 = eps_in[src + i], u1 = u_in[src + j], e1 = eps_in[src + j];
        v = make_float4(u0, e0, u1, e1);
        if (i + 1 < n && !(u1 >= u0 && e1 >= e0)) mono[k] = 0;
      }
    }
    rec[r] = v;
  }
}

void launch_pack_tables(int nd, int ng, int TP, int TT, int TU, const int *np, const int *nt, const int *nu,
                        int tp, int tt, int tu, const float *u_in, const float *eps_in, float4 *rec, int *mono,
                        hipStream_t s) {
  const size_t total = (size_t)nd * ng * TP * TT * TU;
  if (total == 0) return;
  const size_t want = (total + 255) / 256;
  const unsigned nblk = (unsigned)(want < 65536 ? want : 65536);
  hipLaunchKernelGGL(k_pack_tables, dim3(nblk), dim3(256), 0, s, nd, ng, TP, TT, TU, np, nt, nu, tp, tt, tu,
                     u_in, eps_in, rec, mono);
}

void launch_zero(void *p, size_t n, hipStream_t s) {
  if (n == 0) return;
  const size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)(nb < 16384 ? nb : 16384)), dim3(256), 0, s,
                     (unsigned long long *)p, n);
}

// formod_fov for one (ray, channel): neighbours ir-NFOV..ir+NFOV of the
// same package with the same time, linear interpolation in view-point
// altitude, weighted mean over the shape points.  Reads rad0/tau0 (values
// before the convolution), writes rad/tau.  pkg_lo/pkg_hi bound the rays
// of each ray's package.  err is set to 1 if fewer than 2 neighbours exist.
__global__ void k_fov(const double *rad0, const double *tau0, double *rad, double *tau, int nray,
                      int nd, const double *ray_geom, const int *pkg_lo, const int *pkg_hi,
                      const double *dz, const double *w, int n, int *err) {
  JUR_FP_EXACT
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nray * nd) return;
  const int ir = gid / nd, id = gid - ir * nd;
  const int lo = pkg_lo[ir], hi = pkg_hi[ir];
  double z[2 * NFOV + 1], r[2 * NFOV + 1], tt[2 * NFOV + 1];
  int nz = 0;
  const double time = ray_geom[7 * ir];
  const int a = (ir - NFOV > lo) ? ir - NFOV : lo;
  const int b = (ir + 1 + NFOV < hi) ? ir + 1 + NFOV : hi;
  for (int ir2 = a; ir2 < b; ir2++)
    if (ray_geom[7 * ir2] == time) {
      z[nz] = ray_geom[7 * ir2 + 4];
      r[nz] = rad0[(size_t)ir2 * nd + id];
      tt[nz] = tau0[(size_t)ir2 * nd + id];
      nz++;
    }
  if (nz < 2) {
    *err = 1;
    return;
  }
  double wsum = 0, rs = 0, ts = 0;
  const double vpz = ray_geom[7 * ir + 4];
  for (int i = 0; i < n; i++) {
    const double zfov = vpz + dz[i];
    const int idx = locate_irr(z, nz, zfov);
    rs += w[i] * lin(z[idx], r[idx], z[idx + 1], r[idx + 1], zfov);
    ts += w[i] * lin(z[idx], tt[idx], z[idx + 1], tt[idx + 1], zfov);
    wsum += w[i];
  }
  rad[(size_t)ir * nd + id] = rs / wsum;
  tau[(size_t)ir * nd + id] = ts / wsum;
}

// Brightness temperature and the observation mask.
__global__ void k_finalize(double *rad, int nray, int nd, const ChanInfo *chan, int bbt,
                           const unsigned char *mask) {
  JUR_FP_EXACT
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nray * nd) return;
  const int id = gid % nd;
  double v = rad[gid];
  if (bbt) v = bright(v, chan[id].nu);
  if (mask && mask[gid]) v = __longlong_as_double(0x7ff8000000000000LL);
  rad[gid] = v;
}

void launch_srcfunc(const double *st, const double *fnu, const double *f, const int *off,
                    const int *nf, const double *dnu, int nd, double *sr, hipStream_t s) {
  const int n = nd * TBLNS;
  hipLaunchKernelGGL(k_srcfunc, dim3((n + 255) / 256), dim3(256), 0, s, st, fnu, f, off, nf, dnu, nd, sr);
}

void launch_zrange(const atm_t *atm, int nprof, double *zr, hipStream_t s) {
  hipLaunchKernelGGL(k_zrange, dim3((nprof + 63) / 64), dim3(64), 0, s, atm, nprof, zr);
}

void launch_reduce_info(const RayInfo *info, int nray, unsigned long long *out, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, out, 3);
  if (nray > 0)
    hipLaunchKernelGGL(k_reduce_info, dim3((nray + 255) / 256), dim3(256), 0, s, info, nray, out);
}

void launch_fov(const double *rad0, const double *tau0, double *rad, double *tau, int nray, int nd,
                const double *ray_geom, const int *pkg_lo, const int *pkg_hi, const double *dz,
                const double *w, int n, int *err, hipStream_t s) {
  const int tot = nray * nd;
  hipLaunchKernelGGL(k_fov, dim3((tot + 255) / 256), dim3(256), 0, s, rad0, tau0, rad, tau, nray, nd,
                     ray_geom, pkg_lo, pkg_hi, dz, w, n, err);
}

void launch_block_np(const RayInfo *info, int nray, int blk0, int nb, int *npb, hipStream_t s) {
  hipLaunchKernelGGL(k_block_np, dim3((nb + 63) / 64), dim3(64), 0, s, info, nray, blk0, nb, npb);
}

void launch_finalize(double *rad, int nray, int nd, const ChanInfo *chan, int bbt,
                     const unsigned char *mask, hipStream_t s) {
  const int tot = nray * nd;
  if (tot > 0)
    hipLaunchKernelGGL(k_finalize, dim3((tot + 255) / 256), dim3(256), 0, s, rad, nray, nd, chan, bbt, mask);
}

}  // namespace jur
