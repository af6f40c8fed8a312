// FP32 emissivity lookups: the chains of lookup_f64.hip with FP32 table
// interpolation.
//
// Precision plan (budget: 0.01 K brightness temperature against FP64):
//  * The state of a chain is the path emissivity eps_path = 1 - tau_path.
//    The reference's update tau_path *= 1 - eps with eps = 1 - (1 - eps00) /
//    tau_path makes tau_path exactly 1 - eps00, so eps_path = eps00 carries
//    no accumulated rounding, and small emissivities keep full relative
//    precision in float where 1 - tau_path would not.
//  * The column-density side of the emissivity-growth step -- the inverse
//    lookup u(eps_path), the sum u + u_seg and the bracket search on u -- is
//    FP64: u_seg is often 1e-3 of u or less.
//  * Output is the segment emissivity (eps00 - eps_path) / (1 - eps_path);
//    1 marks an opaque segment (tau_path below float resolution).
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) los[((size_t)(f) * L.cap + (size_t)(ip)) * WAVE]

// Column lookups with float interpolation on a double column-density axis;
// boundary branches come from the walk's bracket record as in col_eps.
__device__ __forceinline__ float col_eps_f(const float4 *col, int n, double x, int mono, int &g) {
  float4 r;
  if (mono) {
    g = col_walk<0>(col, n, x, g, r);
  } else {
    const float4 r0 = col[0];
    if (x < (double)r0.x) return __fdividef(r0.y, r0.x) * (float)x;
    const float4 rl = col[n - 1];
    if (x > (double)rl.x) return 1.f - __expf(__logf(1.f - rl.y) / rl.x * (float)x);
    g = col_bisect<0>(col, n, x, r);
  }
  if (x < (double)r.x) return __fdividef(r.y, r.x) * (float)x;
  if (x > (double)r.z) return 1.f - __expf(__logf(1.f - r.w) / r.z * (float)x);
  return r.y + __fdividef(r.w - r.y, r.z - r.x) * (float)(x - (double)r.x);
}

__device__ __forceinline__ double col_u_f(const float4 *col, int n, float x, int mono, int &g) {
  float4 r;
  if (mono) {
    g = col_walk<1>(col, n, (double)x, g, r);
  } else {
    const float4 r0 = col[0];
    if (x < r0.y) return (double)(__fdividef(r0.x, r0.y) * x);
    const float4 rl = col[n - 1];
    if (x > rl.y) return (double)(log1pf(-x) * __fdividef(rl.x, __logf(1.f - rl.y)));
    g = col_bisect<1>(col, n, (double)x, r);
  }
  if (x < r.y) return (double)(__fdividef(r.x, r.y) * x);
  if (x > r.w) return (double)(log1pf(-x) * __fdividef(r.z, __logf(1.f - r.w)));
  return (double)r.x + (double)(__fdividef(r.z - r.x, r.w - r.y) * (x - r.y));
}

template <bool CGA>
__global__ __launch_bounds__(WAVE) void k_lookup_f32(ModelView M, BatchView B, LosView L,
                                                     EpsView E, const RayInfo *info, int blk0,
                                                     int nblk) {
  const int nd = M.nd, ng = M.ng;
  const int w = xcd_remap(blockIdx.x, nblk * nd * ng);
  const int bl = w / (nd * ng);
  const int rem = w - bl * nd * ng;
  const int id = rem / ng;
  const int ig = rem - id * ng;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const TabView &T = M.tab;
  const int tix = ig * nd + id;
  const int npt = T.np[tix];
  if (npt < 30) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  const double *los = L.los + (size_t)bl * L.nfield * L.cap * WAVE + lane;
  float *out = (float *)E.eps + (((size_t)bl * nd + id) * ng + ig) * E.cap * WAVE + lane;
  const int mono = T.mono[tix];
  const int pdir = T.pdir[tix];
  const double *pg = T.p + (size_t)tix * T.TP;
  const int *ntp = T.nt + (size_t)tix * T.TP;

  float ep = 0.f;
  double cu = 0, cp = 0, ct = 0;
  int gi[8];
#pragma unroll
  for (int c = 0; c < 8; c++) gi[c] = T.TU / 2;
  int ipr = -1;
  int ip = 0;
  for (; ip < R.np; ip++) {
    const double p = LOSR(F_P, ip);
    const double t = LOSR(F_T, ip);
    const double useg = LOSR(L.f_u + ig, ip);
    if (CGA) {
      cu += useg;
      cp += useg * p;
      ct += useg * t;
    }
    const double pq = CGA ? cp / cu : p;
    const double tq = CGA ? ct / cu : t;
    ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
    const int nt0 = ntp[ipr], nt1 = ntp[ipr + 1];
    float e = 0.f;
    if (nt0 >= 2 && nt1 >= 2) {
      const size_t c0 = (size_t)tix * T.TP + ipr;
      const double *tg0 = T.t + c0 * T.TT;
      const double *tg1 = tg0 + T.TT;
      const int it0 = locate_reg(tg0, nt0, tq);
      const int it1 = locate_reg(tg1, nt1, tq);
      const int *nu0 = T.nu + c0 * T.TT;
      const int *nu1 = nu0 + T.TT;
      const int n00 = nu0[it0], n01 = nu0[it0 + 1], n10 = nu1[it1], n11 = nu1[it1 + 1];
      if (n00 >= 2 && n01 >= 2 && n10 >= 2 && n11 >= 2) {
        const float4 *k00 = T.rec + (c0 * T.TT + it0) * T.TU;
        const float4 *k01 = k00 + T.TU;
        const float4 *k10 = T.rec + ((c0 + 1) * T.TT + it1) * T.TU;
        const float4 *k11 = k10 + T.TU;
        float e00, e01, e10, e11;
        if (CGA) {
          e00 = col_eps_f(k00, n00, cu, mono, gi[4]);
          e01 = col_eps_f(k01, n01, cu, mono, gi[5]);
          e10 = col_eps_f(k10, n10, cu, mono, gi[6]);
          e11 = col_eps_f(k11, n11, cu, mono, gi[7]);
        } else {
          e00 = col_eps_f(k00, n00, col_u_f(k00, n00, ep, mono, gi[0]) + useg, mono, gi[4]);
          e01 = col_eps_f(k01, n01, col_u_f(k01, n01, ep, mono, gi[1]) + useg, mono, gi[5]);
          e10 = col_eps_f(k10, n10, col_u_f(k10, n10, ep, mono, gi[2]) + useg, mono, gi[6]);
          e11 = col_eps_f(k11, n11, col_u_f(k11, n11, ep, mono, gi[3]) + useg, mono, gi[7]);
        }
        const float w0 = (float)((tq - tg0[it0]) / (tg0[it0 + 1] - tg0[it0]));
        const float w1 = (float)((tq - tg1[it1]) / (tg1[it1 + 1] - tg1[it1]));
        const float a = e00 + (e01 - e00) * w0;
        const float b = e10 + (e11 - e10) * w1;
        float wp;
        if (CGA) {
          const double x0 = pg[ipr], x1 = pg[ipr + 1];
          wp = (pq / x0 > 0 && x1 / x0 > 0) ? (float)(log(pq / x0) / log(x1 / x0))
                                             : (float)((pq - x0) / (x1 - x0));
        } else {
          wp = (float)((pq - pg[ipr]) / (pg[ipr + 1] - pg[ipr]));
        }
        float en = a + (b - a) * wp;
        en = (en < 1.f) ? en : 1.f;
        en = (en > 0.f) ? en : 0.f;
        e = (en >= 1.f) ? 1.f : (en - ep) / (1.f - ep);
        ep = en;
      }
    }
    out[(size_t)ip * WAVE] = e;
    if (ep >= 1.f) {
      // tau_path below float resolution: every later segment is opaque too
      for (ip++; ip < R.np; ip++) out[(size_t)ip * WAVE] = 1.f;
      break;
    }
  }
}

void launch_lookup_f32(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                       const RayInfo *info, int blk0, int nblk, hipStream_t s) {
  const dim3 grid(nblk * M.nd * M.ng), blk(WAVE);
  if (M.nd * M.ng == 0) return;
  if (M.formod == 0)
    hipLaunchKernelGGL(k_lookup_f32<true>, grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  else
    hipLaunchKernelGGL(k_lookup_f32<false>, grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
}

}  // namespace jur
