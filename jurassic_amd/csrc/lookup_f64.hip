// FP64 emissivity lookups (intpol_tbl_ega, src/jurassic.c:3799-3891, and
// intpol_tbl_cga, :3710-3795), one independent chain per (ray, channel,
// emitter).
//
// The EGA recurrence couples the segments of one ray only through each
// emitter's own path transmittance tau_path[id][ig]; emitters and channels are
// independent.  So one wavefront handles one (block of 64 rays, channel,
// emitter), a lane per ray, and writes the segment emissivity eps of every
// LOS point to the emissivity buffer; the radiance kernel combines emitters
// as the reference does (tau_seg *= 1 - eps, emitter order).  A lane keeps
// only tau_path and the bracket indices of its four table corners, so many
// wavefronts stay resident to hide the latency of the table gathers.
//
// Built with -ffp-contract=off; arithmetic follows the reference.
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) los[((size_t)(f) * L.cap + (size_t)(ip)) * WAVE]

template <bool CGA>
__global__ __launch_bounds__(WAVE) void k_lookup_f64(ModelView M, BatchView B, LosView L,
                                                     EpsView E, const RayInfo *info, int blk0,
                                                     int nblk) {
  JUR_FP_EXACT
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
  if (npt < 30) return;  // eps = 0 for every segment; the radiance kernel knows
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  const double *los = L.los + (size_t)bl * L.nfield * L.cap * WAVE + lane;
  double *out = (double *)E.eps + (((size_t)bl * nd + id) * ng + ig) * E.cap * WAVE + lane;
  const int mono = T.mono[tix];
  const int pdir = T.pdir[tix];
  const double *pg = T.p + (size_t)tix * T.TP;
  const int *ntp = T.nt + (size_t)tix * T.TP;

  double tp = 1, cu = 0, cp = 0, ct = 0;
  int gi[8];
#pragma unroll
  for (int c = 0; c < 8; c++) gi[c] = T.TU / 2;
  int ipr = -1;

  for (int ip = 0; ip < R.np; ip++) {
    const double p = LOSR(F_P, ip);
    const double t = LOSR(F_T, ip);
    const double useg = LOSR(L.f_u + ig, ip);
    if (CGA) {
      cu = cu + useg;
      cp = cp + useg * p;
      ct = ct + useg * t;
    }
    double eps;
    if (tp < 1e-9) {
      eps = 1;
    } else {
      const double pq = CGA ? cp / cu : p;
      const double tq = CGA ? ct / cu : t;
      ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
      const int nt0 = ntp[ipr], nt1 = ntp[ipr + 1];
      if (nt0 < 2 || nt1 < 2) {
        eps = 0;
      } else {
        const size_t c0 = (size_t)tix * T.TP + ipr;
        const double *tg0 = T.t + c0 * T.TT;
        const double *tg1 = tg0 + T.TT;
        const int it0 = locate_reg(tg0, nt0, tq);
        const int it1 = locate_reg(tg1, nt1, tq);
        const int *nu0 = T.nu + c0 * T.TT;
        const int *nu1 = nu0 + T.TT;
        const int n00 = nu0[it0], n01 = nu0[it0 + 1], n10 = nu1[it1], n11 = nu1[it1 + 1];
        if (n00 < 2 || n01 < 2 || n10 < 2 || n11 < 2) {
          eps = 0;
        } else {
          const float4 *k00 = T.rec + (c0 * T.TT + it0) * T.TU;
          const float4 *k01 = k00 + T.TU;
          const float4 *k10 = T.rec + ((c0 + 1) * T.TT + it1) * T.TU;
          const float4 *k11 = k10 + T.TU;
          double e00, e01, e10, e11;
          if (CGA) {
            e00 = col_eps(k00, n00, cu, mono, gi[4]);
            e01 = col_eps(k01, n01, cu, mono, gi[5]);
            e10 = col_eps(k10, n10, cu, mono, gi[6]);
            e11 = col_eps(k11, n11, cu, mono, gi[7]);
          } else {
            const double x = 1 - tp;
            e00 = col_eps(k00, n00, col_u(k00, n00, x, mono, gi[0]) + useg, mono, gi[4]);
            e01 = col_eps(k01, n01, col_u(k01, n01, x, mono, gi[1]) + useg, mono, gi[5]);
            e10 = col_eps(k10, n10, col_u(k10, n10, x, mono, gi[2]) + useg, mono, gi[6]);
            e11 = col_eps(k11, n11, col_u(k11, n11, x, mono, gi[3]) + useg, mono, gi[7]);
          }
          e00 = lin(tg0[it0], e00, tg0[it0 + 1], e01, tq);
          e11 = lin(tg1[it1], e10, tg1[it1 + 1], e11, tq);
          e00 = CGA ? logx(pg[ipr], e00, pg[ipr + 1], e11, pq) : lin(pg[ipr], e00, pg[ipr + 1], e11, pq);
          e00 = (e00 < 1) ? e00 : 1;  // MAX(MIN(eps00, 1), 0)
          e00 = (e00 > 0) ? e00 : 0;
          eps = 1 - (1 - e00) / tp;
        }
      }
    }
    tp *= (1 - eps);
    out[(size_t)ip * WAVE] = eps;
  }
}

void launch_lookup_f64(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                       const RayInfo *info, int blk0, int nblk, hipStream_t s) {
  const dim3 grid(nblk * M.nd * M.ng), blk(WAVE);
  if (M.nd * M.ng == 0) return;
  if (M.formod == 0)
    hipLaunchKernelGGL(k_lookup_f64<true>, grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  else
    hipLaunchKernelGGL(k_lookup_f64<false>, grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
}

}  // namespace jur
