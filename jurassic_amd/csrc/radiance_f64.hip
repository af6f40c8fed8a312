// FP64 radiance kernel: emissivity-growth (EGA) or Curtis-Godson (CGA)
// table lookups, continua, source function and radiance accumulation along
// each line of sight (formod_pencil, src/jurassic.c:3196-3338).
//
// Work decomposition: one wavefront per (block of 64 rays, channel); lane =
// ray.  All lanes of a wavefront read the same channel's tables, so lanes
// whose rays follow similar paths share cache lines; the segment data of the
// 64 rays are read with one coalesced access per field.  Each lane walks its
// segments in order (the EGA recurrence is sequential along the ray) and
// keeps, per emitter and table corner, the bracket index of the previous
// segment as the start of the next search.
//
// Built with -ffp-contract=off; operation order follows the reference, so
// results agree with the reference to rounding.
#include "jur_device.hpp"

namespace jur {

struct GasState {
  double tp;        // path transmittance tau_path of this emitter
  int gi[8];        // bracket indices: inverse lookups c00..c11, forward c00..c11
  double cu, cp, ct; // Curtis-Godson running sums
};

// intpol_tbl_ega / intpol_tbl_cga for one (channel, emitter) and one
// segment (src/jurassic.c:3799-3891 and 3710-3795).  Returns eps.
template <bool CGA>
__device__ __forceinline__ double seg_eps(const TabView &T, int tix, double p, double t, double useg,
                                          GasState &s) {
  JUR_FP_EXACT
  const int npt = T.np[tix];
  if (npt < 30) return 0;
  if (s.tp < 1e-9) return 1;
  double pq = p, tq = t;
  if (CGA) {
    pq = s.cp / s.cu;
    tq = s.ct / s.cu;
  }
  const double *pg = T.p + (size_t)tix * T.TP;
  const int ipr = locate_irr(pg, npt, pq);
  const int *ntp = T.nt + (size_t)tix * T.TP;
  const size_t c0 = (size_t)tix * T.TP + ipr;
  const double *tg0 = T.t + c0 * T.TT;
  const double *tg1 = tg0 + T.TT;
  // the reference tests all six sizes together; testing the T sizes first
  // keeps the u-size reads in range and gives the same eps = 0
  if (ntp[ipr] < 2 || ntp[ipr + 1] < 2) return 0;
  const int it0 = locate_reg(tg0, ntp[ipr], tq);
  const int it1 = locate_reg(tg1, ntp[ipr + 1], tq);
  const int *nu0 = T.nu + c0 * T.TT;
  const int *nu1 = nu0 + T.TT;
  const int n00 = nu0[it0], n01 = nu0[it0 + 1], n10 = nu1[it1], n11 = nu1[it1 + 1];
  if (n00 < 2 || n01 < 2 || n10 < 2 || n11 < 2) return 0;
  const float4 *k00 = T.rec + (c0 * T.TT + it0) * T.TU;
  const float4 *k01 = k00 + T.TU;
  const float4 *k10 = T.rec + ((c0 + 1) * T.TT + it1) * T.TU;
  const float4 *k11 = k10 + T.TU;
  const int mono = T.mono[tix];
  double e00, e01, e10, e11;
  if (CGA) {
    const double cu = s.cu;
    e00 = col_eps(k00, n00, cu, mono, s.gi[4]);
    e01 = col_eps(k01, n01, cu, mono, s.gi[5]);
    e10 = col_eps(k10, n10, cu, mono, s.gi[6]);
    e11 = col_eps(k11, n11, cu, mono, s.gi[7]);
  } else {
    const double x = 1 - s.tp;
    double u;
    u = col_u(k00, n00, x, mono, s.gi[0]);
    e00 = col_eps(k00, n00, u + useg, mono, s.gi[4]);
    u = col_u(k01, n01, x, mono, s.gi[1]);
    e01 = col_eps(k01, n01, u + useg, mono, s.gi[5]);
    u = col_u(k10, n10, x, mono, s.gi[2]);
    e10 = col_eps(k10, n10, u + useg, mono, s.gi[6]);
    u = col_u(k11, n11, x, mono, s.gi[3]);
    e11 = col_eps(k11, n11, u + useg, mono, s.gi[7]);
  }
  e00 = lin(tg0[it0], e00, tg0[it0 + 1], e01, tq);
  e11 = lin(tg1[it1], e10, tg1[it1 + 1], e11, tq);
  e00 = CGA ? logx(pg[ipr], e00, pg[ipr + 1], e11, pq) : lin(pg[ipr], e00, pg[ipr + 1], e11, pq);
  e00 = (e00 < 1) ? e00 : 1;  // MAX(MIN(eps00, 1), 0)
  e00 = (e00 > 0) ? e00 : 0;
  return 1 - (1 - e00) / s.tp;
}

#define LOSR(f, ip) base[((size_t)(f) * L.cap + (size_t)(ip)) * WAVE]

template <int NGAS, bool CGA>
__global__ __launch_bounds__(WAVE) void k_radiance_f64(ModelView M, BatchView B, LosView L,
                                                       const RayInfo *info, int blk0, int nblk,
                                                       double *rad_out, double *tau_out) {
  JUR_FP_EXACT
  const int nd = M.nd;
  const int w = xcd_remap(blockIdx.x, nblk * nd);
  const int bl = w / nd;
  const int id = w - bl * nd;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  const atm_t *atm = &B.atm[B.ray_prof[ir]];
  const ChanInfo C = M.chan[id];
  const double nu = C.nu;
  const double *base = L.los + (size_t)bl * L.nfield * L.cap * WAVE + lane;
  const double *sr_col = M.sr + (size_t)id * TBLNS;

  double kcl = 0;
  if (L.f_cl >= 0 && M.ncl > 0)
    kcl = lin(M.clnu[C.icl], atm->clk[C.icl], M.clnu[C.icl + 1], atm->clk[C.icl + 1], nu);
  double sfeps = 1.0;
  if (M.nsf > 0)
    sfeps = lin(M.sfnu[C.isf], atm->sfeps[C.isf], M.sfnu[C.isf + 1], atm->sfeps[C.isf + 1], nu);

  constexpr int NA = NGAS > 0 ? NGAS : 1;
  GasState gs[NA];
#pragma unroll
  for (int g = 0; g < NGAS; g++) {
    gs[g].tp = 1;
    gs[g].cu = gs[g].cp = gs[g].ct = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) gs[g].gi[c] = M.tab.TU / 2;
  }

  double rad = 0, tau = 1, drefl = 0, tau_refl = 1;
  double ds_prev = 0;
  for (int ip = 0; ip < R.np; ip++) {
    const double p = LOSR(F_P, ip);
    const double t = LOSR(F_T, ip);
    const double dsr = LOSR(F_DS, ip);
    // trapezoid rule (src/jurassic.c:4410-4413)
    const double ds = (ip == 0) ? dsr * 0.5 : 0.5 * (ds_prev + dsr);
    ds_prev = dsr;

    double q[NA], u[NA];
#pragma unroll
    for (int g = 0; g < NGAS; g++) {
      q[g] = LOSR(F_Q0 + g, ip);
      u[g] = 10 * q[g] * p / (JUR_KB * t) * ds;  // column density (src/jurassic.c:4418-4419)
      if (CGA) {
        gs[g].cu = gs[g].cu + u[g];
        gs[g].cp = gs[g].cp + u[g] * p;
        gs[g].ct = gs[g].ct + u[g] * t;
      }
    }

    double tau_gas = 1;
#pragma unroll
    for (int g = 0; g < NGAS; g++) {
      const double eps = seg_eps<CGA>(M.tab, g * nd + id, p, t, u[g], gs[g]);
      gs[g].tp *= (1 - eps);
      tau_gas *= (1 - eps);
    }

    // formod_continua (src/jurassic.c:3075-3117)
    double beta = (C.window < M.nw) ? LOSR(L.f_k + C.window, ip) : 0.0;
    if (L.f_cl >= 0) beta += LOSR(L.f_cl, ip) * kcl;
    if (M.ig_co2 >= 0) {
      double uc = 0;
#pragma unroll
      for (int g = 0; g < NGAS; g++)
        if (g == M.ig_co2) uc = u[g];
      beta += ctmco2(M.ctm, nu, p, t, uc) / ds;
    }
    if (M.ig_h2o >= 0) {
      double uh = 0, qh = 0;
#pragma unroll
      for (int g = 0; g < NGAS; g++)
        if (g == M.ig_h2o) {
          uh = u[g];
          qh = q[g];
        }
      beta += ctmh2o(M.ctm, nu, p, t, qh, uh) / ds;
    }
    if (M.ctm_n2) beta += ctmn2(M.ctm, nu, p, t);
    if (M.ctm_o2) beta += ctmo2(M.ctm, nu, p, t);

    const double src = srcfunc(M.st, sr_col, t);

    if (tau_gas > 0) {
      const double eps = 1 - tau_gas * exp(-beta * ds);
      rad += src * eps * tau;
      tau *= (1 - eps);
      // down-welling sum of the reflection term, accumulated forward:
      // D = sum_i src_i eps_i prod_{k>i} (1 - eps_k)
      drefl = drefl * (1 - eps) + src * eps;
      tau_refl *= (1 - eps);
    }
  }

  // surface emission, reflection and solar term (src/jurassic.c:3262-3328)
  if (M.sftype >= 1 && R.sft > 0) {
    const double src_sf = srcfunc(M.st, sr_col, R.sft);
    rad += sfeps * src_sf * tau;
    if (M.sftype >= 2 && sfeps < 1) {
      rad += drefl * tau * (1 - sfeps);
      if (M.sftype >= 3 && R.pad > 0)
        rad += 6.764e-5 / (2. * JUR_PI) * C.planck_sun * tau_refl * (1 - sfeps) * tau * R.rcos;
    }
  }
  rad_out[(size_t)ir * nd + id] = rad;
  tau_out[(size_t)ir * nd + id] = tau;
}

template <int NGAS>
static void launch_ng(const ModelView &M, const BatchView &B, const LosView &L, const RayInfo *info,
                      int blk0, int nblk, double *rad, double *tau, hipStream_t s) {
  const dim3 grid(nblk * M.nd), blk(WAVE);
  if (M.formod == 0)
    hipLaunchKernelGGL((k_radiance_f64<NGAS, true>), grid, blk, 0, s, M, B, L, info, blk0, nblk, rad, tau);
  else
    hipLaunchKernelGGL((k_radiance_f64<NGAS, false>), grid, blk, 0, s, M, B, L, info, blk0, nblk, rad, tau);
}

void launch_radiance_f64(const ModelView &M, const BatchView &B, const LosView &L,
                         const RayInfo *info, int blk0, int nblk, double *rad, double *tau,
                         hipStream_t s) {
  switch (M.ng) {
    case 0: launch_ng<0>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 1: launch_ng<1>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 2: launch_ng<2>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 3: launch_ng<3>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 4: launch_ng<4>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 5: launch_ng<5>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 6: launch_ng<6>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    case 7: launch_ng<7>(M, B, L, info, blk0, nblk, rad, tau, s); break;
    default: launch_ng<8>(M, B, L, info, blk0, nblk, rad, tau, s); break;
  }
}

}  // namespace jur
