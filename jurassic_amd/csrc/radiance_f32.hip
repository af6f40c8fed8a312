// FP32 radiance kernel: the same EGA/CGA recurrence as radiance_f64.hip
// with FP32 table interpolation, continua, source function and accumulation.
//
// Precision plan (error budget 0.01 K in brightness temperature against the
// FP64 path):
//  * The EGA state is carried as the path emissivity eps_path = 1 - tau_path
//    (the reference computes tau_path *= 1 - eps with eps = 1 - (1 - eps00) /
//    tau_path, i.e. tau_path becomes exactly 1 - eps00).  Small emissivities
//    keep full relative precision in float, where 1 - tau_path would not.
//  * The inverse lookup u(eps_path) and the forward lookup eps(u + u_seg) keep
//    their column-density arithmetic (the bracket search on u, u_eq + u_seg)
//    in FP64: u_seg is often 1e-3 of u_eq, below float resolution otherwise.
//  * Segment emissivities combine as 1 - prod(1 - e_g) in the cancellation-
//    free form e + e_g - e e_g; radiance and transmittance are accumulated
//    with compensated (Kahan / two-sum) arithmetic.
// Ray geometry is the FP64 ray tracer's.
#include "jur_device.hpp"

namespace jur {

struct GasStateF {
  float ep;          // path emissivity 1 - tau_path
  int gi[8];
  double cu, cp, ct; // Curtis-Godson running sums (FP64: sums of up to 4096 terms)
};

// Column lookups with float interpolation and a double column-density axis.
__device__ __forceinline__ float col_eps_f(const float4 *col, int n, double x, int mono, int &g) {
  const float4 r0 = col[0];
  if (x < (double)r0.x) return (r0.y / r0.x) * (float)x;
  const float4 rl = col[n - 1];
  if (x > (double)rl.x) {
    const float a = __logf(1.f - rl.y) / rl.x;
    return 1.f - __expf(a * (float)x);
  }
  float4 r;
  g = mono ? col_walk<0>(col, n, x, g, r) : col_bisect<0>(col, n, x, r);
  const float s = (r.w - r.y) / (r.z - r.x);
  return r.y + s * (float)(x - (double)r.x);
}

__device__ __forceinline__ double col_u_f(const float4 *col, int n, float x, int mono, int &g) {
  const float4 r0 = col[0];
  if (x < r0.y) return (double)(r0.x / r0.y) * (double)x;
  const float4 rl = col[n - 1];
  if (x > rl.y) {
    const float a = __logf(1.f - rl.y) / rl.x;
    return (double)(log1pf(-x) / a);
  }
  float4 r;
  g = mono ? col_walk<1>(col, n, (double)x, g, r) : col_bisect<1>(col, n, (double)x, r);
  const float s = (r.z - r.x) / (r.w - r.y);
  return (double)r.x + (double)(s * (x - r.y));
}

// One (channel, emitter) segment update.  Returns the segment emissivity of
// the emitter and updates its path emissivity; *sat is set when the segment
// is opaque for this emitter (the reference's tau_gas == 0 case).
template <bool CGA>
__device__ __forceinline__ float seg_eps_f(const TabView &T, int tix, double p, double t, double useg,
                                           GasStateF &s, bool *sat) {
  const int npt = T.np[tix];
  if (npt < 30) return 0.f;
  if (s.ep >= 1.f) {  // tau_path below float resolution: eps = 1
    *sat = true;
    return 1.f;
  }
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
  if (ntp[ipr] < 2 || ntp[ipr + 1] < 2) return 0.f;
  const int it0 = locate_reg(tg0, ntp[ipr], tq);
  const int it1 = locate_reg(tg1, ntp[ipr + 1], tq);
  const int *nu0 = T.nu + c0 * T.TT;
  const int *nu1 = nu0 + T.TT;
  const int n00 = nu0[it0], n01 = nu0[it0 + 1], n10 = nu1[it1], n11 = nu1[it1 + 1];
  if (n00 < 2 || n01 < 2 || n10 < 2 || n11 < 2) return 0.f;
  const float4 *k00 = T.rec + (c0 * T.TT + it0) * T.TU;
  const float4 *k01 = k00 + T.TU;
  const float4 *k10 = T.rec + ((c0 + 1) * T.TT + it1) * T.TU;
  const float4 *k11 = k10 + T.TU;
  const int mono = T.mono[tix];
  float e00, e01, e10, e11;
  if (CGA) {
    e00 = col_eps_f(k00, n00, s.cu, mono, s.gi[4]);
    e01 = col_eps_f(k01, n01, s.cu, mono, s.gi[5]);
    e10 = col_eps_f(k10, n10, s.cu, mono, s.gi[6]);
    e11 = col_eps_f(k11, n11, s.cu, mono, s.gi[7]);
  } else {
    const float x = s.ep;
    e00 = col_eps_f(k00, n00, col_u_f(k00, n00, x, mono, s.gi[0]) + useg, mono, s.gi[4]);
    e01 = col_eps_f(k01, n01, col_u_f(k01, n01, x, mono, s.gi[1]) + useg, mono, s.gi[5]);
    e10 = col_eps_f(k10, n10, col_u_f(k10, n10, x, mono, s.gi[2]) + useg, mono, s.gi[6]);
    e11 = col_eps_f(k11, n11, col_u_f(k11, n11, x, mono, s.gi[3]) + useg, mono, s.gi[7]);
  }
  const float wt0 = (float)((tq - tg0[it0]) / (tg0[it0 + 1] - tg0[it0]));
  const float wt1 = (float)((tq - tg1[it1]) / (tg1[it1 + 1] - tg1[it1]));
  float a = e00 + (e01 - e00) * wt0;
  float b = e10 + (e11 - e10) * wt1;
  float w;
  if (CGA) {
    const double x0 = pg[ipr], x1 = pg[ipr + 1];
    w = (pq / x0 > 0 && x1 / x0 > 0) ? (float)(log(pq / x0) / log(x1 / x0))
                                      : (float)((pq - x0) / (x1 - x0));
  } else {
    w = (float)((pq - pg[ipr]) / (pg[ipr + 1] - pg[ipr]));
  }
  float e = a + (b - a) * w;
  e = (e < 1.f) ? e : 1.f;
  e = (e > 0.f) ? e : 0.f;
  const float ep0 = s.ep;
  s.ep = e;
  if (e >= 1.f) {
    *sat = true;
    return 1.f;
  }
  // segment emissivity 1 - (1 - e)/(1 - ep0) = (e - ep0)/(1 - ep0)
  return (e - ep0) / (1.f - ep0);
}

#define LOSR(f, ip) base[((size_t)(f) * L.cap + (size_t)(ip)) * WAVE]

// Kahan-compensated sum.
struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

template <int NGAS, bool CGA>
__global__ __launch_bounds__(WAVE) void k_radiance_f32(ModelView M, BatchView B, LosView L,
                                                       const RayInfo *info, int blk0, int nblk,
                                                       double *rad_out, double *tau_out) {
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
  const float nuf = (float)nu;
  const double *base = L.los + (size_t)bl * L.nfield * L.cap * WAVE + lane;
  const double *sr_col = M.sr + (size_t)id * TBLNS;

  float kcl = 0.f;
  if (L.f_cl >= 0 && M.ncl > 0)
    kcl = (float)lin(M.clnu[C.icl], atm->clk[C.icl], M.clnu[C.icl + 1], atm->clk[C.icl + 1], nu);
  float sfeps = 1.f;
  if (M.nsf > 0)
    sfeps = (float)lin(M.sfnu[C.isf], atm->sfeps[C.isf], M.sfnu[C.isf + 1], atm->sfeps[C.isf + 1], nu);

  // channel constants of the continua (src/jurassic.c:1724-1741, 2762-2791,
  // 2841-2917), evaluated once in FP64
  bool co2_on = false, h2o_on = false, n2_on = false, o2_on = false;
  float co2_a = 0, co2_b = 0, co2_c = 0;      // ctw(t) = a dt260 dt296 - b dt230 dt296 + c dt230 dt260
  float h2o_slf = 0, h2o_l2r = 0, h2o_frn = 0;
  float n2_b = 0, n2_beta = 0, o2_b = 0, o2_beta = 0;
  if (M.ig_co2 >= 0) {
    const double xw = nu / 2 + 1;
    if (xw >= 1 && xw < 2001) {
      const int iw = (int)xw;
      const double dw = xw - iw, ew = 1 - dw;
      co2_on = true;
      co2_a = (float)(5.050505e-4 * (ew * M.ctm.co2_230[iw - 1] + dw * M.ctm.co2_230[iw]));
      co2_b = (float)(9.259259e-4 * (ew * M.ctm.co2_260[iw - 1] + dw * M.ctm.co2_260[iw]));
      co2_c = (float)(4.208754e-4 * (ew * M.ctm.co2_296[iw - 1] + dw * M.ctm.co2_296[iw]));
    }
  }
  if (M.ig_h2o >= 0) {
    const double xw = nu / 10 + 1;
    if (xw >= 1 && xw < 2001) {
      const int iw = (int)xw;
      const double dw = xw - iw, ew = 1 - dw;
      const double cw296 = ew * M.ctm.h2o_296[iw - 1] + dw * M.ctm.h2o_296[iw];
      const double cw260 = ew * M.ctm.h2o_260[iw - 1] + dw * M.ctm.h2o_260[iw];
      const double cwfrn = ew * M.ctm.h2o_frn[iw - 1] + dw * M.ctm.h2o_frn[iw];
      double sfac = 1;
      if (!(nu <= 820 || nu >= 960)) {
        const double xx = (nu - 820) / 10;
        const int ix = (int)xx;
        const double dx = xx - ix;
        sfac = (1 - dx) * M.ctm.h2o_xfcrev[ix] + dx * M.ctm.h2o_xfcrev[ix + 1];
      }
      const double vf2 = (nu - 370) * (nu - 370), vf6 = vf2 * vf2 * vf2;
      const double fscal = 36100 / (vf2 + vf6 * 1e-8 + 36100) * -.25 + 1;
      h2o_on = true;
      h2o_slf = (float)(sfac * cw296);
      h2o_l2r = (float)(log2(cw260 / cw296) / 36.0);
      h2o_frn = (float)(cwfrn * fscal);
    }
  }
  if (M.ctm_n2 && !(nu < M.ctm.n2_nu[0] || nu > M.ctm.n2_nu[97])) {
    const int idx = locate_reg(M.ctm.n2_nu, 98, nu);
    n2_on = true;
    n2_b = (float)lin(M.ctm.n2_nu[idx], M.ctm.n2_b[idx], M.ctm.n2_nu[idx + 1], M.ctm.n2_b[idx + 1], nu);
    n2_beta = (float)lin(M.ctm.n2_nu[idx], M.ctm.n2_beta[idx], M.ctm.n2_nu[idx + 1], M.ctm.n2_beta[idx + 1], nu);
  }
  if (M.ctm_o2 && !(nu < M.ctm.o2_nu[0] || nu > M.ctm.o2_nu[89])) {
    const int idx = locate_reg(M.ctm.o2_nu, 90, nu);
    o2_on = true;
    o2_b = (float)lin(M.ctm.o2_nu[idx], M.ctm.o2_b[idx], M.ctm.o2_nu[idx + 1], M.ctm.o2_b[idx + 1], nu);
    o2_beta = (float)lin(M.ctm.o2_nu[idx], M.ctm.o2_beta[idx], M.ctm.o2_nu[idx + 1], M.ctm.o2_beta[idx + 1], nu);
  }
  const float st0 = (float)M.st[0];
  const float inv_dst = (float)(1.0 / (M.st[1] - M.st[0]));

  constexpr int NA = NGAS > 0 ? NGAS : 1;
  GasStateF gs[NA];
#pragma unroll
  for (int g = 0; g < NGAS; g++) {
    gs[g].ep = 0.f;
    gs[g].cu = gs[g].cp = gs[g].ct = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) gs[g].gi[c] = M.tab.TU / 2;
  }

  KSum rad;
  float tau_hi = 1.f, tau_lo = 0.f;   // two-sum compensated transmittance
  float drefl = 0.f, tau_refl = 1.f;
  double ds_prev = 0;
  bool dead = false;
  for (int ip = 0; ip < R.np && !dead; ip++) {
    const double p = LOSR(F_P, ip);
    const double t = LOSR(F_T, ip);
    const double dsr = LOSR(F_DS, ip);
    const double ds = (ip == 0) ? dsr * 0.5 : 0.5 * (ds_prev + dsr);
    ds_prev = dsr;

    double u[NA];
    float qf[NA];
#pragma unroll
    for (int g = 0; g < NGAS; g++) {
      const double q = LOSR(F_Q0 + g, ip);
      qf[g] = (float)q;
      u[g] = 10 * q * p / (JUR_KB * t) * ds;
      if (CGA) {
        gs[g].cu += u[g];
        gs[g].cp += u[g] * p;
        gs[g].ct += u[g] * t;
      }
    }

    bool sat = false;
    float etot = 0.f;  // 1 - prod_g (1 - e_g), cancellation-free
#pragma unroll
    for (int g = 0; g < NGAS; g++) {
      const float e = seg_eps_f<CGA>(M.tab, g * nd + id, p, t, u[g], gs[g], &sat);
      etot = etot + e - etot * e;
    }
    if (sat) {
      // tau_gas == 0: the reference skips this and (its path transmittance
      // staying below 1e-9) every later segment
      dead = true;
      break;
    }

    const float pf = (float)p, tf = (float)t, dsf = (float)ds;
    float beta = (C.window < M.nw) ? (float)LOSR(L.f_k + C.window, ip) : 0.f;
    if (L.f_cl >= 0) beta += (float)LOSR(L.f_cl, ip) * kcl;
    if (co2_on) {
      float uc = 0.f;
#pragma unroll
      for (int g = 0; g < NGAS; g++)
        if (g == M.ig_co2) uc = (float)u[g];
      const float d230 = tf - 230.f, d260 = tf - 260.f, d296 = tf - 296.f;
      const float ctw = co2_a * d260 * d296 - co2_b * d230 * d296 + co2_c * d230 * d260;
      beta += uc * pf * (float)(1.0 / (JUR_NA * 1000.0 * JUR_P0)) * ctw / dsf;
    }
    if (h2o_on) {
      float uh = 0.f, qh = 0.f;
#pragma unroll
      for (int g = 0; g < NGAS; g++)
        if (g == M.ig_h2o) {
          uh = (float)u[g];
          qh = qf[g];
        }
      const float slf = h2o_slf * exp2f(h2o_l2r * (296.f - tf));
      const float a1 = nuf * uh * tanhf(0.7193876f / tf * nuf);
      const float a3 = pf * (float)(1.0 / JUR_P0) * (qh * slf + (1.f - qh) * h2o_frn) * 1e-20f;
      beta += a1 * (296.f / tf) * a3 / dsf;
    }
    if (n2_on) {
      const float r = pf * (float)(273.0 / JUR_P0) / tf;
      beta += 0.1f * r * r * __expf(n2_beta * (1.f / 296.f - 1.f / tf)) * 0.79f * n2_b *
              (0.79f + 0.21f * (1.294f - 0.4545f * tf / 296.f));
    }
    if (o2_on) {
      const float r = pf * (float)(273.0 / JUR_P0) / tf;
      beta += 0.1f * r * r * __expf(o2_beta * (1.f / 296.f - 1.f / tf)) * 0.21f * o2_b;
    }

    // source function: linear in the uniform 1200-point temperature grid
    int it = (int)((tf - st0) * inv_dst);
    it = it < 0 ? 0 : (it > TBLNS - 2 ? TBLNS - 2 : it);
    const float s0 = (float)sr_col[it], s1 = (float)sr_col[it + 1];
    const float src = s0 + (s1 - s0) * ((tf - (float)M.st[it]) * inv_dst);

    // e = 1 - (1 - etot) exp(-beta ds)
    const float c = -expm1f(-beta * dsf);
    const float e = etot + c - etot * c;
    const float d = tau_hi * e;          // tau_{i-1} - tau_i
    rad.add(src * d);
    {  // tau -= d with a two-sum error term
      const float s = tau_hi - d;
      const float bb = s - tau_hi;
      tau_lo += (tau_hi - (s - bb)) + (-d - bb);
      tau_hi = s;
    }
    drefl = drefl - drefl * e + src * e;
    tau_refl = tau_refl - tau_refl * e;
  }

  double radd = (double)rad.s - (double)rad.c;
  const double taud = (double)tau_hi + (double)tau_lo;
  if (M.sftype >= 1 && R.sft > 0) {
    const double src_sf = srcfunc(M.st, sr_col, R.sft);
    radd += (double)sfeps * src_sf * taud;
    if (M.sftype >= 2 && sfeps < 1.f) {
      radd += (double)drefl * taud * (1.0 - sfeps);
      if (M.sftype >= 3 && R.pad > 0)
        radd += 6.764e-5 / (2. * JUR_PI) * C.planck_sun * tau_refl * (1.0 - sfeps) * taud * R.rcos;
    }
  }
  rad_out[(size_t)ir * nd + id] = radd;
  tau_out[(size_t)ir * nd + id] = taud;
}

template <int NGAS>
static void launch_ng(const ModelView &M, const BatchView &B, const LosView &L, const RayInfo *info,
                      int blk0, int nblk, double *rad, double *tau, hipStream_t s) {
  const dim3 grid(nblk * M.nd), blk(WAVE);
  if (M.formod == 0)
    hipLaunchKernelGGL((k_radiance_f32<NGAS, true>), grid, blk, 0, s, M, B, L, info, blk0, nblk, rad, tau);
  else
    hipLaunchKernelGGL((k_radiance_f32<NGAS, false>), grid, blk, 0, s, M, B, L, info, blk0, nblk, rad, tau);
}

void launch_radiance_f32(const ModelView &M, const BatchView &B, const LosView &L,
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
