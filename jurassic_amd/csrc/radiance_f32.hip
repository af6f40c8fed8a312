// FP32 radiance kernel: combines the FP32 segment emissivities of the lookup
// kernel with continua and the source function, accumulating radiance and
// transmittance along each line of sight (formod_pencil,
// src/jurassic.c:3233-3334) in float with compensated sums.
//
// * Emitters combine as e = 1 - prod_g (1 - e_g) in the cancellation-free
//   form e + e_g - e e_g, and the continuum as e + c - e c with
//   c = 1 - exp(-beta ds).
// * With d_i = tau_{i-1} e_i the transmittance is updated by a two-sum
//   (tau -= d with its rounding error carried) and the radiance
//   sum_i src_i d_i is a Kahan sum.
// * A segment with an opaque emitter (the reference's tau_gas == 0) is
//   skipped together with all later ones, as the reference does once
//   tau_path falls below 1e-9.
// * The continua divide by T and ds through their hardware reciprocals
//   (v_rcp_f32, 1 ulp), formed once per segment: a correctly rounded float
//   division costs about ten instructions, and radiances move by a few
//   float ulp.
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) LOS_AT(los, L, f, ip)

struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__global__ __launch_bounds__(WAVE) void k_radiance_f32(ModelView M, BatchView B, LosView L,
                                                       EpsView E, const RayInfo *info, int blk0,
                                                       int nblk, double *rad_out, double *tau_out) {
  const int nd = M.nd, ng = M.ng;
  int bl, id;
  if (!xcd_block_map(blockIdx.x, nd, nblk, bl, id)) return;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  const atm_t *atm = &B.atm[B.ray_prof[ir]];
  const ChanInfo C = M.chan[id];
  const double nu = C.nu;
  const float nuf = (float)nu;
  // the LOS fields rounded to float by the ray tracer (this kernel rounds
  // every field it reads, so the copy reads half the bytes for the same bits)
  const float *los = L.losf + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const int npb = E.npb[bl];
  const float *eps_g = (const float *)E.eps + E.off[bl] + (size_t)id * ng * npb * WAVE + lane;
  const double *sr_col = M.sr + (size_t)id * TBLNS;

  float kcl = 0.f;
  if (L.f_cl >= 0 && M.ncl > 0)
    kcl = (float)lin(M.clnu[C.icl], atm->clk[C.icl], M.clnu[C.icl + 1], atm->clk[C.icl + 1], nu);
  float sfeps = 1.f;
  if (M.nsf > 0)
    sfeps = (float)lin(M.sfnu[C.isf], atm->sfeps[C.isf], M.sfnu[C.isf + 1], atm->sfeps[C.isf + 1], nu);
  const bool has_k = C.window < M.nw;

  // channel constants of the continua (src/jurassic.c:1724-1741, 2762-2791,
  // 2841-2917), evaluated once in FP64
  bool co2_on = false, h2o_on = false, n2_on = false, o2_on = false;
  float co2_a = 0, co2_b = 0, co2_c = 0;  // ctw = a d260 d296 - b d230 d296 + c d230 d260
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
  const float kctm = (float)(1.0 / (JUR_NA * 1000.0 * JUR_P0));

  // emitters with a table (the others have eps = 0)
  unsigned gmask = 0;
  for (int g = 0; g < ng; g++)
    if (M.tab.np[g * nd + id] >= 30) gmask |= 1u << g;

  KSum rad;
  float tau_hi = 1.f, tau_lo = 0.f;
  float drefl = 0.f, tau_refl = 1.f;
  float epp[NG];  // path emissivity of each emitter before the segment
  for (int g = 0; g < NG; g++) epp[g] = 0.f;
  // The next point's emissivities and LOS fields are loaded one iteration
  // ahead, so that a point's arithmetic does not wait for its own loads
  // (absent fields read F_DS in their place and are not used).
  const int f_k = has_k ? L.f_k + C.window : F_DS;
  const int f_cl = L.f_cl >= 0 ? L.f_cl : F_DS;
  const int f_uc = co2_on ? L.f_u + M.ig_co2 : F_DS;
  const int f_uh = h2o_on ? L.f_u + M.ig_h2o : F_DS;
  const int f_qh = h2o_on ? F_Q0 + M.ig_h2o : F_DS;
  float en[NG];
  float pn = 0, tn = 0, dsn = 0, kn = 0, cln = 0, ucn = 0, uhn = 0, qhn = 0;
  auto load = [&](int ip) {
#pragma unroll
    for (int g = 0; g < NG; g++)
      en[g] = (g < ng && (gmask >> g & 1u)) ? eps_g[((size_t)g * npb + ip) * WAVE] : 0.f;
    pn = LOSR(F_P, ip);
    tn = LOSR(F_T, ip);
    dsn = LOSR(F_DS, ip);
    kn = LOSR(f_k, ip);
    cln = LOSR(f_cl, ip);
    ucn = LOSR(f_uc, ip);
    uhn = LOSR(f_uh, ip);
    qhn = LOSR(f_qh, ip);
  };
  // the source-function bracket of the next point (its T row of sr_col),
  // loaded at the end of the iteration before
  double s0n = 0, s1n = 0, stn = 0;
  auto load_src = [&](float t) {
    int it = (int)((t - st0) * inv_dst);
    it = it < 0 ? 0 : (it > TBLNS - 2 ? TBLNS - 2 : it);
    s0n = sr_col[it];
    s1n = sr_col[it + 1];
    stn = M.st[it];
  };
  if (R.np > 0) {
    load(0);
    load_src(tn);
  }
  for (int ip = 0; ip < R.np; ip++) {
    float ec[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) ec[g] = en[g];
    const float pc = pn, tc = tn, dsc = dsn, kc = kn, clc = cln, ucc = ucn, uhc = uhn, qhc = qhn;
    const double s0c = s0n, s1c = s1n, stc = stn;
    load(ip + 1 < R.np ? ip + 1 : ip);

    bool opaque = false;
    float etot = 0.f;
#pragma unroll
    for (int g = 0; g < NG; g++) {
      if (g >= ng || !(gmask >> g & 1u)) continue;
      const float ep = ec[g];
      opaque |= (ep >= 1.f);
      // segment emissivity 1 - (1 - ep)/(1 - epp) = (ep - epp)/(1 - epp)
      const float e = fdiv(ep - epp[g], 1.f - epp[g]);
      epp[g] = ep;
      etot = etot + e - etot * e;
    }
    if (opaque) break;

    const float pf = (float)pc;
    const float tf = (float)tc;
    const float dsf = (float)dsc;
    const float rt = frcp(tf), rds = frcp(dsf);
    float beta = has_k ? (float)kc : 0.f;
    if (L.f_cl >= 0) beta += (float)clc * kcl;
    if (co2_on) {
      const float uc = (float)ucc;
      const float d230 = tf - 230.f, d260 = tf - 260.f, d296 = tf - 296.f;
      const float ctw = co2_a * d260 * d296 - co2_b * d230 * d296 + co2_c * d230 * d260;
      beta += uc * pf * kctm * ctw * rds;
    }
    if (h2o_on) {
      const float uh = (float)uhc;
      const float qh = (float)qhc;
      const float slf = h2o_slf * exp2f(h2o_l2r * (296.f - tf));
      const float a1 = nuf * uh * tanh_pos_f(0.7193876f * rt * nuf);
      const float a3 = pf * (float)(1.0 / JUR_P0) * (qh * slf + (1.f - qh) * h2o_frn) * 1e-20f;
      beta += a1 * (296.f * rt) * a3 * rds;
    }
    if (n2_on) {
      const float r = pf * (float)(273.0 / JUR_P0) * rt;
      beta += 0.1f * r * r * __expf(n2_beta * (1.f / 296.f - rt)) * 0.79f * n2_b *
              (0.79f + 0.21f * (1.294f - 0.4545f * tf / 296.f));
    }
    if (o2_on) {
      const float r = pf * (float)(273.0 / JUR_P0) * rt;
      beta += 0.1f * r * r * __expf(o2_beta * (1.f / 296.f - rt)) * 0.21f * o2_b;
    }

    // source function: linear in the uniform 1200-point temperature grid
    const float s0 = (float)s0c, s1 = (float)s1c;
    const float src = s0 + (s1 - s0) * ((tf - (float)stc) * inv_dst);

    const float c = -expm1f(-beta * dsf);
    const float e = etot + c - etot * c;
    const float d = tau_hi * e;
    rad.add(src * d);
    {  // tau -= d with its rounding error carried in tau_lo
      const float s = tau_hi - d;
      const float bb = s - tau_hi;
      tau_lo += (tau_hi - (s - bb)) + (-d - bb);
      tau_hi = s;
    }
    drefl = drefl - drefl * e + src * e;
    tau_refl = tau_refl - tau_refl * e;
    load_src(tn);
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

void launch_radiance_f32(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                         const RayInfo *info, int blk0, int nblk, double *rad, double *tau,
                         hipStream_t s) {
  if (M.nd == 0) return;
  hipLaunchKernelGGL(k_radiance_f32, dim3(padded_blocks(nblk) * M.nd), dim3(WAVE), 0, s, M, B, L, E, info, blk0,
                     nblk, rad, tau);
}

}  // namespace jur
