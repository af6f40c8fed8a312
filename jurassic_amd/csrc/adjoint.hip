// Jacobian of the forward model by a reverse sweep: a hand-written adjoint
// of formod_pencil (src/jurassic.c:3196-3338) with the line of sight held
// fixed.  It replaces the n+1 finite-difference formod calls of kernel()
// (src/jurassic.c:4004-4099).
//
// Per ray and channel, with LOS points i = 0..N-1 (N: points before the first
// segment whose gas transmittance is 0, which the reference skips together
// with every later one):
//   P_g,i   path transmittance of emitter g after point i (P_g,-1 = 1);
//           e_g,i = 1 - P_g,i is the path emissivity that the lookup kernels
//           store, and  e_g,i = E(e_g,i-1, p_i, T_i, u_g,i)  (EGA) or
//           E(U_g,i, sum(u p)/U, sum(u T)/U)  (CGA, U = sum of u up to i)
//   s_i     segment transmittance  prod_g (P_g,i / P_g,i-1) exp(-beta_i ds_i)
//   tau_i   observer-to-point transmittance  prod_{k<i} s_k
//   D_i     down-welling sum of the reflection term,
//           D_{i+1} = D_i s_i + src_i (1 - s_i)
//   rad     = sum_i src_i (1 - s_i) tau_i + surface terms in (tau_N, D_N)
// The reverse sweep runs i = N-1..0 with tb = d rad / d tau_{i+1} and
// db = d rad / d D_{i+1} (seeded by the surface terms):
//   G_i = d rad / d s_i = tau_i (tb - src_i) + db (D_i - src_i)
//   d rad / d src_i = (1 - s_i)(tau_i + db)
//   tb <- src_i (1 - s_i) + tb s_i,   db <- db s_i
//   H_i = G_i s_i,   d rad / d beta_i = -H_i ds_i
//   d rad / d e_g,i (through s_i and s_i+1) = (H_i+1 - H_i) / P_g,i
//   L_g,i = that + L_g,i+1 dE_g,i+1 / d e_g,i         (EGA chain)
// The table partials dE/d(e_prev, p, T, u) are those of the brackets that
// the lookup used: the FP32 EGA lookup over monotone tables stores them per
// step; otherwise they are recomputed from the stored path emissivities
// (piecewise-linear in u, linear in T, linear (EGA) or logarithmic (CGA) in
// p; zero where eps00 is clamped).  Continua and the source function are
// differentiated analytically.
//
// Point adjoints go to atmospheric levels through intpol_atm's weights
// (src/jurassic.c:3685-3706: LOGY in p, LIN in T, q, k) and the column
// densities u = 10 q p / (kB T) ds (src/jurassic.c:4417-4419).
//
// k_adj_acc, one wavefront per (block of 64 rays, channel), lane = ray: a
// forward sweep (tau_i, D_i) and the reverse sweep of the accumulation (H_i,
// kept in place of tau_i).  k_adj_chain, one wavefront per (block, channel,
// emitter) like the lookup kernels: the emitter's chain in reverse with its
// table partials (stored by the FP32 EGA lookup, recomputed otherwise).  A lane sums the contributions of the current
// level pair in registers and adds them to its row of the block tile
// kt[(block, channel)][column][lane] when the pair changes, so tile accesses
// of a wavefront are contiguous; T and p columns, which every emitter chain
// of a row updates, take atomic adds.  k_jac_store
// then writes the tile into row-major K, scaled by d BT / d rad when
// brightness temperatures are written.
//
// Not differentiated (the path is fixed): refraction's dependence on p and T
// and the surface altitude/pressure (sfz, sfp), which only move the path.
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) LOS_AT(los, L, f, ip)

// The slopes below are those of the arithmetic each forward path runs, not
// of the exact piecewise-linear table: the FP64 path interpolates with LIN
// on float table entries, whose slope C forms in float (lin_f), and the FP32
// path with reciprocal-multiply float slopes.  Within one bracket a chain
// step then has d e_new / d e_prev = s_eps s_u = 1 + O(1e-7) rather than 1,
// which compounds over the ~1000 steps of a chain; the adjoint follows the
// forward model it differentiates.

// FP64 path: intpol_tbl_eps at x from its u-bracket record (eps_from_rec):
// value e and slope de/du.
__device__ __forceinline__ double eps_vs(const float4 &r, double x, double &e) {
  if (x < (double)r.x) {
    const double s = (double)(r.y / r.x);
    e = s * x;
    return s;
  }
  if (x > (double)r.z) {
    const double a = log((double)(1.f - r.w)) / (double)r.z;
    const double ex = exp(a * x);
    e = 1 - ex;
    return ex > 0 ? -a * ex : 0.0;
  }
  const double s = (double)((r.w - r.y) / (r.z - r.x));
  e = (double)r.y + s * (x - (double)r.x);
  return s;
}

// FP64 path: intpol_tbl_u at x from its eps-bracket record (u_from_rec):
// value u and slope du/deps.
__device__ __forceinline__ double u_vs(const float4 &r, double x, double &u) {
  if (x < (double)r.y) {
    const double s = (double)(r.x / r.y);
    u = s * x;
    return s;
  }
  if (x > (double)r.w) {
    const double a = log((double)(1.f - r.w)) / (double)r.z;
    u = log1p(-x) / a;
    return -1.0 / ((1.0 - x) * a);
  }
  const double s = (double)((r.z - r.x) / (r.w - r.y));
  u = (double)r.x + s * (x - (double)r.y);
  return s;
}

// FP32 path: col_eps_bis and col_u_bis with their slopes.
__device__ __forceinline__ float col_eps_bis_vs(const float4 *col, int n, double x, float &s) {
  const float4 r0 = col[0];
  if (x < (double)r0.x) {
    s = fdiv(r0.y, r0.x);
    return s * (float)x;
  }
  const float4 rl = col[n - 1];
  if (x > (double)rl.x) {
    const float a = fdiv(__logf(1.f - rl.y), rl.x);
    const float ex = __expf(a * (float)x);
    s = ex > 0.f ? -a * ex : 0.f;
    return 1.f - ex;
  }
  float4 r;
  col_bisect<0>(col, n, x, r);
  s = fdiv(r.w - r.y, r.z - r.x);
  return r.y + s * (float)(x - (double)r.x);
}

__device__ __forceinline__ double col_u_bis_vs(const float4 *col, int n, float x, float &s) {
  const float4 r0 = col[0];
  if (x < r0.y) {
    s = fdiv(r0.x, r0.y);
    return (double)(s * x);
  }
  const float4 rl = col[n - 1];
  if (x > rl.y) {
    const float c = fdiv(rl.x, __logf(1.f - rl.y));
    s = -c / (1.f - x);
    return (double)(log1pf(-x) * c);
  }
  float4 r;
  col_bisect<1>(col, n, (double)x, r);
  s = fdiv(r.z - r.x, r.w - r.y);
  return (double)r.x + (double)(s * (x - r.y));
}

// EGA corner: eps(u(e_prev) + u_seg) with d/d e_prev and d/d u_seg.
__device__ __forceinline__ void ega_corner(const float4 *col, int n, int mono, double ep, double useg,
                                           int &g, double &e, double &dep, double &du) {
  float4 r, r2;
  int i;
  if (mono) i = col_walk<1>(col, n, ep, g, r);
  else i = col_bisect<1>(col, n, ep, r);
  double u;
  const double dudx = u_vs(r, ep, u);
  const double arg = u + useg;
  if (mono) col_walk_r<0>(col, n, arg, i, r, r2);
  else col_bisect<0>(col, n, arg, r2);
  du = eps_vs(r2, arg, e);
  dep = du * dudx;
  g = i;
}

// CGA corner: eps(U) and d/dU.
__device__ __forceinline__ void cga_corner(const float4 *col, int n, int mono, double cu, int &g,
                                           double &e, double &du) {
  float4 r;
  g = mono ? col_walk<0>(col, n, cu, g, r) : col_bisect<0>(col, n, cu, r);
  du = eps_vs(r, cu, e);
}

// Partials of one chain step e_new = E(e_prev, pq, tq, x) (x = u_seg for
// EGA, the column sum U for CGA).  A cell without data (nt or nu < 2) gives
// eps = 0, i.e. e_new = e_prev.
struct LkPart {
  double dep, dp, dt, du;
};

template <bool CGA>
__device__ __forceinline__ LkPart lookup_partials(const TabView &T, int tix, int npt, int mono, int pdir,
                                                  double ep, double pq, double tq, double x, int &ipr,
                                                  int *gi) {
  LkPart d{0, 0, 0, 0};
  const double *pg = T.p + (size_t)tix * T.TP;
  ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
  const double pa = pg[ipr], pb = pg[ipr + 1];
  const int *ntp = T.nt + (size_t)tix * T.TP;
  const int nt0 = ntp[ipr], nt1 = ntp[ipr + 1];
  if (nt0 < 2 || nt1 < 2) {
    d.dep = 1;
    return d;
  }
  const double *tg0 = T.t + ((size_t)tix * T.TP + ipr) * T.TT;
  int j0 = (int)((tq - tg0[0]) / (tg0[1] - tg0[0]));
  j0 = j0 < 0 ? 0 : (j0 > nt0 - 2 ? nt0 - 2 : j0);
  int j1 = (int)((tq - tg0[T.TT]) / (tg0[T.TT + 1] - tg0[T.TT]));
  j1 = j1 < 0 ? 0 : (j1 > nt1 - 2 ? nt1 - 2 : j1);
  const size_t c0 = (size_t)tix * T.TP + ipr;
  int nc[4];
  nc[0] = T.nu[c0 * T.TT + j0];
  nc[1] = T.nu[c0 * T.TT + j0 + 1];
  nc[2] = T.nu[(c0 + 1) * T.TT + j1];
  nc[3] = T.nu[(c0 + 1) * T.TT + j1 + 1];
  if (nc[0] < 2 || nc[1] < 2 || nc[2] < 2 || nc[3] < 2) {
    d.dep = 1;
    return d;
  }
  // eps00 = sum_c cw_c e_c is linear in the corner values: every partial is a
  // weighted sum over the corners, accumulated corner by corner
  const double ta0 = tg0[j0], tb0 = tg0[j0 + 1];
  const double ta1 = tg0[T.TT + j1], tb1 = tg0[T.TT + j1 + 1];
  const double s0 = 1 / (tb0 - ta0), s1 = 1 / (tb1 - ta1);
  const double w0 = (tq - ta0) * s0, w1 = (tq - ta1) * s1;
  double wp, dwp;
  if (CGA && pq / pa > 0 && pb / pa > 0) {
    const double l = 1 / log(pb / pa);
    wp = log(pq / pa) * l;
    dwp = l / pq;
  } else {
    dwp = 1 / (pb - pa);
    wp = (pq - pa) * dwp;
  }
  const double q0 = 1 - wp, q1 = wp;
  const double cw[4] = {q0 * (1 - w0), q0 * w0, q1 * (1 - w1), q1 * w1};  // value
  const double tw[4] = {-q0 * s0, q0 * s0, -q1 * s1, q1 * s1};            // d/dT
  const double pw[4] = {-(1 - w0) * dwp, -w0 * dwp, (1 - w1) * dwp, w1 * dwp};  // d/dp
  const float4 *base0 = T.rec + (c0 * T.TT + j0) * T.TU;
  const float4 *base1 = T.rec + ((c0 + 1) * T.TT + j1) * T.TU;
  double en = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const float4 *col = (c < 2 ? base0 : base1) + (c & 1) * T.TU;
    double e, de, du;
    if (CGA) {
      cga_corner(col, nc[c], mono, x, gi[c], e, du);
      de = 0;
    } else {
      ega_corner(col, nc[c], mono, ep, x, gi[c], e, de, du);
    }
    en += cw[c] * e;
    d.dep += cw[c] * de;
    d.du += cw[c] * du;
    d.dt += tw[c] * e;
    d.dp += pw[c] * e;
  }
  if (en < 0 || en > 1) return LkPart{0, 0, 0, 0};  // MAX(MIN(eps00, 1), 0) holds it constant
  return d;
}


// FP32 path: partials of one chain step of k_lookup_f32, with its float
// cell weights, its knot-relative update inside a bracket and its
// reciprocal-multiply slopes.
template <bool CGA>
__device__ __forceinline__ LkPart lookup_partials_f32(const TabView &T, int tix, int npt, int mono, int pdir,
                                                      float ep, double pq, float tq, double x, int &ipr,
                                                      int *gi) {
  LkPart d{0, 0, 0, 0};
  const double *pg = T.p + (size_t)tix * T.TP;
  ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
  const double pa = pg[ipr], pb = pg[ipr + 1];
  const int *ntp = T.nt + (size_t)tix * T.TP;
  const int nt0 = ntp[ipr], nt1 = ntp[ipr + 1];
  if (nt0 < 2 || nt1 < 2) {
    d.dep = 1;
    return d;
  }
  const double *tg0 = T.t + ((size_t)tix * T.TP + ipr) * T.TT;
  int j0 = (int)((tq - (float)tg0[0]) * frcp((float)(tg0[1] - tg0[0])));
  j0 = j0 < 0 ? 0 : (j0 > nt0 - 2 ? nt0 - 2 : j0);
  int j1 = (int)((tq - (float)tg0[T.TT]) * frcp((float)(tg0[T.TT + 1] - tg0[T.TT])));
  j1 = j1 < 0 ? 0 : (j1 > nt1 - 2 ? nt1 - 2 : j1);
  const size_t c0 = (size_t)tix * T.TP + ipr;
  int nc[4];
  nc[0] = T.nu[c0 * T.TT + j0];
  nc[1] = T.nu[c0 * T.TT + j0 + 1];
  nc[2] = T.nu[(c0 + 1) * T.TT + j1];
  nc[3] = T.nu[(c0 + 1) * T.TT + j1 + 1];
  if (nc[0] < 2 || nc[1] < 2 || nc[2] < 2 || nc[3] < 2) {
    d.dep = 1;
    return d;
  }
  const float ta0 = (float)tg0[j0], dt0 = frcp((float)(tg0[j0 + 1] - tg0[j0]));
  const float ta1 = (float)tg0[T.TT + j1], dt1 = frcp((float)(tg0[T.TT + j1 + 1] - tg0[T.TT + j1]));
  const float w0 = (tq - ta0) * dt0, w1 = (tq - ta1) * dt1;
  float wp, dwp;
  if (CGA && pq / pa > 0 && pb / pa > 0) {
    const float lpr = (float)(1.0 / log(pb / pa));
    wp = __logf((float)(pq / pa)) * lpr;
    dwp = lpr / (float)pq;
  } else {
    dwp = frcp((float)(pb - pa));
    wp = (float)(pq - pa) * dwp;
  }
  const float q0 = 1.f - wp, q1 = wp;
  const float cw[4] = {q0 * (1.f - w0), q0 * w0, q1 * (1.f - w1), q1 * w1};
  const float tw[4] = {-q0 * dt0, q0 * dt0, -q1 * dt1, q1 * dt1};
  const float pw[4] = {-(1.f - w0) * dwp, -w0 * dwp, (1.f - w1) * dwp, w1 * dwp};
  const float4 *base0 = T.rec + (c0 * T.TT + j0) * T.TU;
  const float4 *base1 = T.rec + ((c0 + 1) * T.TT + j1) * T.TU;
  float ec[4];
  float dep = 0.f, du = 0.f, dt = 0.f, dp = 0.f;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const float4 *col = (c < 2 ? base0 : base1) + (c & 1) * T.TU;
    const int n = nc[c];
    float e, de, su, se;
    if (CGA) {
      if (mono) {
        float4 r;
        gi[c] = col_walk<0>(col, n, x, gi[c], r);
        e = eps_rec_f_vs(r, x, se);
      } else {
        e = col_eps_bis_vs(col, n, x, se);
      }
      de = 0.f;
    } else if (!mono) {
      const double u = col_u_bis_vs(col, n, ep, su);
      e = col_eps_bis_vs(col, n, u + x, se);
      de = se * su;
    } else {
      e = ega_corner_f32_d(col, n, ep, x, gi[c], de, se);
    }
    ec[c] = e;
    dep += cw[c] * de;
    du += cw[c] * se;
    dt += tw[c] * e;
    dp += pw[c] * e;
  }
  const float a = ec[0] + (ec[1] - ec[0]) * w0;
  const float b = ec[2] + (ec[3] - ec[2]) * w1;
  const float en = a + (b - a) * wp;
  if (en < 0.f || en > 1.f) return d;  // clamped: constant
  d.dep = dep;
  d.du = du;
  d.dt = dt;
  d.dp = dp;
  return d;
}

// Continuum extinction beta of one segment (formod_continua,
// src/jurassic.c:3075-3117) and its partials.
struct BetaGrad {
  double dp, dt, du_co2, du_h2o, dq_h2o;
};

__device__ __forceinline__ double beta_grad(const ModelView &M, const CtmChan &K, double nu, double p,
                                            double t, double ds, double q_h2o, double u_co2,
                                            double u_h2o, double kwin, double clx, BetaGrad &g) {
  g.dp = g.dt = g.du_co2 = g.du_h2o = g.dq_h2o = 0;
  double beta = kwin + clx;
  if (M.ig_co2 >= 0 && K.co2) {
    const double dt230 = t - 230, dt260 = t - 260, dt296 = t - 296;
    const double ctw = dt260 * 5.050505e-4 * dt296 * K.co2_230 - dt230 * 9.259259e-4 * dt296 * K.co2_260 +
                       dt230 * 4.208754e-4 * dt260 * K.co2_296;
    const double dctw = 5.050505e-4 * K.co2_230 * (dt260 + dt296) -
                        9.259259e-4 * K.co2_260 * (dt230 + dt296) +
                        4.208754e-4 * K.co2_296 * (dt230 + dt260);
    const double f = 1 / JUR_NA / 1000 / JUR_P0 / ds;
    beta += u_co2 * p * ctw * f;
    g.du_co2 = p * ctw * f;
    g.dp += u_co2 * ctw * f;
    g.dt += u_co2 * p * dctw * f;
  }
  if (M.ig_h2o >= 0 && K.h2o) {
    const double S = K.h2o_slf * ctm_h2o_pow(K, (296 - t) / (296 - 260));
    const double dS = K.h2o_ratio > 0 ? S * log(K.h2o_ratio) * (-1.0 / 36) : 0.0;
    const double bq = .7193876 * nu;
    const double th = tanh_pos(bq / t);
    const double a1u = nu * th;                            // a1 per unit column
    const double da1u = nu * (1 - th * th) * (-bq / (t * t));
    const double a2 = 296 / t, da2 = -296 / (t * t);
    const double mix = q_h2o * S + (1 - q_h2o) * K.h2o_frn;
    const double a3 = p / JUR_P0 * mix * 1e-20;
    const double da3 = p / JUR_P0 * q_h2o * dS * 1e-20;
    const double rds = 1 / ds;
    beta += a1u * u_h2o * a2 * a3 * rds;
    g.du_h2o = a1u * a2 * a3 * rds;
    g.dq_h2o = a1u * u_h2o * a2 * (p / JUR_P0 * (S - K.h2o_frn) * 1e-20) * rds;
    g.dp += a1u * u_h2o * a2 * (mix / JUR_P0 * 1e-20) * rds;
    g.dt += u_h2o * (da1u * a2 * a3 + a1u * da2 * a3 + a1u * a2 * da3) * rds;
  }
  if (M.ctm_n2 && K.n2) {
    const double qn = 0.79, r = p / JUR_P0 * 273.0 / t;
    const double ex = exp(K.n2_beta * (1 / 296.0 - 1 / t));
    const double F = qn + (1 - qn) * (1.294 - 0.4545 * t / 296.0);
    const double base = 0.1 * r * r * ex * qn * K.n2_b;
    const double c = base * F;
    beta += c;
    g.dp += 2 * c / p;
    g.dt += c * (-2 / t + K.n2_beta / (t * t)) + base * (1 - qn) * (-0.4545 / 296.0);
  }
  if (M.ctm_o2 && K.o2) {
    const double r = p / JUR_P0 * 273.0 / t;
    const double c = 0.1 * r * r * exp(K.o2_beta * (1 / 296.0 - 1 / t)) * 0.21 * K.o2_b;
    beta += c;
    g.dp += 2 * c / p;
    g.dt += c * (-2 / t + K.o2_beta / (t * t));
  }
  return beta;
}

// Per-lane constants of beta_grad_f.
struct CtmChanF {
  float h2o_l2r;    // log2(cw260 / cw296) / 36: S = slf 2^(l2r (296 - T))
  double h2o_dlr;   // d ln S / dT = -ln(cw260 / cw296) / 36
  double bq;        // 0.7193876 nu
};

__device__ __forceinline__ CtmChanF ctm_chan_f(const CtmChan &K, double nu) {
  CtmChanF f;
  f.h2o_l2r = K.h2o_ratio > 0 ? (float)(log2(K.h2o_ratio) / 36.0) : 0.f;
  f.h2o_dlr = K.h2o_ratio > 0 ? log(K.h2o_ratio) * (-1.0 / 36) : 0.0;
  f.bq = .7193876 * nu;
  return f;
}

// beta_grad for the Jacobian of the FP32 forward model, whose continua are
// evaluated in float (radiance_f32.hip): the transcendentals (the H2O
// self-continuum power, tanh, the N2/O2 exponentials) in float, and the
// divisions by T, p and ds as products with the reciprocals rt, rp, rds;
// the products and sums stay FP64 (the column densities and the continuum
// coefficients span ranges beyond float).
__device__ __forceinline__ double beta_grad_f(const ModelView &M, const CtmChan &K, const CtmChanF &KF,
                                              double nu, double p, double t, double rt, double rp, double rds,
                                              double q_h2o, double u_co2, double u_h2o, double kwin, double clx,
                                              BetaGrad &g) {
  g.dp = g.dt = g.du_co2 = g.du_h2o = g.dq_h2o = 0;
  double beta = kwin + clx;
  if (M.ig_co2 >= 0 && K.co2) {
    const double dt230 = t - 230, dt260 = t - 260, dt296 = t - 296;
    const double ctw = dt260 * 5.050505e-4 * dt296 * K.co2_230 - dt230 * 9.259259e-4 * dt296 * K.co2_260 +
                       dt230 * 4.208754e-4 * dt260 * K.co2_296;
    const double dctw = 5.050505e-4 * K.co2_230 * (dt260 + dt296) -
                        9.259259e-4 * K.co2_260 * (dt230 + dt296) +
                        4.208754e-4 * K.co2_296 * (dt230 + dt260);
    const double f = (1 / JUR_NA / 1000 / JUR_P0) * rds;
    beta += u_co2 * p * ctw * f;
    g.du_co2 = p * ctw * f;
    g.dp += u_co2 * ctw * f;
    g.dt += u_co2 * p * dctw * f;
  }
  if (M.ig_h2o >= 0 && K.h2o) {
    const double S = K.h2o_slf * (double)exp2f(KF.h2o_l2r * (float)(296 - t));
    const double dS = S * KF.h2o_dlr;
    const double bq = KF.bq;
    const double th = (double)tanh_pos_f((float)(bq * rt));
    const double a1u = nu * th;
    const double da1u = nu * (1 - th * th) * (-bq * rt * rt);
    const double a2 = 296 * rt, da2 = -296 * rt * rt;
    const double mix = q_h2o * S + (1 - q_h2o) * K.h2o_frn;
    const double pp0 = p * (1 / JUR_P0);
    const double a3 = pp0 * mix * 1e-20;
    const double da3 = pp0 * q_h2o * dS * 1e-20;
    beta += a1u * u_h2o * a2 * a3 * rds;
    g.du_h2o = a1u * a2 * a3 * rds;
    g.dq_h2o = a1u * u_h2o * a2 * (pp0 * (S - K.h2o_frn) * 1e-20) * rds;
    g.dp += a1u * u_h2o * a2 * (mix * (1 / JUR_P0) * 1e-20) * rds;
    g.dt += u_h2o * (da1u * a2 * a3 + a1u * da2 * a3 + a1u * a2 * da3) * rds;
  }
  if (M.ctm_n2 && K.n2) {
    const double qn = 0.79, r = p * (273.0 / JUR_P0) * rt;
    const double ex = (double)__expf((float)(K.n2_beta * (1 / 296.0 - rt)));
    const double F = qn + (1 - qn) * (1.294 - t * (0.4545 / 296.0));
    const double base = 0.1 * r * r * ex * qn * K.n2_b;
    const double c = base * F;
    beta += c;
    g.dp += 2 * c * rp;
    g.dt += c * (-2 * rt + K.n2_beta * rt * rt) + base * (1 - qn) * (-0.4545 / 296.0);
  }
  if (M.ctm_o2 && K.o2) {
    const double r = p * (273.0 / JUR_P0) * rt;
    const double c = 0.1 * r * r * (double)__expf((float)(K.o2_beta * (1 / 296.0 - rt))) * 0.21 * K.o2_b;
    beta += c;
    g.dp += 2 * c * rp;
    g.dt += c * (-2 * rt + K.o2_beta * rt * rt);
  }
  return beta;
}

// formod_srcfunc value and its slope in T.
__device__ __forceinline__ double srcfunc_d(const double *st, const double *sr_col, double t, double &d) {
  const int it = locate_reg(st, TBLNS, t);
  d = (sr_col[it + 1] - sr_col[it]) / (st[it + 1] - st[it]);
  return sr_col[it] + d * (t - st[it]);
}

// srcfunc_d on the uniform grid with its inverse spacing inv_dst in hand
// (products instead of the two divisions; the FP32 Jacobian's sweeps).
__device__ __forceinline__ double srcfunc_du(const double *st, double inv_dst, const double *sr_col, double t,
                                             double &d) {
  int it = (int)((t - st[0]) * inv_dst);
  it = it < 0 ? 0 : (it > TBLNS - 2 ? TBLNS - 2 : it);
  d = (sr_col[it + 1] - sr_col[it]) * inv_dst;
  return sr_col[it] + d * (t - st[it]);
}

// Stored chain state: the FP32 lookups store e (float), the FP64 lookups in
// path mode store P (double).
__device__ __forceinline__ double trans_of(float s) { return 1.0 - (double)s; }
__device__ __forceinline__ double trans_of(double s) { return s; }
__device__ __forceinline__ double emis_of(float s) { return (double)s; }
__device__ __forceinline__ double emis_of(double s) { return 1.0 - s; }

// Level bracket of altitude z in the atmosphere (locate_irr, with the
// previous bracket kept while it still holds) and the LIN weight.
__device__ __forceinline__ int level_of(const atm_t *atm, double z, int lev, double &w) {
  const int n = atm->np;
  if (!(lev >= 0 && lev <= n - 2 && atm->z[lev] <= z && z < atm->z[lev + 1]))
    lev = locate_irr(atm->z, n, z);
  const double z0 = atm->z[lev], z1 = atm->z[lev + 1];
  w = (z - z0) / (z1 - z0);
  return lev;
}

// Sums of point adjoints of the current level pair (wa, wa+1) for NQ level
// quantities (JUR_XMAP_LEV quantity index qi[q], -1 = none); flushed into the
// lane's tile row when the pair changes.
// Quantities in the mask ATOM are added atomically: columns that other
// wavefronts of the same row also update, and columns a wavefront owns,
// where an atomic add without return saves the flush the read of the tile
// (one wavefront's adds to one address are applied in its program order, so
// the sums keep their order).  With CCOL the pair's K columns are loaded
// when the pair is entered, so that its flush does not wait for them (at
// the cost of registers, which the FP64 kernels cannot spare).
// The pair's altitudes and pressures and the reciprocals the weights need are
// kept while the pair lasts (a pair spans ~10 LOS points at RAYDZ = 0.1 km).
// A path leaves a pair into a neighbouring one: that pair costs one round
// trip for the new level's altitude and pressure, where a search of the
// level grid (locate_irr) would take one per bisection step.
// With PATH, add() also returns the point adjoints projected on the altitude
// of the point, sum_q v[q] dQ_q/dz (intpol_atm's slopes within the pair),
// which the adjoint of the refracted path needs.
__device__ __forceinline__ double level_val(const atm_t *atm, int qi, int l) {
  return qi == 0 ? atm->p[l] : (qi == 1 ? atm->t[l] : (qi < 2 + NG ? atm->q[qi - 2][l] : atm->k[qi - 2 - NG][l]));
}

template <int NQ, unsigned ATOM = 0, bool CCOL = true, bool PATH = false>
struct LevelWindow {
  int wa = -1;
  double z0 = 0, z1 = 0, rz = 0, P0 = 0, P1 = 0, rp0 = 0, rp1 = 0;
  bool plog = false;
  double acc[NQ][2];
  double sl[NQ];   // PATH: slope of each quantity in the pair (p: of log p, or of p)
  int col[NQ][2];  // K columns of the pair's levels (loaded when the pair is entered)
  __device__ __forceinline__ void flush(const int *xm, const int *qi, double *kt) {
    if (wa < 0) return;
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (qi[q] < 0) continue;
        const int c = CCOL ? col[q][k] : xm[JUR_XMAP_LEV(qi[q], wa + k)];
        if (c < 0) continue;
        if (ATOM >> q & 1u) unsafeAtomicAdd(&kt[(size_t)c * WAVE], acc[q][k]);
        else kt[(size_t)c * WAVE] += acc[q][k];
      }
  }
  // Point adjoints v[q] at altitude z; quantity q == qp is pressure (LOGY
  // weights p/p_level, src/jurassic.c:3685-3706), the others use the LIN
  // weights.
  __device__ __forceinline__ double add(const atm_t *atm, double z, int &lev, const double *v, int qp, double p,
                                        const int *xm, const int *qi, double *kt) {
    if (!(wa >= 0 && z0 <= z && z < z1)) {
      // the pair locate_irr returns: a neighbour of the current pair (with
      // its clamps at the first and last pair), else a search
      int nl = -1;
      if (wa >= 0 && z0 < z1) {
        if (z < z0) {
          if (wa > 0) {
            const double zl = atm->z[wa - 1], Pl = atm->p[wa - 1];
            if (zl <= z || wa == 1) {
              nl = wa - 1;
              z1 = z0;
              z0 = zl;
              P1 = P0;
              P0 = Pl;
              rp1 = rp0;
              rp0 = 1 / Pl;
            }
          }
        } else {
          const int n = atm->np;
          if (wa + 2 <= n - 1) {
            const double zh = atm->z[wa + 2], Ph = atm->p[wa + 2];
            if (z < zh || wa + 2 == n - 1) {
              nl = wa + 1;
              z0 = z1;
              z1 = zh;
              P0 = P1;
              P1 = Ph;
              rp0 = rp1;
              rp1 = 1 / Ph;
            }
          }
        }
      }
      if (nl < 0) {
        double w;
        nl = level_of(atm, z, lev, w);
        z0 = atm->z[nl];
        z1 = atm->z[nl + 1];
        P0 = atm->p[nl];
        P1 = atm->p[nl + 1];
        rp0 = 1 / P0;
        rp1 = 1 / P1;
      }
      rz = 1 / (z1 - z0);
      plog = P1 / P0 > 0;
      if (PATH) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          if (qi[q] < 0) sl[q] = 0;
          else if (q == qp) sl[q] = plog ? log(P1 / P0) * rz : (P1 - P0) * rz;
          else sl[q] = (level_val(atm, qi[q], nl + 1) - level_val(atm, qi[q], nl)) * rz;
        }
      }
      if (nl != wa) {
        flush(xm, qi, kt);
        wa = nl;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          acc[q][0] = acc[q][1] = 0;
          if (CCOL) {
            col[q][0] = qi[q] >= 0 ? xm[JUR_XMAP_LEV(qi[q], nl)] : -1;
            col[q][1] = qi[q] >= 0 ? xm[JUR_XMAP_LEV(qi[q], nl + 1)] : -1;
          }
        }
      }
      lev = nl;
    }
    const double w = (z - z0) * rz;
    double dz = 0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      if (q == qp && plog) {
        acc[q][0] += (1 - w) * (p * rp0) * v[q];
        acc[q][1] += w * (p * rp1) * v[q];
      } else {
        acc[q][0] += (1 - w) * v[q];
        acc[q][1] += w * v[q];
      }
      if (PATH) dz += v[q] * ((q == qp && plog) ? p * sl[q] : sl[q]);
    }
    return dz;
  }
};

// Reverse sweep of one emitter chain per wavefront: L_g,i from H_i and the
// recomputed table partials; T and p entries are shared with the other
// emitters of the row (atomic adds), q_g entries belong to this chain.
template <bool CGA, typename SF, bool FAST, bool PATH>
__global__ __launch_bounds__(WAVE) void k_adj_chain(ModelView M, BatchView B, LosView L, EpsView E,
                                                    const RayInfo *info, int blk0, int nblk, JacView J) {
  const int nd = M.nd, ng = M.ng;
  int bl, rem;
  if (!xcd_block_map(blockIdx.x, nd * ng, nblk, bl, rem)) return;
  const int id = rem / ng;
  const int g = rem - id * ng;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const int tix = g * nd + id;
  const int npt = M.tab.np[tix];
  if (npt < 30) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;
  const int ipf = B.ray_prof[ir];
  const int im = J.prof_map[ipf];
  if (im < 0 || im >= J.nmap) return;
  const int *xm = J.xmap + (size_t)im * JUR_XMAP_LEN;
  const atm_t *atm = &B.atm[ipf];
  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const int npb = E.npb[bl];
  const SF *epsb = (const SF *)E.eps + E.off[bl] + (size_t)id * ng * npb * WAVE + lane;
#define EPS_AT(g, ip) epsb[((size_t)(g) * npb + (ip)) * WAVE]
  double *sc = J.scr + J.soff[bl] + (size_t)id * npb * J.nscr * WAVE + lane;
#define SCR(k, ip) sc[((size_t)(ip) * J.nscr + (k)) * WAVE]
  double *kt = J.kt + ((size_t)bl * nd + id) * J.ncol * WAVE + lane;
  const int nef = J.nef[((size_t)bl * nd + id) * WAVE + lane];
  // the path adjoint's point values: this emitter's part of d rad / d z_i
  // and d rad / d ds_i (the other emitters of the row add theirs)
  auto path_add = [&](int ip, double gz, double gd) {
    if (PATH) {
      unsafeAtomicAdd(&SCR(J.kz, ip), gz);
      unsafeAtomicAdd(&SCR(J.kz + 1, ip), gd);
    }
  };
  const int mono = M.tab.mono[tix], pdir = M.tab.pdir[tix];
  const int fu = L.f_u + g;
  double cu = 0, cp = 0, ct = 0;  // Curtis-Godson sums up to the current point
  if (CGA)
    for (int ip = 0; ip < nef; ip++) {
      const double u = LOSR(fu, ip);
      cu += u;
      cp += u * LOSR(F_P, ip);
      ct += u * LOSR(F_T, ip);
    }
  const int qi[3] = {1, 0, 2 + g};
  LevelWindow<3, 7u, FAST, PATH> W;
  int lev = -1, ipr = -1;
  int gi[4];
#pragma unroll
  for (int c = 0; c < 4; c++) gi[c] = M.tab.TU / 2;
  double carry = 0, Hn = 0, sU = 0, sP = 0, sT = 0;
  double trc = nef > 0 ? trans_of(EPS_AT(g, nef - 1)) : 1;
  constexpr bool F32 = sizeof(SF) == sizeof(float);
  if constexpr (FAST) {
    static_assert(F32 && !CGA, "the fast chain is the FP32 EGA one");
    {
      // FP32 EGA on monotone tables: k_lookup_f32 stored each step's
      // partials, so the chain is the recurrence alone.  Each point's
      // inputs (partials, LOS values, the emissivity before it, H) are
      // loaded one point ahead.
      const double *losb = L.los + (size_t)bl * L.cap * L.nfield * WAVE;
      const unsigned nfield = L.nfield;
      const float *epsg = (const float *)E.eps + E.off[bl] + ((size_t)id * ng + g) * npb * WAVE;
      const float4 *pab = E.part + E.off[bl] + ((size_t)id * ng + g) * npb * WAVE;
      const double *scb = J.scr + J.soff[bl] + (size_t)id * npb * J.nscr * WAVE;
      const unsigned nscr = J.nscr;
#define LOSQ(f, ip) ld32(losb, ((unsigned)(ip) * nfield + (unsigned)(f)) * WAVE + lane)
#define EPSQ(ip) ld32(epsg, (unsigned)(ip) * WAVE + lane)
      double p = 0, t = 0, useg = 0, Hc = 0, zc = 0, dsc = 0;
      float epf = 0.f;
      float4 dc = make_float4(1.f, 0.f, 0.f, 0.f);
      if (nef > 0) {
        dc = ld32(pab, (unsigned)(nef - 1) * WAVE + lane);
        p = LOSQ(F_P, nef - 1);
        t = LOSQ(F_T, nef - 1);
        useg = LOSQ(fu, nef - 1);
        epf = nef > 1 ? EPSQ(nef - 2) : 0.f;
        Hc = ld32(scb, (unsigned)(nef - 1) * nscr * WAVE + lane);
        zc = LOSQ(F_Z, nef - 1);
        dsc = LOSQ(F_DS, nef - 1);
      }
      for (int ip = nef - 1; ip >= 0; ip--) {
        const int ipn = ip > 0 ? ip - 1 : 0;
        const float4 dx = ld32(pab, (unsigned)ipn * WAVE + lane);
        const double pn = LOSQ(F_P, ipn), tn = LOSQ(F_T, ipn), un = LOSQ(fu, ipn);
        const float en_f = ipn > 0 ? EPSQ(ipn - 1) : 0.f;
        const double Hx = ld32(scb, (unsigned)ipn * nscr * WAVE + lane);
        const double zx = LOSQ(F_Z, ipn), dsx = LOSQ(F_DS, ipn);
        const LkPart d{dc.x, dc.y, dc.z, dc.w};
        const double H = Hc, z = zc, ds = dsc;
        const double trp = 1.0 - (double)epf;
        // the chain rule's divisions by the transmittance, T and p as float
        // reciprocals (relative error 1e-7, far below the FP32 lookups')
        const double rt = (double)frcp((float)t), rp = (double)frcp((float)p);
        const double lam = (Hn - H) * (double)frcp((float)trc) + carry;
        const double u_bar = lam * d.du, p_bar = lam * d.dp, t_bar = lam * d.dt;
        carry = lam * d.dep;
        const double ucoef = 10 * p * rt * ds * (1 / JUR_KB);
        double v[3];
        v[0] = t_bar - u_bar * useg * rt;
        v[1] = p_bar + u_bar * useg * rp;
        v[2] = u_bar * ucoef;
        const double gz = W.add(atm, z, lev, v, 1, p, xm, qi, kt);
        path_add(ip, gz, u_bar * useg / ds);
        Hn = H;
        trc = ip > 0 ? trp : 1.0;
        p = pn;
        t = tn;
        useg = un;
        epf = en_f;
        Hc = Hx;
        zc = zx;
        dsc = dsx;
        dc = dx;
      }
      W.flush(xm, qi, kt);
#undef LOSQ
#undef EPSQ
      return;
    }
  } else {
  for (int ip = nef - 1; ip >= 0; ip--) {
    const double H = SCR(0, ip);
    double epp = 0, trp = 1;
    float epf = 0.f;  // the stored float path emissivity (FP32 path)
    if (ip > 0) {
      const SF sp = EPS_AT(g, ip - 1);
      epp = emis_of(sp);
      trp = trans_of(sp);
      epf = (float)sp;
    }
    const double z = LOSR(F_Z, ip), p = LOSR(F_P, ip), t = LOSR(F_T, ip), ds = LOSR(F_DS, ip);
    const double useg = LOSR(fu, ip);
    const double lam = (Hn - H) / trc + carry;
    double u_bar, p_bar, t_bar;
    if (CGA) {
      if (cu > 0) {
        const double pq = cp / cu, tq = ct / cu;
        const LkPart d = F32 ? lookup_partials_f32<true>(M.tab, tix, npt, mono, pdir, epf, pq, (float)tq, cu,
                                                         ipr, gi)
                             : lookup_partials<true>(M.tab, tix, npt, mono, pdir, epp, pq, tq, cu, ipr, gi);
        sU += lam * (d.du - (d.dp * pq + d.dt * tq) / cu);
        sP += lam * d.dp / cu;
        sT += lam * d.dt / cu;
        carry = lam * d.dep;
      } else {
        carry = lam;
      }
      u_bar = sU + p * sP + t * sT;
      p_bar = useg * sP;
      t_bar = useg * sT;
      cu -= useg;
      cp -= useg * p;
      ct -= useg * t;
    } else {
      const LkPart d = F32 ? lookup_partials_f32<false>(M.tab, tix, npt, mono, pdir, epf, p, (float)t, useg,
                                                        ipr, gi)
                           : lookup_partials<false>(M.tab, tix, npt, mono, pdir, epp, p, t, useg, ipr, gi);
      u_bar = lam * d.du;
      p_bar = lam * d.dp;
      t_bar = lam * d.dt;
      carry = lam * d.dep;
    }
    const double ucoef = 10 * p / (JUR_KB * t) * ds;
    double v[3];
    v[0] = t_bar - u_bar * useg / t;
    v[1] = p_bar + u_bar * useg / p;
    v[2] = u_bar * ucoef;
    const double gz = W.add(atm, z, lev, v, 1, p, xm, qi, kt);
    path_add(ip, gz, u_bar * useg / ds);
    Hn = H;
    trc = trp;
  }
  W.flush(xm, qi, kt);
  }
#undef EPS_AT
#undef SCR
}

// The FP32 EGA chains of all emitters of a row in one wavefront (EM >= ng,
// at most four; k_lookup_f32 stored each step's table partials).  The emitters
// share the point's LOS values, H and level pair: their T and p adjoints are
// summed per point before they enter the level window, and with PATH their
// parts of d rad / d z_i and d rad / d ds_i are added to the accumulation's
// in one read-modify-write per point.
template <int EM, bool PATH>
__global__ __launch_bounds__(WAVE) void k_adj_chains_f32(ModelView M, BatchView B, LosView L, EpsView E,
                                                         const RayInfo *info, int blk0, int nblk, JacView J) {
  const int nd = M.nd, ng = M.ng;
  int bl, id;
  if (!xcd_block_map(blockIdx.x, nd, nblk, bl, id)) return;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;
  const int ipf = B.ray_prof[ir];
  const int im = J.prof_map[ipf];
  if (im < 0 || im >= J.nmap) return;
  const int *xm = J.xmap + (size_t)im * JUR_XMAP_LEN;
  const atm_t *atm = &B.atm[ipf];
  const int npb = E.npb[bl];
  const int nef = J.nef[((size_t)bl * nd + id) * WAVE + lane];
  double *kt = J.kt + ((size_t)bl * nd + id) * J.ncol * WAVE + lane;
  const double *losb = L.los + (size_t)bl * L.cap * L.nfield * WAVE;
  const unsigned nfield = L.nfield;
  const float *eps0 = (const float *)E.eps + E.off[bl] + (size_t)id * ng * npb * WAVE;
  const float4 *pab0 = E.part + E.off[bl] + (size_t)id * ng * npb * WAVE;
  double *scb = J.scr + J.soff[bl] + (size_t)id * npb * J.nscr * WAVE;
  const unsigned nscr = J.nscr;
#define LOSQ(f, ip) ld32(losb, ((unsigned)(ip) * nfield + (unsigned)(f)) * WAVE + lane)
#define EPSQ(g, ip) ld32(eps0, ((unsigned)(g) * npb + (unsigned)(ip)) * WAVE + lane)
#define PARQ(g, ip) ld32(pab0, ((unsigned)(g) * npb + (unsigned)(ip)) * WAVE + lane)
#define SCQ(k, ip) scb[((size_t)(ip) * nscr + (k)) * WAVE + lane]
  // quantities of the level window: T, p, then q of each emitter
  int qi[2 + EM];
  bool on[EM];
  qi[0] = 1;
  qi[1] = 0;
#pragma unroll
  for (int g = 0; g < EM; g++) {
    on[g] = g < ng && M.tab.np[g * nd + id] >= 30;
    qi[2 + g] = on[g] ? 2 + g : -1;
  }
  LevelWindow<2 + EM, (1u << (2 + EM)) - 1u, true, PATH> W;
  int lev = -1;
  double carry[EM], trc[EM], useg[EM];
  float epf[EM];
  float4 dc[EM];
  double p = 0, t = 0, Hc = 0, zc = 0, dsc = 0, gzc = 0, gdc = 0, Hn = 0;
#pragma unroll
  for (int g = 0; g < EM; g++) {
    carry[g] = 0;
    trc[g] = 1;
    useg[g] = 0;
    epf[g] = 0.f;
    dc[g] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (nef > 0) {
    const int ip = nef - 1;
    p = LOSQ(F_P, ip);
    t = LOSQ(F_T, ip);
    zc = LOSQ(F_Z, ip);
    dsc = LOSQ(F_DS, ip);
    Hc = SCQ(0, ip);
    if (PATH) {
      gzc = SCQ(J.kz, ip);
      gdc = SCQ(J.kz + 1, ip);
    }
#pragma unroll
    for (int g = 0; g < EM; g++) {
      if (!on[g]) continue;
      trc[g] = 1.0 - (double)EPSQ(g, ip);
      dc[g] = PARQ(g, ip);
      useg[g] = LOSQ(L.f_u + g, ip);
      epf[g] = ip > 0 ? EPSQ(g, ip - 1) : 0.f;
    }
  }
  for (int ip = nef - 1; ip >= 0; ip--) {
    // the next point's inputs
    const int ipn = ip > 0 ? ip - 1 : 0;
    const double pn = LOSQ(F_P, ipn), tn = LOSQ(F_T, ipn);
    const double zx = LOSQ(F_Z, ipn), dsx = LOSQ(F_DS, ipn);
    const double Hx = SCQ(0, ipn);
    double gzx = 0, gdx = 0;
    if (PATH) {
      gzx = SCQ(J.kz, ipn);
      gdx = SCQ(J.kz + 1, ipn);
    }
    float4 dx[EM];
    double un[EM];
    float en[EM];
#pragma unroll
    for (int g = 0; g < EM; g++) {
      dx[g] = on[g] ? PARQ(g, ipn) : make_float4(0.f, 0.f, 0.f, 0.f);
      un[g] = on[g] ? LOSQ(L.f_u + g, ipn) : 0.0;
      en[g] = (on[g] && ipn > 0) ? EPSQ(g, ipn - 1) : 0.f;
    }
    const double H = Hc, ds = dsc;
    // the chain rule's divisions by the transmittance, T and p as float
    // reciprocals (relative error 1e-7, far below the FP32 lookups')
    const double rt = (double)frcp((float)t), rp = (double)frcp((float)p);
    const double ucoef = 10 * p * rt * ds * (1 / JUR_KB);
    double v[2 + EM];
    double vt = 0, vp = 0, gu = 0;
#pragma unroll
    for (int g = 0; g < EM; g++) {
      v[2 + g] = 0;
      if (!on[g]) continue;
      const double lam = (Hn - H) * (double)frcp((float)trc[g]) + carry[g];
      // the stored partials are (d eps_prev, d p, d T, d u), as LkPart
      const double u_bar = lam * dc[g].w;
      carry[g] = lam * dc[g].x;
      vt += lam * dc[g].z - u_bar * useg[g] * rt;
      vp += lam * dc[g].y + u_bar * useg[g] * rp;
      v[2 + g] = u_bar * ucoef;
      gu += u_bar * useg[g];
      trc[g] = ip > 0 ? 1.0 - (double)epf[g] : 1.0;
    }
    v[0] = vt;
    v[1] = vp;
    const double gz = W.add(atm, zc, lev, v, 1, p, xm, qi, kt);
    if (PATH) {
      SCQ(J.kz, ip) = gzc + gz;
      SCQ(J.kz + 1, ip) = gdc + gu * rcp_d(ds);
    }
    Hn = H;
    p = pn;
    t = tn;
    zc = zx;
    dsc = dsx;
    Hc = Hx;
    gzc = gzx;
    gdc = gdx;
#pragma unroll
    for (int g = 0; g < EM; g++) {
      dc[g] = dx[g];
      useg[g] = un[g];
      epf[g] = en[g];
    }
  }
  W.flush(xm, qi, kt);
#undef LOSQ
#undef EPSQ
#undef PARQ
#undef SCQ
}

// ------------------------------------------------------------------
// Adjoint of the refracted path (raytrace, src/jurassic.c:4201-4438).
//
// Below 60 km the march bends the ray by the refractivity N = 7.753e-5 p/T
// (jurassic.h:197-198) at the point (n = 1 + N) and at four probes (the
// step's midpoint and three points 0.02 km off it, for the gradient), so
// every LOS point after the first depends on p and T of the levels the
// path passed.  The reverse sweep runs each step of the march backwards,
// step k taking the state (x_k, e_k) to
//   x_k+1 = x_k + ds_k/2 (e_k + e'),  e' = v/|v|,  v = n e_k + ds_k grad N,
//   ds_k = min(RAYDS, RAYDZ / |e_k . x_k/|x_k||),
// with the adjoint (X, E) of (x_k+1, e_k+1) and the point values
// d rad / d z_k and d rad / d ds_k that the accumulation and the emitter
// chains stored (ds_k entering through the trapezoid rule, the last point's
// through the clip of the step that left the altitude range, k_ray_finish).
// Each probe's and point's N goes to its two levels with intpol_atm's
// weights (LOGY in p, LIN in T).
//
// The reverse step is linear in (X, E) and its coefficients do not depend
// on the channel: k_path_coef recomputes each forward step once (in
// parallel over steps) and stores the coefficients; k_adj_path applies them
// to the adjoints of CH channels per lane.
// ------------------------------------------------------------------

// p, t and their altitude slopes at z in level pair lo (weight w), and the
// LOGY weights of p on the two levels
struct PtAt {
  int lo;
  double w, p, t, dp, dt, fp0, fp1;
};
__device__ __forceinline__ PtAt pt_at(const atm_t *atm, double z, int hint) {
  PtAt r;
  r.lo = level_of(atm, z, hint, r.w);
  const double z0 = atm->z[r.lo], z1 = atm->z[r.lo + 1];
  const double P0 = atm->p[r.lo], P1 = atm->p[r.lo + 1];
  const double T0 = atm->t[r.lo], T1 = atm->t[r.lo + 1];
  const double rz = 1 / (z1 - z0);
  if (P1 / P0 > 0) {
    const double lq = log(P1 / P0) * rz;
    r.p = P0 * exp(lq * (z - z0));
    r.dp = r.p * lq;
    r.fp0 = r.p / P0 * (1 - r.w);
    r.fp1 = r.p / P1 * r.w;
  } else {
    r.p = P0 + (P1 - P0) * rz * (z - z0);
    r.dp = (P1 - P0) * rz;
    r.fp0 = 1 - r.w;
    r.fp1 = r.w;
  }
  r.t = T0 + (T1 - T0) * rz * (z - z0);
  r.dt = (T1 - T0) * rz;
  return r;
}

constexpr double PATH_CR = 7.753e-05, PATH_HP = 0.02, PATH_ZREF = 60;

// One forward step of the march from (x, e), recomputed: its length, the
// refractive index and gradient, the new direction e1, and the p and T of
// the point (pk) and of the four probes (pq, at xq).
struct MarchStep {
  double r, xh[3], ds, n, ng[3], nv, e1[3];
  bool rf;
  PtAt pk, pq[4];
  double xq[4][3];
};
__device__ __forceinline__ void march_step(const ModelView &M, const atm_t *atm, const double *x, const double *e,
                                           int &hint, MarchStep &S) {
  S.r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int i = 0; i < 3; i++) S.xh[i] = x[i] / S.r;
  const double z = S.r - JUR_RE;
  S.ds = M.rayds;
  if (M.raydz > 0) {
    const double ca = fabs((e[0] * x[0] + e[1] * x[1] + e[2] * x[2]) / S.r);
    if (ca != 0) S.ds = (M.rayds < M.raydz / ca) ? M.rayds : M.raydz / ca;
  }
  S.rf = M.refrac && z <= PATH_ZREF;
  S.n = 1;
  S.ng[0] = S.ng[1] = S.ng[2] = 0;
  if (S.rf) {
    S.pk = pt_at(atm, z, hint);
    hint = S.pk.lo;
    S.n = 1 + PATH_CR * S.pk.p / S.pk.t;
    for (int i = 0; i < 3; i++) S.xq[0][i] = x[i] + 0.5 * S.ds * e[i];
#pragma unroll
    for (int q = 1; q < 4; q++)
#pragma unroll
      for (int i = 0; i < 3; i++) S.xq[q][i] = S.xq[0][i] + (i == q - 1 ? PATH_HP : 0.0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double zq =
          sqrt(S.xq[q][0] * S.xq[q][0] + S.xq[q][1] * S.xq[q][1] + S.xq[q][2] * S.xq[q][2]) - JUR_RE;
      S.pq[q] = pt_at(atm, zq, S.pk.lo);
    }
    const double N0 = PATH_CR * S.pq[0].p / S.pq[0].t;
    for (int i = 0; i < 3; i++) S.ng[i] = (PATH_CR * S.pq[i + 1].p / S.pq[i + 1].t - N0) / PATH_HP;
  }
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = e[i] * S.n + S.ds * S.ng[i];
  S.nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int i = 0; i < 3; i++) S.e1[i] = v[i] / S.nv;
}

// Coefficients of the reverse of step k, per channel with the adjoint (X, E)
// of (x_k+1, e_k+1) and the point values gz = d rad / d z_k, gs = the
// step-length share of d rad / d ds:
//   V   = (G - e1 (e1 . G)) / |v|,  G = E + ds/2 X
//   gds = (e + e1)/2 . X + NGM . V + gs
//   X'  = X + MX V + (gz + SK e . V) XH + gds A
//   E'  = ds/2 (X + MX V) + n V + gds B
// MX V is the probes' share (d x_k of the four refractivities), NGM = grad N
// + MX^T e / 2, A and B the step length's dependence on (x_k, e_k).  The
// refractivities' adjoints e . V (point) and ds/HP (-sum V, V_0, V_1, V_2)
// (probes) go to the levels LB, LB + 1, LB + 2 as TAU_j . V (temperature)
// and PI_j . V (pressure); LB = -1: above 60 km, no level terms.  LB = -2:
// the five points span more than three levels; then slots PC_W + 4q hold
// point q's factors dN/dT (1 - w), dN/dT w, dN/dp f_lo, dN/dp f_hi and
// PC_LO its level pair (8 bits per point).
constexpr int PC_HDS = 0, PC_E = 1, PC_E1 = 4, PC_RNV = 7, PC_N = 8, PC_NGM = 9, PC_MX = 12, PC_SK = 21,
              PC_XH = 22, PC_A = 25, PC_B = 28, PC_LB = 31, PC_TAU = 32, PC_PI = 41, PC_W = 32, PC_LO = 52;
constexpr int PC_USED = PC_LO + 1;  // coefficients in use (JUR_NPCO: the stride)
static_assert(PC_USED <= JUR_NPCO && JUR_NPCO * WAVE * 8 % 1024 == 0 && NP <= 256,
              "reverse-step coefficient layout");
constexpr int PCO_STEPS = 16;  // steps per lane of k_path_coef

// The point's and the probes' refractivity adjoints as linear forms of V:
// point e . V, probe 0 -f (V_0 + V_1 + V_2), probe q f V_q-1 (f = ds/HP).
__device__ __forceinline__ void point_forms(const double *e, double f, double c[5][3]) {
  for (int i = 0; i < 3; i++) {
    c[0][i] = e[i];
    c[1][i] = -f;
    for (int q = 0; q < 3; q++) c[2 + q][i] = i == q ? f : 0.0;
  }
}

// LM (lane-major, for the small-batch sweep that stages each lane's step in
// LDS): step k of lane l at ((k 64 + l) JUR_NPCO + f); otherwise
// coefficient-major, (k JUR_NPCO + f) 64 + l.
template <bool LM>
__global__ __launch_bounds__(WAVE) void k_path_coef(ModelView M, BatchView B, LosView L, const RayInfo *info,
                                                    int blk0, int nblk, int nchunk, JacView J) {
  int bl, chk;
  if (!xcd_block_map(blockIdx.x, nchunk, nblk, bl, chk)) return;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK || R.np < 2) return;
  const int k0 = chk * PCO_STEPS;
  const int k1 = k0 + PCO_STEPS < R.np - 1 ? k0 + PCO_STEPS : R.np - 1;
  if (k0 >= k1) return;
  const atm_t *atm = &B.atm[B.ray_prof[ir]];
  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  double *pc = J.pco + J.soff[bl] / ((size_t)M.nd * J.nscr) * JUR_NPCO + (LM ? (size_t)lane * JUR_NPCO : lane);
  int hint = -1;
  for (int k = k0; k < k1; k++) {
    double x[3], e[3];
    for (int i = 0; i < 3; i++) {
      x[i] = LOSR(L.f_xe + i, k);
      e[i] = LOSR(L.f_xe + 3 + i, k);
    }
    MarchStep S;
    march_step(M, atm, x, e, hint, S);
    double *o = pc + (size_t)k * JUR_NPCO * WAVE;
#define PCW(f, v) o[LM ? (size_t)(f) : (size_t)(f) * WAVE] = (v)
    PCW(PC_HDS, 0.5 * S.ds);
    PCW(PC_RNV, 1 / S.nv);
    PCW(PC_N, S.n);
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    if (M.raydz > 0) {
      const double cc = e[0] * S.xh[0] + e[1] * S.xh[1] + e[2] * S.xh[2];
      if (cc != 0 && M.raydz / fabs(cc) < M.rayds) {
        const double cdz = -M.raydz * (cc > 0 ? 1.0 : -1.0) / (cc * cc);
        for (int i = 0; i < 3; i++) {
          b[i] = cdz * S.xh[i];
          a[i] = cdz * (e[i] - cc * S.xh[i]) / S.r;
        }
      }
    }
    double mx[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, ngm[3], sk = 0;
    double tau[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, pi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < 3; i++) ngm[i] = S.ng[i];
    int lb = -1;
    if (S.rf) {
      const double f = S.ds / PATH_HP;
      double u[4][3];
      for (int q = 0; q < 4; q++) {
        const PtAt &P = S.pq[q];
        const double sq = PATH_CR / P.t * P.dp - PATH_CR * P.p / (P.t * P.t) * P.dt;
        const double rq = sqrt(S.xq[q][0] * S.xq[q][0] + S.xq[q][1] * S.xq[q][1] + S.xq[q][2] * S.xq[q][2]);
        for (int i = 0; i < 3; i++) u[q][i] = sq * S.xq[q][i] / rq;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) mx[i][j] = f * (u[j + 1][i] - u[0][i]);
      for (int j = 0; j < 3; j++) ngm[j] += 0.5 * (e[0] * mx[0][j] + e[1] * mx[1][j] + e[2] * mx[2][j]);
      sk = PATH_CR / S.pk.t * S.pk.dp - PATH_CR * S.pk.p / (S.pk.t * S.pk.t) * S.pk.dt;
      const int lo = min(min(min(S.pk.lo, S.pq[0].lo), min(S.pq[1].lo, S.pq[2].lo)), S.pq[3].lo);
      const int hi = max(max(max(S.pk.lo, S.pq[0].lo), max(S.pq[1].lo, S.pq[2].lo)), S.pq[3].lo);
      if (hi > lo + 1) {
        lb = -2;
        auto wide = [&](int q, const PtAt &P) {
          const double dNdt = -PATH_CR * P.p / (P.t * P.t), dNdp = PATH_CR / P.t;
          PCW(PC_W + 4 * q, dNdt * (1 - P.w));
          PCW(PC_W + 4 * q + 1, dNdt * P.w);
          PCW(PC_W + 4 * q + 2, dNdp * P.fp0);
          PCW(PC_W + 4 * q + 3, dNdp * P.fp1);
        };
        wide(0, S.pk);
        wide(1, S.pq[0]);
        wide(2, S.pq[1]);
        wide(3, S.pq[2]);
        wide(4, S.pq[3]);
        const unsigned long long los5 = (unsigned long long)S.pk.lo | (unsigned long long)S.pq[0].lo << 8 |
                                        (unsigned long long)S.pq[1].lo << 16 |
                                        (unsigned long long)S.pq[2].lo << 24 | (unsigned long long)S.pq[3].lo << 32;
        PCW(PC_LO, (double)los5);
      } else {
        lb = lo;
        double c[5][3];
        point_forms(e, f, c);
        // point P with form c: its two levels' shares (P.lo is lb or lb + 1)
        auto addpt = [&](const PtAt &P, const double *cq) {
          const double dNdt = -PATH_CR * P.p / (P.t * P.t), dNdp = PATH_CR / P.t;
          const double t0 = dNdt * (1 - P.w), t1 = dNdt * P.w, p0 = dNdp * P.fp0, p1 = dNdp * P.fp1;
          const bool up = P.lo != lb;
#pragma unroll
          for (int i = 0; i < 3; i++) {
            tau[0][i] += up ? 0.0 : t0 * cq[i];
            tau[1][i] += (up ? t0 : t1) * cq[i];
            tau[2][i] += up ? t1 * cq[i] : 0.0;
            pi[0][i] += up ? 0.0 : p0 * cq[i];
            pi[1][i] += (up ? p0 : p1) * cq[i];
            pi[2][i] += up ? p1 * cq[i] : 0.0;
          }
        };
        addpt(S.pk, c[0]);
        addpt(S.pq[0], c[1]);
        addpt(S.pq[1], c[2]);
        addpt(S.pq[2], c[3]);
        addpt(S.pq[3], c[4]);
      }
    }
    for (int i = 0; i < 3; i++) {
      PCW(PC_E + i, e[i]);
      PCW(PC_E1 + i, S.e1[i]);
      PCW(PC_NGM + i, ngm[i]);
      PCW(PC_XH + i, S.xh[i]);
      PCW(PC_A + i, a[i]);
      PCW(PC_B + i, b[i]);
      for (int j = 0; j < 3; j++) {
        PCW(PC_MX + 3 * i + j, mx[i][j]);
        if (lb != -2) {
          PCW(PC_TAU + 3 * i + j, tau[i][j]);
          PCW(PC_PI + 3 * i + j, pi[i][j]);
        }
      }
    }
    PCW(PC_SK, sk);
    PCW(PC_LB, (double)lb);
#undef PCW
  }
}

// The reverse march of CH channels of one ray (a lane): the adjoints (X, E)
// of the state, the level window, and the clip of the last step.  Level sums
// are kept for a window of three levels (the step's LB and the two above)
// per channel and added to the K tile when the window moves.
template <int CH>
struct PathRev {
  const int *xm;
  int nlev, nc;
  double *kt0;
  size_t ktstride;
  double X[CH][3], Ev[CH][3], Xp[CH][3], gzp[CH], gdn[CH], Tw[CH][3], Pw[CH][3];
  int wb;

  // The clipped last point: x_l = x_p + frac (x_cut - x_p), its raw step
  // replacing point N-2's: ds_cut frac.  gz: d rad / d z of point N-1; gdl,
  // gdm: d rad / d ds of points N-1 and N-2.
  __device__ __forceinline__ void init(const ModelView &M, const RayTail &T, const double *xl, const double *gz,
                                       const double *gdl, const double *gdm) {
    wb = -1;
    const double zp = T.zprev, zc = T.zcut;
    const double zb = zc < T.zmin ? T.zmin : T.zmax;
    const double frac = (zb - zp) / (zc - zp);
    const double rl = sqrt(xl[0] * xl[0] + xl[1] * xl[1] + xl[2] * xl[2]);
    const double rc = sqrt(T.xcut[0] * T.xcut[0] + T.xcut[1] * T.xcut[1] + T.xcut[2] * T.xcut[2]);
    // d ds / d (x, e) at the cut point
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    if (M.raydz > 0) {
      double xh[3];
      for (int i = 0; i < 3; i++) xh[i] = T.xcut[i] / rc;
      const double cc = T.ecut[0] * xh[0] + T.ecut[1] * xh[1] + T.ecut[2] * xh[2];
      if (cc != 0 && M.raydz / fabs(cc) < M.rayds) {
        const double cdz = -M.raydz * (cc > 0 ? 1.0 : -1.0) / (cc * cc);
        for (int i = 0; i < 3; i++) {
          b[i] = cdz * xh[i];
          a[i] = cdz * (T.ecut[i] - cc * xh[i]) / rc;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CH; c++) {
      gdn[c] = gdm[c];
      const double gs = 0.5 * (gdm[c] + gdl[c]);  // d rad / d (clipped raw ds of point N-2)
      double gfrac = T.dscut * gs;
      for (int i = 0; i < 3; i++) {
        const double xli = gz[c] * xl[i] / rl;
        Xp[c][i] = (1 - frac) * xli;
        X[c][i] = frac * xli;
        gfrac += xli * (T.xcut[i] - T.xprev[i]);
      }
      gzp[c] = gfrac * (frac - 1) / (zc - zp);
      const double gzc = -gfrac * frac / (zc - zp);
      for (int i = 0; i < 3; i++) {
        X[c][i] += gzc * T.xcut[i] / rc + frac * gs * a[i];
        Ev[c][i] = frac * gs * b[i];
      }
    }
  }

  __device__ __forceinline__ void flush_level(int j) {
    const int l = wb + j;
    if (l >= nlev) return;
    const int ct = xm[JUR_XMAP_LEV(1, l)], cp = xm[JUR_XMAP_LEV(0, l)];
#pragma unroll
    for (int c = 0; c < CH; c++) {
      if (c >= nc) continue;
      double *kt = kt0 + (size_t)c * ktstride;
      if (ct >= 0) kt[(size_t)ct * WAVE] += Tw[c][j];
      if (cp >= 0) kt[(size_t)cp * WAVE] += Pw[c][j];
    }
  }

  // The reverse of march step k (of N points) with its coefficients cf and
  // the channels' point values gz = d rad / d z_k, gd = d rad / d ds_k.
  // LAY 0: coefficient f at cf[f]; LAY 1 (a lane's coefficients staged in
  // LDS by 16 B direct-to-LDS loads, pair i at i KiB + lane 16 B): at
  // cf[(f / 2) 128 + f % 2] with cf at the lane's first pair.
  template <int LAY = 0>
  __device__ __forceinline__ void step(int k, int N, const double *cf, const double *gzv, const double *gdv) {
#define PCR(f) (LAY == 0 ? cf[f] : cf[((f) >> 1) * 128 + ((f) & 1)])
    const double hds = PCR(PC_HDS), rnv = PCR(PC_RNV), n = PCR(PC_N), sk = PCR(PC_SK);
    double e[3], e1[3], ngm[3], xh[3], a[3], b[3], mx[9];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      e[i] = PCR(PC_E + i);
      e1[i] = PCR(PC_E1 + i);
      ngm[i] = PCR(PC_NGM + i);
      xh[i] = PCR(PC_XH + i);
      a[i] = PCR(PC_A + i);
      b[i] = PCR(PC_B + i);
    }
#pragma unroll
    for (int i = 0; i < 9; i++) mx[i] = PCR(PC_MX + i);
    const int lb = (int)PCR(PC_LB);
    double V[CH][3];
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const double gz = gzv[c], gd = gdv[c];
      const double gs = (k == N - 2) ? 0.0 : 0.5 * (gd + gdn[c]);  // point N-2's raw ds was clipped
      gdn[c] = gd;
      double G[3];
      for (int i = 0; i < 3; i++) G[i] = Ev[c][i] + hds * X[c][i];
      const double pe = e1[0] * G[0] + e1[1] * G[1] + e1[2] * G[2];
      double v[3];
      for (int i = 0; i < 3; i++) v[i] = (G[i] - e1[i] * pe) * rnv;
      const double gn = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
      double gds = gs;
      for (int i = 0; i < 3; i++) gds += 0.5 * (e[i] + e1[i]) * X[c][i] + ngm[i] * v[i];
      const double gzk = gz + sk * gn;
      for (int i = 0; i < 3; i++) {
        const double xg = X[c][i] + mx[3 * i] * v[0] + mx[3 * i + 1] * v[1] + mx[3 * i + 2] * v[2];
        X[c][i] = xg + gzk * xh[i] + gds * a[i];
        Ev[c][i] = hds * xg + n * v[i] + gds * b[i];
        V[c][i] = v[i];
      }
    }
    if (lb >= 0) {
      if (lb != wb) {
        if (wb >= 0 && lb == wb + 1) {
          flush_level(0);
#pragma unroll
          for (int c = 0; c < CH; c++) {
            Tw[c][0] = Tw[c][1];
            Tw[c][1] = Tw[c][2];
            Tw[c][2] = 0;
            Pw[c][0] = Pw[c][1];
            Pw[c][1] = Pw[c][2];
            Pw[c][2] = 0;
          }
        } else if (wb >= 0 && lb == wb - 1) {
          flush_level(2);
#pragma unroll
          for (int c = 0; c < CH; c++) {
            Tw[c][2] = Tw[c][1];
            Tw[c][1] = Tw[c][0];
            Tw[c][0] = 0;
            Pw[c][2] = Pw[c][1];
            Pw[c][1] = Pw[c][0];
            Pw[c][0] = 0;
          }
        } else {
          if (wb >= 0)
            for (int j = 0; j < 3; j++) flush_level(j);
#pragma unroll
          for (int c = 0; c < CH; c++)
            for (int j = 0; j < 3; j++) Tw[c][j] = Pw[c][j] = 0;
        }
        wb = lb;
      }
      for (int j = 0; j < 3; j++) {
        double tj[3], pj[3];
        for (int i = 0; i < 3; i++) {
          tj[i] = PCR(PC_TAU + 3 * j + i);
          pj[i] = PCR(PC_PI + 3 * j + i);
        }
#pragma unroll
        for (int c = 0; c < CH; c++) {
          Tw[c][j] += tj[0] * V[c][0] + tj[1] * V[c][1] + tj[2] * V[c][2];
          Pw[c][j] += pj[0] * V[c][0] + pj[1] * V[c][1] + pj[2] * V[c][2];
        }
      }
    } else if (lb == -2) {
      // points over more than three levels: straight to the tile
      const unsigned long long los5 = (unsigned long long)PCR(PC_LO);
      const double f = hds * (2 / PATH_HP);
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const int lo = (int)(los5 >> (8 * q) & 255u);
        const double w0 = PCR(PC_W + 4 * q), w1 = PCR(PC_W + 4 * q + 1);
        const double w2 = PCR(PC_W + 4 * q + 2), w3 = PCR(PC_W + 4 * q + 3);
        const int t0 = xm[JUR_XMAP_LEV(1, lo)], t1 = xm[JUR_XMAP_LEV(1, lo + 1)];
        const int p0 = xm[JUR_XMAP_LEV(0, lo)], p1 = xm[JUR_XMAP_LEV(0, lo + 1)];
#pragma unroll
        for (int c = 0; c < CH; c++) {
          if (c >= nc) continue;
          const double g = q == 0 ? e[0] * V[c][0] + e[1] * V[c][1] + e[2] * V[c][2]
                                  : (q == 1 ? -f * (V[c][0] + V[c][1] + V[c][2]) : f * V[c][q >= 2 ? q - 2 : 0]);
          double *kt = kt0 + (size_t)c * ktstride;
          if (t0 >= 0) kt[(size_t)t0 * WAVE] += g * w0;
          if (t1 >= 0) kt[(size_t)t1 * WAVE] += g * w1;
          if (p0 >= 0) kt[(size_t)p0 * WAVE] += g * w2;
          if (p1 >= 0) kt[(size_t)p1 * WAVE] += g * w3;
        }
      }
    }
#undef PCR
  }

  // after step N-2 (which leads to the cut point): the clip's dependence on
  // x_N-2 joins its adjoint (cf: step N-2's coefficients)
  template <int LAY = 0>
  __device__ __forceinline__ void join_clip(const double *cf) {
#pragma unroll
    for (int c = 0; c < CH; c++)
      for (int i = 0; i < 3; i++) {
        const int f = PC_XH + i;
        X[c][i] += Xp[c][i] + gzp[c] * (LAY == 0 ? cf[f] : cf[(f >> 1) * 128 + (f & 1)]);
      }
  }

  __device__ __forceinline__ void finish() {
    if (wb >= 0)
      for (int j = 0; j < 3; j++) flush_level(j);
  }
};

// A step's inputs: its coefficients and the channels' point values.
template <int CH>
struct PathStepIn {
  double c[PC_USED];
  double gz[CH], gd[CH];
};

// Forward sweep, surface seeds, reverse sweep of the accumulation: writes
// rad, tau, H_i (scratch), the number of points before the first opaque
// segment, and the K entries of continua, source function, window
// extinction, clouds and surface.
// EMF > 0 (FP32 EGA with stored partials, ng <= EMF): the reverse sweep also
// runs the emitter chains of the row, point by point, as k_adj_chains_f32
// does; H_i is used where it is formed, and the chains' T, p and q adjoints
// join the accumulation's in one level window.  PFU (with PATH): it also
// runs the reverse march of the refracted path (PathRev, one channel) point
// by point from the stored step coefficients, so that the row's whole
// reverse sweep is one pass over the path: the form for small batches,
// where each kernel lasts as long as one wavefront's sweep.
template <typename SF, bool PATH, int EMF = 0, bool PFU = false>
__global__ __launch_bounds__(WAVE) void k_adj_acc(ModelView M, BatchView B, LosView L, EpsView E,
                                                  const RayInfo *info, const RayTail *tail, int blk0, int nblk,
                                                  JacView J, double *rad_out, double *tau_out) {
  const int nd = M.nd, ng = M.ng;
  int bl, id;
  if (!xcd_block_map(blockIdx.x, nd, nblk, bl, id)) return;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;
  const int ipf = B.ray_prof[ir];
  const int im = J.prof_map[ipf];
  const atm_t *atm = &B.atm[ipf];
  const ChanInfo C = M.chan[id];
  const double nu = C.nu;
  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const int npb = E.npb[bl];
  const SF *epsb = (const SF *)E.eps + E.off[bl] + (size_t)id * ng * npb * WAVE + lane;
#define EPS_AT(g, ip) epsb[((size_t)(g) * npb + (ip)) * WAVE]
  double *sc = J.scr + J.soff[bl] + (size_t)id * npb * J.nscr * WAVE + lane;
#define SCR(k, ip) sc[((size_t)(ip) * J.nscr + (k)) * WAVE]
  double *kt = J.kt + ((size_t)bl * nd + id) * J.ncol * WAVE + lane;
#define KT(col) kt[(size_t)(col) * WAVE]
  const double *sr_col = M.sr + (size_t)id * TBLNS;
  const double inv_dst = 1 / (M.st[1] - M.st[0]);

  unsigned gm = 0;  // emitters with a table for this channel
  for (int g = 0; g < ng; g++)
    if (M.tab.np[g * nd + id] >= 30) gm |= 1u << g;
  // product of the path transmittances of those emitters after point ip
  auto qprod = [&](int ip) {
    double q = 1;
    for (int g = 0; g < ng; g++)
      if (gm >> g & 1u) q *= trans_of(EPS_AT(g, ip));
    return q;
  };

  double kcl = 0, wcl = 0;
  const bool cloud = L.f_cl >= 0 && M.ncl > 0;
  if (cloud) {
    const double x0 = M.clnu[C.icl], x1 = M.clnu[C.icl + 1];
    kcl = lin(x0, atm->clk[C.icl], x1, atm->clk[C.icl + 1], nu);
    wcl = (nu - x0) / (x1 - x0);
  }
  double sfeps = 1.0, wsf = 0;
  if (M.nsf > 0) {
    const double x0 = M.sfnu[C.isf], x1 = M.sfnu[C.isf + 1];
    sfeps = lin(x0, atm->sfeps[C.isf], x1, atm->sfeps[C.isf + 1], nu);
    wsf = (nu - x0) / (x1 - x0);
  }
  const bool has_k = C.window < M.nw;
  const CtmChan K = ctm_chan(M.ctm, nu);
  // The Jacobian of the FP32 forward model (float continua and source
  // function interpolation) takes the continua's transcendentals and the
  // segment transmittance's exponential in float; the sweeps stay FP64.
  constexpr bool F32 = sizeof(SF) == sizeof(float);
  const CtmChanF KF = ctm_chan_f(K, nu);
  // Segment terms of point (p, t, ds): beta and its gradient, the source
  // function and its slope, and s = tg exp(-beta ds); rt = 1/t, rp = 1/p.
  auto segment = [&](double p, double t, double ds, double q_h2o, double u_co2, double u_h2o, double kwin,
                     double clx, double tg, BetaGrad &bg, double &src, double &dsrc, double &rt,
                     double &rp, double &beta) {
    if constexpr (F32) {
      rt = rcp_d(t);
      rp = rcp_d(p);
      beta = beta_grad_f(M, K, KF, nu, p, t, rt, rp, rcp_d(ds), q_h2o, u_co2, u_h2o, kwin, clx, bg);
      src = srcfunc_du(M.st, inv_dst, sr_col, t, dsrc);
      return tg * (double)__expf((float)(-beta * ds));
    } else {
      rt = 1 / t;
      rp = 1 / p;
      beta = beta_grad(M, K, nu, p, t, ds, q_h2o, u_co2, u_h2o, kwin, clx, bg);
      src = srcfunc_d(M.st, sr_col, t, dsrc);
      return tg * exp(-beta * ds);
    }
  };
  const int fq_h2o = M.ig_h2o >= 0 ? F_Q0 + M.ig_h2o : -1;
  const int fu_h2o = M.ig_h2o >= 0 ? L.f_u + M.ig_h2o : -1;
  const int fu_co2 = M.ig_co2 >= 0 ? L.f_u + M.ig_co2 : -1;

  // Point inputs of the sweeps, loaded one point ahead so that a point's
  // arithmetic does not wait for its own loads (absent fields read F_DS in
  // their place and are not used).  The emitters' stored values are loaded
  // with them and multiplied (in qprod's order) one iteration later.
  struct PtIn {
    double z, p, t, ds, qh, uc, uh, kw, cl;
  };
  const int f_qh = fq_h2o >= 0 ? fq_h2o : F_DS, f_uc = fu_co2 >= 0 ? fu_co2 : F_DS;
  const int f_uh = fu_h2o >= 0 ? fu_h2o : F_DS, f_kw = has_k ? L.f_k + C.window : F_DS;
  const int f_cl = cloud ? L.f_cl : F_DS;
  auto ld_pt = [&](int ip, PtIn &x, bool withz) {
    if (withz) x.z = LOSR(F_Z, ip);
    x.p = LOSR(F_P, ip);
    x.t = LOSR(F_T, ip);
    x.ds = LOSR(F_DS, ip);
    x.qh = LOSR(f_qh, ip);
    x.uc = LOSR(f_uc, ip);
    x.uh = LOSR(f_uh, ip);
    x.kw = LOSR(f_kw, ip);
    x.cl = LOSR(f_cl, ip);
  };
  SF en[NG];
  auto ld_eps = [&](int ip) {
#pragma unroll
    for (int g = 0; g < NG; g++) en[g] = (g < ng && (gm >> g & 1u)) ? EPS_AT(g, ip) : SF(0);
  };
  auto prod_en = [&]() {
    double q = 1;
#pragma unroll
    for (int g = 0; g < NG; g++)
      if (g < ng && (gm >> g & 1u)) q *= trans_of(en[g]);
    return q;
  };
  PtIn nx{};

  // ---------------- forward sweep: tau_i (and D_i) per point ----------------
  double rad = 0, tau = 1, dr = 0;
  int nef = R.np;
  {
    double qp = 1, qn = 1;
    if (R.np > 0) {
      ld_pt(0, nx, false);
      ld_eps(0);
      qn = prod_en();
    }
    for (int ip = 0; ip < R.np; ip++) {
      const PtIn cu = nx;
      const double q = qn;
      const int ipn = ip + 1 < R.np ? ip + 1 : ip;
      ld_pt(ipn, nx, false);
      ld_eps(ipn);
      if (!(q > 0)) {
        nef = ip;
        break;
      }
      const double tg = F32 ? q * rcp_d(qp) : q / qp;  // qp > 0: q > 0 was checked
      qp = q;
      BetaGrad bg;
      double src, dsrc, rt, rp, beta;
      const double s = segment(cu.p, cu.t, cu.ds, fq_h2o >= 0 ? cu.qh : 0.0, fu_co2 >= 0 ? cu.uc : 0.0,
                               fu_h2o >= 0 ? cu.uh : 0.0, has_k ? cu.kw : 0.0, cloud ? cu.cl * kcl : 0.0, tg,
                               bg, src, dsrc, rt, rp, beta);
      SCR(0, ip) = tau;
      if (J.nscr > 1) SCR(1, ip) = dr;
      rad += src * (1 - s) * tau;
      dr = dr * s + src * (1 - s);
      tau *= s;
      qn = prod_en();
    }
  }

  // surface emission, reflection and solar term, and their adjoint seeds
  double tb = 0, db = 0, sft_bar = 0, sfeps_bar = 0;
  if (M.sftype >= 1 && R.sft > 0) {
    double dsrc_sf;
    const double src_sf = srcfunc_d(M.st, sr_col, R.sft, dsrc_sf);
    rad += sfeps * src_sf * tau;
    tb += sfeps * src_sf;
    sfeps_bar += src_sf * tau;
    sft_bar = sfeps * tau * dsrc_sf;
    if (M.sftype >= 2 && sfeps < 1) {
      rad += dr * tau * (1 - sfeps);
      tb += dr * (1 - sfeps);
      db += tau * (1 - sfeps);
      sfeps_bar -= dr * tau;
      if (M.sftype >= 3 && R.pad > 0) {
        const double ks = 6.764e-5 / (2. * JUR_PI) * C.planck_sun * R.rcos;
        rad += ks * tau * (1 - sfeps) * tau;
        tb += 2 * ks * tau * (1 - sfeps);
        sfeps_bar -= ks * tau * tau;
      }
    }
  }
  rad_out[(size_t)ir * nd + id] = rad;
  tau_out[(size_t)ir * nd + id] = tau;
  J.nef[((size_t)bl * nd + id) * WAVE + lane] = nef;
  if (im < 0 || im >= J.nmap) return;  // no state map: K row stays zero
  const int *xm = J.xmap + (size_t)im * JUR_XMAP_LEN;

  // ---------------- reverse sweep of the accumulation ----------------
  // H_i replaces tau_i in the scratch; continua, source function, window
  // extinction and clouds contribute here, the emitter chains below.
  double clz_bar = 0, cldz_bar = 0, clk_bar = 0;
  // the surface temperature taken from the last LOS point's T: its slope
  // moves d rad / d z of that point (added here where the path adjoint runs
  // in this sweep, after the sweep otherwise)
  double gz_sft = 0;
  if (PFU && M.sftype >= 1 && R.sft > 0 && !(M.nsf > 0 && atm->sft > 0) && R.np > 0 && R.np - 1 < nef) {
    double w;
    const int lv = level_of(atm, LOSR(F_Z, R.np - 1), -1, w);
    gz_sft = sft_bar * (atm->t[lv + 1] - atm->t[lv]) / (atm->z[lv + 1] - atm->z[lv]);
  }
  {
    // level quantities: T, p, the window's extinction, then CO2 and H2O, or
    // with the chains every emitter's q
    constexpr int NQW = EMF > 0 ? 3 + EMF : 5;
    int qi[NQW];
    qi[0] = 1;
    qi[1] = 0;
    qi[2] = has_k ? 2 + NG + C.window : -1;
    if constexpr (EMF > 0) {
#pragma unroll
      for (int g = 0; g < EMF; g++) qi[3 + g] = g < ng ? 2 + g : -1;
    } else {
      qi[3] = M.ig_co2 >= 0 ? 2 + M.ig_co2 : -1;
      qi[4] = M.ig_h2o >= 0 ? 2 + M.ig_h2o : -1;
    }
    LevelWindow<NQW, (1u << NQW) - 1u, F32, PATH> W;
    // the path adjoint's point values d rad / d z_i, d rad / d ds_i: zero
    // beyond the first opaque segment, the emitter chains add theirs
    if (PATH && !PFU)
      for (int ip = nef; ip < R.np; ip++) SCR(J.kz, ip) = SCR(J.kz + 1, ip) = 0.0;
    int lev = -1;
    double qi_cur = nef > 0 ? qprod(nef - 1) : 1;
    // the chains' state: Lambda carry, the path transmittance after the
    // point (as the float the chain divides by), the point's stored partials
    // (dep, dp, dT, du) and column density, loaded one point ahead
    constexpr int EC = EMF > 0 ? EMF : 1;
    bool on[EC];
    double carry[EC], ucur[EC], unx[EC];
    float trc[EC];
    float4 dcur[EC], dnx[EC];
    const float4 *pab = EMF > 0 ? E.part + E.off[bl] + (size_t)id * ng * npb * WAVE + lane : nullptr;
#pragma unroll
    for (int g = 0; g < EC; g++) {
      on[g] = EMF > 0 && g < ng && (gm >> g & 1u);
      carry[g] = ucur[g] = unx[g] = 0;
      trc[g] = (on[g] && nef > 0) ? (float)trans_of(EPS_AT(g, nef - 1)) : 1.f;
      dcur[g] = dnx[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    double Hn = 0;
    // the path's reverse march (PFU): its state, the held values of the last
    // point, and the next step's coefficients loaded one point ahead
    const int N = R.np;
    PathRev<1> P;
    bool pstart = false;
    double gz_last = 0, gd_last = 0;
    // the next step's coefficients are staged in LDS (one wavefront per
    // workgroup: 27 KiB), copied by the wavefront's direct-to-LDS loads
    // while the next point's accumulation runs, instead of in registers
    const char *pcg = nullptr;
    double *pcs = nullptr;
    // (lane-major coefficients: each lane copies its own step's 432 B; the
    // lanes of a wavefront are at different steps)
    auto pc_fetch = [&](int k) {
      if constexpr (PFU) {
        // the reads of the last step's coefficients have completed before
        // the copy overwrites them; the memory clobber keeps the compiler
        // from moving those reads below the copy, whose LDS write it does
        // not see
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        const char *g = pcg + ((size_t)k * WAVE + lane) * JUR_NPCO * sizeof(double);
#pragma unroll
        for (int i = 0; i < JUR_NPCO / 2; i++)
          __builtin_amdgcn_global_load_lds((const void *)(g + i * 16),
                                           (__attribute__((address_space(3))) void *)((char *)pcs + i * 1024),
                                           16, 0, 0);
      }
    };
    if constexpr (PFU) {
      __shared__ double pc_lds[JUR_NPCO * WAVE];
      pcs = pc_lds;
      P.xm = xm;
      P.nlev = atm->np;
      P.nc = 1;
      P.kt0 = kt;
      P.ktstride = (size_t)J.ncol * WAVE;
      pcg = (const char *)(J.pco + J.soff[bl] / ((size_t)nd * J.nscr) * JUR_NPCO);
      const int k0 = nef - 1 < N - 2 ? nef - 1 : N - 2;
      if (k0 >= 0) pc_fetch(k0);
    }
    // point ip's inputs and scratch, and the emitters' values at ip - 1,
    // loaded one iteration ahead
    double sc0n = 0, sc1n = 0;
    auto ld_rev = [&](int ip) {
      ld_pt(ip, nx, true);
      sc0n = SCR(0, ip);
      sc1n = J.nscr > 1 ? SCR(1, ip) : 0.0;
      if (ip > 0) ld_eps(ip - 1);
      if constexpr (EMF > 0) {
#pragma unroll
        for (int g = 0; g < EMF; g++)
          if (on[g]) {
            dnx[g] = pab[((size_t)g * npb + ip) * WAVE];
            unx[g] = LOSR(L.f_u + g, ip);
          }
      }
    };
    if (nef > 0) ld_rev(nef - 1);
    for (int ip = nef - 1; ip >= 0; ip--) {
      const PtIn cu = nx;
      const double tau_i = sc0n, dr_i = sc1n;
      const double qm = ip > 0 ? prod_en() : 1;
      float em1[EC];  // the emitters' path emissivities before point ip
#pragma unroll
      for (int g = 0; g < EC; g++) {
        em1[g] = EMF > 0 ? (float)en[g] : 0.f;
        dcur[g] = dnx[g];
        ucur[g] = unx[g];
      }
      if (ip > 0) ld_rev(ip - 1);
      const double tg = F32 ? qi_cur * rcp_d(qm) : qi_cur / qm;
      qi_cur = qm;
      const double z = cu.z, p = cu.p, t = cu.t, ds = cu.ds;
      const double q_h2o = fq_h2o >= 0 ? cu.qh : 0.0;
      const double u_co2 = fu_co2 >= 0 ? cu.uc : 0.0;
      const double u_h2o = fu_h2o >= 0 ? cu.uh : 0.0;
      const double aux = cloud ? cu.cl : 0.0;
      BetaGrad bg;
      double src, dsrc, rt, rp, beta;
      const double s = segment(p, t, ds, q_h2o, u_co2, u_h2o, has_k ? cu.kw : 0.0, aux * kcl, tg, bg, src, dsrc,
                               rt, rp, beta);
      const double G = tau_i * (tb - src) + db * (dr_i - src);
      const double src_bar = (1 - s) * (tau_i + db);
      tb = src * (1 - s) + tb * s;
      db = db * s;
      const double H = G * s;
      if (EMF == 0) SCR(0, ip) = H;
      const double beta_bar = -H * ds;

      // u = 10 q p / (kB T) ds
      const double ucoef = F32 ? 10 * p * rt * ds * (1 / JUR_KB) : 10 * p / (JUR_KB * t) * ds;
      const double uc_bar = beta_bar * bg.du_co2, uh_bar = beta_bar * bg.du_h2o;
      const double uu = uc_bar * u_co2 + uh_bar * u_h2o;
      double v[NQW];
      v[0] = beta_bar * bg.dt + src_bar * dsrc - (F32 ? uu * rt : uu / t);  // T
      v[1] = beta_bar * bg.dp + (F32 ? uu * rp : uu / p);                   // p
      v[2] = beta_bar;                                                                  // k
      const double vco2 = uc_bar * ucoef, vh2o = uh_bar * ucoef + beta_bar * bg.dq_h2o;
      double gu = 0;  // the chains' sum of u_bar u_seg (d rad / d ds_i through u)
      if constexpr (EMF > 0) {
        // the chains (as k_adj_chains_f32: float reciprocals of T, p and the
        // transmittance, whose 1e-7 is far below the FP32 lookups')
        const double rtf = (double)frcp((float)t), rpf = (double)frcp((float)p);
        const double ucf = 10 * p * rtf * ds * (1 / JUR_KB);
        double vt = 0, vp = 0;
#pragma unroll
        for (int g = 0; g < EMF; g++) {
          double vq = (g == M.ig_co2 ? vco2 : 0.0) + (g == M.ig_h2o ? vh2o : 0.0);
          if (on[g]) {
            const double lam = (Hn - H) * (double)frcp(trc[g]) + carry[g];
            const double u_bar = lam * dcur[g].w;
            carry[g] = lam * dcur[g].x;
            vt += lam * dcur[g].z - u_bar * ucur[g] * rtf;
            vp += lam * dcur[g].y + u_bar * ucur[g] * rpf;
            vq += u_bar * ucf;
            gu += u_bar * ucur[g];
            trc[g] = ip > 0 ? (float)(1.0 - (double)em1[g]) : 1.f;
          }
          v[3 + g] = vq;
        }
        v[0] += vt;
        v[1] += vp;
        Hn = H;
      } else {
        v[3] = vco2;  // q_co2
        v[4] = vh2o;  // q_h2o
      }
      const double gz = W.add(atm, z, lev, v, 1, p, xm, qi, kt);
      if (PATH) {
        // altitude: the level quantities' slopes and the cloud profile's;
        // segment length: exp(-beta ds) (the continua's beta does not
        // depend on ds, their u ~ ds cancels their 1/ds), the column
        // densities' part comes from the emitter chains
        double gzc = gz;
        if (cloud && atm->cldz > 0) gzc -= beta_bar * kcl * aux * (z - atm->clz) / (atm->cldz * atm->cldz);
        const double gdc = EMF > 0 ? -H * beta + gu * rcp_d(ds) : -H * beta;
        if constexpr (PFU) {
          if (ip == N - 1) {
            gz_last = gzc + gz_sft;
            gd_last = gdc;
          } else {
            if (!pstart) {
              // the first step reached: the clip of the last step, with
              // point N-2's d rad / d ds where this is that point
              double xl[3];
              for (int i = 0; i < 3; i++) xl[i] = LOSR(L.f_xe + i, N - 1);
              const double gdm = ip == N - 2 ? gdc : 0.0;
              P.init(M, tail[ir], xl, &gz_last, &gd_last, &gdm);
              pstart = true;
            }
            // the copy of this step's coefficients has landed in LDS: the
            // direct-to-LDS loads count as vector memory loads, which the
            // compiler does not track for LDS reads (the memory clobber
            // keeps the reads below the wait)
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            P.template step<1>(ip, N, pcs + 2 * lane, &gzc, &gdc);
            if (ip == N - 2) P.template join_clip<1>(pcs + 2 * lane);
            // the next step's coefficients, in flight during the next
            // point's accumulation
            if (ip > 0) pc_fetch(ip - 1);
          }
        } else {
          SCR(J.kz, ip) = gzc;
          SCR(J.kz + 1, ip) = gdc;
        }
      }

      if (cloud) {
        clk_bar += beta_bar * aux;
        if (atm->cldz > 0) {
          const double cl_bar = beta_bar * kcl;
          const double dz = (z - atm->clz) / atm->cldz;
          clz_bar += cl_bar * aux * dz / atm->cldz;
          cldz_bar += cl_bar * aux * dz * dz / atm->cldz;
        }
      }
    }
    if (PFU && pstart) P.finish();
    W.flush(xm, qi, kt);
  }

  // surface terms
  if (M.sftype >= 1 && R.sft > 0) {
    if (M.nsf > 0 && atm->sft > 0) {
      const int c = xm[JUR_XMAP_SFT];
      if (c >= 0) KT(c) += sft_bar;
    } else if (R.np > 0) {
      // the surface temperature is the LOS temperature of the last point
      double w;
      const int lv = level_of(atm, LOSR(F_Z, R.np - 1), -1, w);
      int c;
      if ((c = xm[JUR_XMAP_LEV(1, lv)]) >= 0) KT(c) += (1 - w) * sft_bar;
      if ((c = xm[JUR_XMAP_LEV(1, lv + 1)]) >= 0) KT(c) += w * sft_bar;
      if (PATH && !PFU && R.np - 1 < nef)
        SCR(J.kz, R.np - 1) += sft_bar * (atm->t[lv + 1] - atm->t[lv]) / (atm->z[lv + 1] - atm->z[lv]);
    }
    if (M.nsf > 0) {
      int c;
      if ((c = xm[JUR_XMAP_SFEPS(C.isf)]) >= 0) KT(c) += (1 - wsf) * sfeps_bar;
      if (C.isf + 1 < NSF && (c = xm[JUR_XMAP_SFEPS(C.isf + 1)]) >= 0) KT(c) += wsf * sfeps_bar;
    }
  }
  if (cloud) {
    int c;
    if ((c = xm[JUR_XMAP_CLK(C.icl)]) >= 0) KT(c) += (1 - wcl) * clk_bar;
    if (C.icl + 1 < NCL && (c = xm[JUR_XMAP_CLK(C.icl + 1)]) >= 0) KT(c) += wcl * clk_bar;
    if ((c = xm[JUR_XMAP_CLZ]) >= 0) KT(c) += clz_bar;
    if ((c = xm[JUR_XMAP_CLDZ]) >= 0) KT(c) += cldz_bar;
  }
#undef EPS_AT
#undef SCR
#undef KT
}

// The reverse sweep over the steps, CH channels per lane (one ray).  With
// PF the inputs of the next step are loaded while a step is computed (for
// small batches, where a wavefront's sweep is the kernel's duration).
template <int CH, bool PF>
__global__ __launch_bounds__(WAVE) void k_adj_path(ModelView M, BatchView B, LosView L, EpsView E,
                                                   const RayInfo *info, const RayTail *tail, int blk0, int nblk,
                                                   JacView J) {
  const int nd = M.nd;
  const int ngrp = (nd + CH - 1) / CH;
  int bl, grp;
  if (!xcd_block_map(blockIdx.x, ngrp, nblk, bl, grp)) return;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK || R.np < 2) return;
  const int ipf = B.ray_prof[ir];
  const int im = J.prof_map[ipf];
  if (im < 0 || im >= J.nmap) return;
  const int *xm = J.xmap + (size_t)im * JUR_XMAP_LEN;
  const atm_t *atm = &B.atm[ipf];
  const int id0 = grp * CH;
  const int nc = nd - id0 < CH ? nd - id0 : CH;
  const int npb = E.npb[bl];
  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const double *sc0 = J.scr + J.soff[bl] + (size_t)id0 * npb * J.nscr * WAVE + lane;
  const size_t scstride = (size_t)npb * J.nscr * WAVE;
#define GZ(c, ip) sc0[(c) * scstride + ((size_t)(ip) * J.nscr + J.kz) * WAVE]
#define GD(c, ip) sc0[(c) * scstride + ((size_t)(ip) * J.nscr + J.kz + 1) * WAVE]
  const double *pc = J.pco + J.soff[bl] / ((size_t)nd * J.nscr) * JUR_NPCO + lane;
  const RayTail T = tail[ir];
  const int N = R.np;
  PathRev<CH> P;
  P.xm = xm;
  P.nlev = atm->np;
  P.nc = nc;
  P.kt0 = J.kt + ((size_t)bl * nd + id0) * J.ncol * WAVE + lane;
  P.ktstride = (size_t)J.ncol * WAVE;
  {
    double xl[3], gz[CH], gdl[CH], gdm[CH];
    for (int i = 0; i < 3; i++) xl[i] = LOSR(L.f_xe + i, N - 1);
#pragma unroll
    for (int c = 0; c < CH; c++) {
      gz[c] = c < nc ? GZ(c, N - 1) : 0.0;
      gdl[c] = c < nc ? GD(c, N - 1) : 0.0;
      gdm[c] = c < nc ? GD(c, N - 2) : 0.0;
    }
    P.init(M, T, xl, gz, gdl, gdm);
  }
  auto load_in = [&](int k, PathStepIn<CH> &in) {
    const double *o = pc + (size_t)k * JUR_NPCO * WAVE;
#pragma unroll
    for (int f = 0; f < PC_USED; f++) in.c[f] = o[(size_t)f * WAVE];
#pragma unroll
    for (int c = 0; c < CH; c++) {
      in.gz[c] = c < nc ? GZ(c, k) : 0.0;
      in.gd[c] = c < nc ? GD(c, k) : 0.0;
    }
  };
  PathStepIn<CH> cur, nxt;
  load_in(N - 2, cur);
  if (PF && N >= 3) load_in(N - 3, nxt);
  P.step(N - 2, N, cur.c, cur.gz, cur.gd);
  P.join_clip(cur.c);
  for (int k = N - 3; k >= 0; k--) {
    if (PF) {
      cur = nxt;
      if (k > 0) load_in(k - 1, nxt);
    } else {
      load_in(k, cur);
    }
    P.step(k, N, cur.c, cur.gz, cur.gd);
  }
  P.finish();
#undef GZ
#undef GD
}

// Tile -> row-major K[ray][channel][col], with d BT / d rad (BRIGHT,
// src/jurassic.h:126-127) applied to rows written as brightness temperature
// and NaN rows for masked radiances.  One workgroup per (ray block,
// channel); an LDS transpose of 64 columns at a time makes both the tile
// reads and the row writes contiguous.
__global__ __launch_bounds__(WAVE) void k_jac_store(const double *kt, int nd, int ncol, int blk0,
                                                    int nray, const double *rad, const ChanInfo *chan,
                                                    int write_bbt, const unsigned char *mask, double *K) {
  __shared__ double tile[WAVE][WAVE + 1];
  __shared__ double scale[WAVE];
  const int bl = blockIdx.x / nd, id = blockIdx.x - bl * nd;
  const int lane = threadIdx.x;
  const int ray0 = (blk0 + bl) * WAVE;
  const int ir = ray0 + lane;
  double sc = 1;
  if (ir < nray) {
    const size_t o = (size_t)ir * nd + id;
    if (mask && mask[o]) {
      sc = __builtin_nan("");
    } else if (write_bbt) {
      const double nu = chan[id].nu, r = rad[o];
      const double a = JUR_C1 * (nu * nu * nu);
      const double lg = log1p(a / r);
      sc = JUR_C2 * nu * a / (lg * lg * r * (r + a));
    }
  }
  scale[lane] = sc;
  const double *src = kt + ((size_t)bl * nd + id) * ncol * WAVE;
  const int nr = nray - ray0 < WAVE ? nray - ray0 : WAVE;
  for (int c0 = 0; c0 < ncol; c0 += WAVE) {
    const int nc = ncol - c0 < WAVE ? ncol - c0 : WAVE;
    __syncthreads();
    for (int c = 0; c < nc; c++) tile[c][lane] = src[(size_t)(c0 + c) * WAVE + lane];
    __syncthreads();
    if (lane < nc)
      for (int r = 0; r < nr; r++) K[((size_t)(ray0 + r) * nd + id) * ncol + c0 + lane] = tile[lane][r] * scale[r];
  }
}

constexpr int PATH_CH = 4;  // channels per lane of k_adj_path

template <bool CGA, typename SF, bool PATH>
static void launch_adj(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                       const RayInfo *info, const void *tail, int blk0, int nblk, int maxnp, const JacView &J,
                       double *rad, double *tau, hipStream_t s) {
  const int nb8 = padded_blocks(nblk);
  // small batches (at most two wavefronts per SIMD), FP32 EGA with stored
  // partials and at most four emitters: the emitter chains and the path
  // adjoint run in the accumulation's reverse sweep, one pass instead of
  // three that each last as long as the longest ray's sweep
  bool one = false;
  if constexpr (!CGA && sizeof(SF) == sizeof(float))
    one = E.part && M.ng >= 1 && M.ng <= 4 && (size_t)nb8 * M.nd < 2048;
  if (PATH) {
    // reverse-step coefficients of steps 0 .. np - 2 (lane-major for the
    // one-sweep form)
    const int nchunk = (maxnp - 1 + PCO_STEPS - 1) / PCO_STEPS;
    if (nchunk > 0) {
      if (one)
        hipLaunchKernelGGL(k_path_coef<true>, dim3(nb8 * nchunk), dim3(WAVE), 0, s, M, B, L, info, blk0, nblk,
                           nchunk, J);
      else
        hipLaunchKernelGGL(k_path_coef<false>, dim3(nb8 * nchunk), dim3(WAVE), 0, s, M, B, L, info, blk0, nblk,
                           nchunk, J);
    }
  }
  bool fused = false;
  if constexpr (!CGA && sizeof(SF) == sizeof(float)) {
    if (one) {
      const dim3 grow(nb8 * M.nd), blk(WAVE);
      const RayTail *tl = (const RayTail *)tail;
      if (M.ng <= 2)
        hipLaunchKernelGGL((k_adj_acc<SF, PATH, 2, PATH>), grow, blk, 0, s, M, B, L, E, info, tl, blk0, nblk, J,
                           rad, tau);
      else if (M.ng == 3)
        hipLaunchKernelGGL((k_adj_acc<SF, PATH, 3, PATH>), grow, blk, 0, s, M, B, L, E, info, tl, blk0, nblk, J,
                           rad, tau);
      else
        hipLaunchKernelGGL((k_adj_acc<SF, PATH, 4, PATH>), grow, blk, 0, s, M, B, L, E, info, tl, blk0, nblk, J,
                           rad, tau);
      fused = true;
    }
  }
  if (fused) return;
  hipLaunchKernelGGL((k_adj_acc<SF, PATH>), dim3(nb8 * M.nd), dim3(WAVE), 0, s, M, B, L, E, info,
                     (const RayTail *)tail, blk0, nblk, J, rad, tau);
  if (M.ng > 0) {
    const dim3 grid(nb8 * M.nd * M.ng), blk(WAVE);
    bool fast = false;
    if constexpr (!CGA && sizeof(SF) == sizeof(float)) {
      if (E.part && M.ng <= 4) {
        const dim3 grow(nb8 * M.nd);
        if (M.ng <= 2)
          hipLaunchKernelGGL((k_adj_chains_f32<2, PATH>), grow, blk, 0, s, M, B, L, E, info, blk0, nblk, J);
        else if (M.ng == 3)
          hipLaunchKernelGGL((k_adj_chains_f32<3, PATH>), grow, blk, 0, s, M, B, L, E, info, blk0, nblk, J);
        else
          hipLaunchKernelGGL((k_adj_chains_f32<4, PATH>), grow, blk, 0, s, M, B, L, E, info, blk0, nblk, J);
        fast = true;
      } else if (E.part) {
        hipLaunchKernelGGL((k_adj_chain<false, float, true, PATH>), grid, blk, 0, s, M, B, L, E, info, blk0,
                           nblk, J);
        fast = true;
      }
    }
    if (!fast)
      hipLaunchKernelGGL((k_adj_chain<CGA, SF, false, PATH>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk, J);
  }
  if (PATH) {
    // large batches: PATH_CH channels per lane share each step's
    // coefficients; small ones (under two wavefronts per SIMD that way):
    // one channel per lane, the next step's inputs loaded ahead
    const int ngrp = (M.nd + PATH_CH - 1) / PATH_CH;
    if ((size_t)nb8 * ngrp >= 2048)
      hipLaunchKernelGGL((k_adj_path<PATH_CH, false>), dim3(nb8 * ngrp), dim3(WAVE), 0, s, M, B, L, E, info,
                         (const RayTail *)tail, blk0, nblk, J);
    else
      hipLaunchKernelGGL((k_adj_path<1, true>), dim3(nb8 * M.nd), dim3(WAVE), 0, s, M, B, L, E, info,
                         (const RayTail *)tail, blk0, nblk, J);
  }
}

void launch_adjoint(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                    const RayInfo *info, const void *tail, int blk0, int nblk, int maxnp, const JacView &J,
                    bool fp32, double *rad, double *tau, hipStream_t s) {
  if (M.nd == 0) return;
  const bool path = J.kz >= 0;
#define JUR_ADJ(CGA, SF)                                                                         \
  (path ? launch_adj<CGA, SF, true>(M, B, L, E, info, tail, blk0, nblk, maxnp, J, rad, tau, s)  \
        : launch_adj<CGA, SF, false>(M, B, L, E, info, tail, blk0, nblk, maxnp, J, rad, tau, s))
  if (M.formod == 0) {
    if (fp32) JUR_ADJ(true, float);
    else JUR_ADJ(true, double);
  } else {
    if (fp32) JUR_ADJ(false, float);
    else JUR_ADJ(false, double);
  }
#undef JUR_ADJ
}

void launch_jac_store(const double *kt, int nd, int ncol, int blk0, int nblk, int nray, const double *rad,
                      const ChanInfo *chan, int write_bbt, const unsigned char *mask, double *K,
                      hipStream_t s) {
  if (nd == 0 || ncol == 0 || nblk == 0) return;
  hipLaunchKernelGGL(k_jac_store, dim3(nblk * nd), dim3(WAVE), 0, s, kt, nd, ncol, blk0, nray, rad, chan,
                     write_bbt, mask, K);
}

}  // namespace jur
