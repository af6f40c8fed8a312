// FP64 radiance kernel: combines the per-emitter segment emissivities of the
// lookup kernel with continua and the source function and accumulates
// radiance and transmittance along each line of sight (formod_pencil,
// src/jurassic.c:3233-3334).
//
// One wavefront per (block of 64 rays, channel), lane = ray.  Every segment
// field is read with one coalesced access per wavefront; the work per segment
// is a few FP64 operations, so the kernel streams.
//
// Built with -ffp-contract=off; operation order follows the reference,
// except that the per-segment continua and source function multiply by
// reciprocals (1/T, 1/ds, the grid's inverse spacing, constants folded at
// compile time) where the reference divides, and the H2O continuum's power
// and tanh are evaluated through exp/expm1 (jur_device.hpp, CtmChan):
// radiances within a few ulp of the reference expressions.
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) LOS_AT(los, L, f, ip)

__global__ __launch_bounds__(WAVE) void k_radiance_f64(ModelView M, BatchView B, LosView L,
                                                       EpsView E, const RayInfo *info, int blk0,
                                                       int nblk, double *rad_out, double *tau_out) {
  JUR_FP_EXACT
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
  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const int npb = E.npb[bl];
  const double *eps_g = (const double *)E.eps + E.off[bl] + (size_t)id * ng * npb * WAVE + lane;
  const double *sr_col = M.sr + (size_t)id * TBLNS;
  const double inv_dst = 1 / (M.st[1] - M.st[0]);

  double kcl = 0;
  if (L.f_cl >= 0 && M.ncl > 0)
    kcl = lin(M.clnu[C.icl], atm->clk[C.icl], M.clnu[C.icl + 1], atm->clk[C.icl + 1], nu);
  double sfeps = 1.0;
  if (M.nsf > 0)
    sfeps = lin(M.sfnu[C.isf], atm->sfeps[C.isf], M.sfnu[C.isf + 1], atm->sfeps[C.isf + 1], nu);
  const bool has_k = C.window < M.nw;
  const CtmChan K = ctm_chan(M.ctm, nu);

  double rad = 0, tau = 1, drefl = 0, tau_refl = 1;
  for (int ip = 0; ip < R.np; ip++) {
    const double p = LOSR(F_P, ip);
    const double t = LOSR(F_T, ip);
    const double ds = LOSR(F_DS, ip);

    // tau_seg = prod_g (1 - eps_g), emitter order (intpol_tbl_ega)
    double tau_gas = 1;
    for (int g = 0; g < ng; g++) {
      const double eps = (M.tab.np[g * nd + id] < 30) ? 0.0 : eps_g[((size_t)g * npb + ip) * WAVE];
      tau_gas *= (1 - eps);
    }

    // formod_continua (src/jurassic.c:3075-3117)
    const double rt = 1 / t, rds = 1 / ds;
    double beta = has_k ? LOSR(L.f_k + C.window, ip) : 0.0;
    if (L.f_cl >= 0) beta += LOSR(L.f_cl, ip) * kcl;
    if (M.ig_co2 >= 0) beta += ctmco2_seg(K, p, t, LOSR(L.f_u + M.ig_co2, ip)) * rds;
    if (M.ig_h2o >= 0)
      beta += ctmh2o_seg(K, nu, p, t, rt, LOSR(F_Q0 + M.ig_h2o, ip), LOSR(L.f_u + M.ig_h2o, ip)) * rds;
    if (M.ctm_n2) beta += ctmn2_seg(K, p, t, rt);
    if (M.ctm_o2) beta += ctmo2_seg(K, p, rt);

    const double src = srcfunc_u(M.st, inv_dst, sr_col, t);

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

void launch_radiance_f64(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                         const RayInfo *info, int blk0, int nblk, double *rad, double *tau,
                         hipStream_t s) {
  if (M.nd == 0) return;
  hipLaunchKernelGGL(k_radiance_f64, dim3(padded_blocks(nblk) * M.nd), dim3(WAVE), 0, s, M, B, L, E, info, blk0,
                     nblk, rad, tau);
}

}  // namespace jur
