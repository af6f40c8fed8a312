// Ray tracing (FP64) and hydrostatic equilibrium on the device.
//
// One lane per ray.  The lane marches its line of sight through its own
// atmosphere exactly as raytrace (src/jurassic.c:4201-4438) does and writes
// the LOS points into the ray block's field-major buffer (see LosView), then
// finds the tangent point (tangent_point, src/jurassic.c:5223-5263) and the
// solar-incidence factor of formod_pencil (src/jurassic.c:3297-3326).
// Built with -ffp-contract=off: every operation rounds as in the reference.
#include "jur_device.hpp"

namespace jur {

// Level search of intpol_atm: the bisection of locate_irr, or, for a
// strictly monotone altitude grid, a walk from the previous point's level
// (the LOS moves by a fraction of a level per step), which returns the same
// index.
// anp is atm->np, read once by the caller: a load of it in the march would
// wait (vmcnt) for the LOS stores issued before it.
__device__ __forceinline__ int atm_level(const atm_t *atm, int anp, int zdir, double z, int &g) {
  g = (zdir != 0 && g >= 0) ? irr_walk(atm->z, anp, z, g, zdir) : locate_irr(atm->z, anp, z);
  return g;
}

// Pressure and temperature interpolation of intpol_atm
// (src/jurassic.c:3685-3706) with the level-pair coefficients cached: LOGY
// evaluates y0 exp(log(y1/y0)/(x1-x0) (x-x0)) and LIN y0 + (y1-y0)/(x1-x0)
// (x-x0); the quotients depend only on the two levels, so keeping them while
// the ray stays between the same levels gives bit-identical values.
struct LevelCache {
  int ip = -1;
  double z0, z1, p0, pq, t0, tq;
  bool plog;
};

__device__ __forceinline__ void intpol_pt(const atm_t *atm, int anp, int zdir, int &g, LevelCache &lc,
                                          double z, double *p, double *t) {
  JUR_FP_EXACT
  // the walk from g returns g when z is still between the cached levels
  bool stay = false;
  if (lc.ip >= 0 && g == lc.ip) {
    if (zdir > 0) stay = (g == 0 || !(lc.z0 > z)) && (g == anp - 2 || lc.z1 > z);
    else if (zdir < 0) stay = (g == 0 || !(lc.z0 <= z)) && (g == anp - 2 || lc.z1 <= z);
  }
  const int ip = stay ? g : atm_level(atm, anp, zdir, z, g);
  if (ip != lc.ip) {
    lc.ip = ip;
    const double x0 = atm->z[ip], x1 = atm->z[ip + 1];
    const double p0 = atm->p[ip], p1 = atm->p[ip + 1];
    lc.z0 = x0;
    lc.z1 = x1;
    lc.p0 = p0;
    lc.plog = (p1 / p0 > 0);
    lc.pq = lc.plog ? log(p1 / p0) / (x1 - x0) : (p1 - p0) / (x1 - x0);
    lc.t0 = atm->t[ip];
    lc.tq = (atm->t[ip + 1] - lc.t0) / (x1 - x0);
  }
  *p = lc.plog ? lc.p0 * exp(lc.pq * (z - lc.z0)) : lc.p0 + lc.pq * (z - lc.z0);
  *t = lc.t0 + lc.tq * (z - lc.z0);
}

// Altitude of a Cartesian point: the z output of cart2geo
// (src/jurassic.c:108-119), for the refraction probes, which use nothing else.
__device__ __forceinline__ double cart2z(const double *x) {
  JUR_FP_EXACT
  return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) - JUR_RE;
}

// sza (src/jurassic.c:5183-5219).
__device__ __forceinline__ double sza(double sec, double lon, double lat) {
  JUR_FP_EXACT
  const double D = sec / 86400 - 0.5;
  const double g = (357.529 + 0.98560028 * D) * (JUR_PI / 180.0);
  const double q = 280.459 + 0.98564736 * D;
  const double L = (q + 1.915 * sin(g) + 0.020 * sin(2 * g)) * (JUR_PI / 180.0);
  const double e = (23.439 - 0.00000036 * D) * (JUR_PI / 180.0);
  const double dec = asin(sin(e) * sin(L));
  const double ra = atan2(cos(e) * sin(L), cos(L));
  const double GMST = 18.697374558 + 24.06570982441908 * D;
  const double LST = GMST + lon / 15;
  const double h = LST / 12 * JUR_PI - ra;
  const double latr = lat * (JUR_PI / 180.0);
  return acos(sin(latr) * sin(dec) + cos(latr) * cos(dec) * cos(h)) * (180.0 / JUR_PI);
}

#define LOSW(f, ip) LOS_AT(base, L, f, ip)
// a LOS field and, for the FP32 forward pass, its float copy (of the
// fields the FP32 radiance kernel reads: the column densities of CO2 and
// H2O and the H2O mixing ratio among the per-gas fields)
// (wr: the lane writes; with several lanes per ray only the first)
#define LOSW3(f, ip, v, copy)                                    \
  do {                                                           \
    const double v_ = (v);                                       \
    if (wr) {                                                    \
      LOSW(f, ip) = v_;                                          \
      if (basef && (copy)) LOS_AT(basef, L, f, ip) = (float)v_;  \
    }                                                            \
  } while (0)
#define LOSW2(f, ip, v) LOSW3(f, ip, v, true)

// RL lanes per ray.  RL = 4 (small batches, where a wavefront's march is the
// kernel's duration): the four lanes of a ray march it together and each
// evaluates one of the four refraction probes of a step, which the lanes
// then exchange; every lane does the same operations on the same values, so
// the LOS is the same to the last bit, and the first lane writes it.  A
// wavefront then covers a quarter of a 64-ray block.
template <int RL>
__global__ __launch_bounds__(WAVE) void k_raytrace(ModelView M, BatchView B, LosView L,
                                                   RayInfo *info, RayTail *tail, int blk0) {
  JUR_FP_EXACT
  const int lane = threadIdx.x;
  const int lb = blockIdx.x / RL;  // ray block of the chunk
  const int rib = RL == 1 ? lane : (blockIdx.x % RL) * (WAVE / RL) + lane / RL;  // ray in the block
  const int sub = RL == 1 ? 0 : lane % RL;
  const bool wr = sub == 0;
  const int blk = blk0 + lb;
  const int ir = blk * WAVE + rib;
  if (ir >= B.nray) return;
  double *base = L.los + (size_t)lb * L.cap * L.nfield * WAVE + rib;
  float *basef = L.losf ? L.losf + (size_t)lb * L.cap * L.nfield * WAVE + rib : nullptr;

  const double *g = B.ray_geom + (size_t)7 * ir;
  const int iprof = B.ray_prof[ir];
  if (iprof < 0 || iprof >= B.nprof || B.atm[iprof].np < 2 || B.atm[iprof].np > NP) {
    RayInfo E{};
    E.status = JUR_ERR_ARG;
    if (wr) info[ir] = E;
    return;
  }
  const atm_t *atm = &B.atm[iprof];
  const double obsz = g[1], obslon = g[2], obslat = g[3];
  const double vpz = g[4], vplon = g[5], vplat = g[6];

  const double h = 0.02, zrefrac = 60;
  double ex0[3], ex1[3], n, ngr[3], norm, p, t, x[3], xh[3], xobs[3], xvp[3];
  double z = 1e99, zmax, zmin;
  int np = 0;

  RayInfo R;
  R.np = 0;
  R.status = JUR_OK;
  R.sft = -999;
  R.tpz = vpz;
  R.tplon = vplon;
  R.tplat = vplat;
  R.rcos = 0;
  R.pad = 0;

  // altitude range (gsl_stats_minmax semantics, src/jurassic.c:4223) and
  // monotonicity of the altitude grid
  zmin = zmax = atm->z[0];
  int zasc = 1, zdesc = 1;
  const int anp = atm->np;
  for (int i = 0; i < anp; i++) {
    const double v = atm->z[i];
    if (v < zmin) zmin = v;
    if (v > zmax) zmax = v;
    if (isnan(v)) { zmin = zmax = v; zasc = zdesc = 0; break; }
    if (i > 0) {
      if (!(v > atm->z[i - 1])) zasc = 0;
      if (!(v < atm->z[i - 1])) zdesc = 0;
    }
  }
  const int zdir = zasc ? 1 : (zdesc ? -1 : 0);
  int zg = -1;  // level of the previous LOS point
  LevelCache lc;
  // level pair of intpol_atm: [0] z0 [1] z1 [2] p0 [3] p quotient [4] t0
  // [5] t slope [6..] q0[ng], q slopes[ng], k0[nw], k slopes[nw]; field-major,
  // lane fastest
  extern __shared__ double lvl[];
#define LV(f) lvl[(f) * WAVE + lane]
  int lip = -1;
  bool lplog = false;
  if (M.nsf > 0) {
    zmin = (atm->sfz > zmin) ? atm->sfz : zmin;
    if (atm->sfp > 0) {
      const int ip = locate_irr(atm->p, atm->np, atm->sfp);
      const double zip = lin(log(atm->p[ip]), atm->z[ip], log(atm->p[ip + 1]), atm->z[ip + 1],
                             log(atm->sfp));
      zmin = (zip > zmin) ? zip : zmin;
    }
  }

  if (obsz < zmin) {
    R.status = JUR_ERR_OBSERVER;
    if (wr) info[ir] = R;
    return;
  }
  if (vpz > zmax) {
    if (wr) info[ir] = R;
    return;
  }

  geo2cart(obsz, obslon, obslat, xobs);
  geo2cart(vpz, vplon, vplat, xvp);
  for (int i = 0; i < 3; i++) ex0[i] = xvp[i] - xobs[i];
  norm = sqrt(ex0[0] * ex0[0] + ex0[1] * ex0[1] + ex0[2] * ex0[2]);
  for (int i = 0; i < 3; i++) ex0[i] /= norm;
  for (int i = 0; i < 3; i++) x[i] = xobs[i];

  // entry point of an observer above the atmosphere (bisection to 1 m)
  if (obsz > zmax) {
    double dmax = norm, dmin = 0;
    while (fabs(dmin - dmax) > 0.001) {
      const double d = (dmax + dmin) / 2;
      for (int i = 0; i < 3; i++) x[i] = xobs[i] + d * ex0[i];
      z = cart2z(x);
      if (z <= zmax && z > zmax - 0.001) break;
      if (z < zmax - 0.0005) dmax = d; else dmin = d;
    }
  }

  // Cartesian positions of the points whose longitude and latitude are
  // needed after the march: the first and the last point, and the two
  // neighbours of the first minimum of z (tangent_point); cart2geo of the
  // same x gives the values the reference stores per point.
  // The positions go straight to the ray's tail (they change rarely), the
  // rest of MinTrack's state stays in registers.
  RayTail *tl = tail + ir;
  double zprev = 0, xprev[3] = {0, 0, 0};
  double zlow = 0;
  int imin = 0;
  bool after_pending = false, nan_seen = false;
  double dscut = 0;
  double ds1 = 0, ds2 = 0;  // raw step lengths of the last two points written
  for (;;) {
    double ds = M.rayds;
    if (M.raydz > 0) {
      norm = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      const double rn = 1 / norm;
      for (int i = 0; i < 3; i++) xh[i] = x[i] * rn;
      const double cosa = fabs(ex0[0] * xh[0] + ex0[1] * xh[1] + ex0[2] * xh[2]);
      if (cosa != 0) ds = (M.rayds < M.raydz / cosa) ? M.rayds : M.raydz / cosa;
    }

    z = cart2z(x);

    if (z < zmin || z > zmax) {
      // The step left the altitude range: k_ray_finish clips it to the
      // range and writes the last point (its cart2geo/geo2cart would hold
      // registers for the whole march here).  With no point written yet
      // the entry search ended outside the range, so there is no previous
      // point to clip to (the reference would read los->z[-1]): the ray has
      // no LOS points.
      dscut = ds;
      break;
    }

    // intpol_atm at the LOS point.  The level pair and the quotients of
    // LOGY (p) and LIN (t, q, k) are kept in LDS while the point stays
    // between the same two levels (the level walk would return the same
    // index), so that a step loads nothing from the atmosphere.
    {
      bool stay = false;
      if (lip >= 0) {
        const double za = LV(0), zb = LV(1);
        if (zdir > 0) stay = (lip == 0 || !(za > z)) && (lip == anp - 2 || zb > z);
        else if (zdir < 0) stay = (lip == 0 || !(za <= z)) && (lip == anp - 2 || zb <= z);
      }
      if (!stay) {
        const int ipa = atm_level(atm, anp, zdir, z, zg);
        if (ipa != lip) {
          lip = ipa;
          const double x0 = atm->z[ipa], x1 = atm->z[ipa + 1];
          const double y0 = atm->p[ipa], y1 = atm->p[ipa + 1];
          LV(0) = x0;
          LV(1) = x1;
          lplog = (y1 / y0 > 0);
          LV(2) = y0;
          LV(3) = lplog ? log(y1 / y0) / (x1 - x0) : (y1 - y0) / (x1 - x0);
          LV(4) = atm->t[ipa];
          LV(5) = (atm->t[ipa + 1] - atm->t[ipa]) / (x1 - x0);
          for (int ig = 0; ig < M.ng; ig++) {
            LV(6 + ig) = atm->q[ig][ipa];
            LV(6 + M.ng + ig) = (atm->q[ig][ipa + 1] - atm->q[ig][ipa]) / (x1 - x0);
          }
          for (int iw = 0; iw < M.nw; iw++) {
            LV(6 + 2 * M.ng + iw) = atm->k[iw][ipa];
            LV(6 + 2 * M.ng + M.nw + iw) = (atm->k[iw][ipa + 1] - atm->k[iw][ipa]) / (x1 - x0);
          }
        }
      }
    }
    const double dzl = z - LV(0);
    p = lplog ? LV(2) * exp(LV(3) * dzl) : LV(2) + LV(3) * dzl;
    t = LV(4) + LV(5) * dzl;

    if (np >= L.cap) {
      // capacity exceeded: the reference limit NLOS -> error, a smaller
      // working capacity -> the host retries with NLOS
      R.status = (L.cap >= NLOS) ? JUR_ERR_LOS : 100;
      if (wr) info[ir] = R;
      return;
    }
    // the trapezoid segment length of the point and its column densities
    // (src/jurassic.c:4410-4419) from its raw step length ds and the
    // previous point's: 0.5 (ds_prev + ds), the first point 0.5 ds (the
    // clip of the last step changes the last two points, which k_ray_finish
    // rewrites)
    const double dtr = np == 0 ? ds * 0.5 : 0.5 * (ds1 + ds);
    LOSW3(F_Z, np, z, false);
    LOSW2(F_P, np, p);
    LOSW2(F_T, np, t);
    const double ucol = 10 * p / (JUR_KB * t) * dtr;  // column density per unit vmr
    for (int ig = 0; ig < M.ng; ig++) {
      const double q = LV(6 + ig) + LV(6 + M.ng + ig) * dzl;
      LOSW3(F_Q0 + ig, np, q, ig == M.ig_h2o);
      LOSW3(L.f_u + ig, np, q * ucol, ig == M.ig_co2 || ig == M.ig_h2o);
    }
    for (int iw = 0; iw < M.nw; iw++)
      LOSW2(L.f_k + iw, np, LV(6 + 2 * M.ng + iw) + LV(6 + 2 * M.ng + M.nw + iw) * dzl);
    LOSW2(F_DS, np, dtr);
    if (L.f_cl >= 0) {
      const double aux = (M.ncl > 0 && atm->cldz > 0)
                             ? exp(-0.5 * (((z - atm->clz) / atm->cldz) * ((z - atm->clz) / atm->cldz)))
                             : 0.0;
      LOSW2(L.f_cl, np, aux);
    }
    if (L.f_xe >= 0)
      for (int i = 0; i < 3; i++) {
        LOSW3(L.f_xe + i, np, x[i], false);
        LOSW3(L.f_xe + 3 + i, np, ex0[i], false);
      }
    // MinTrack::add, with the raw step lengths at the minimum and after it
    if (!nan_seen) {
      const bool lower = (np == 0) || z < zlow;
      if (lower) zlow = z;
      if (lower || isnan(z)) {
        imin = np;
        if (wr) {
          tl->ds_imin = ds;
          if (np > 0)
            for (int i = 0; i < 3; i++) tl->mt.xbefore[i] = xprev[i];
        }
        after_pending = true;
        nan_seen = isnan(z);
      } else if (after_pending) {
        if (wr) {
          for (int i = 0; i < 3; i++) tl->mt.xafter[i] = x[i];
          tl->ds_imin1 = ds;
        }
        after_pending = false;
      }
    } else if (after_pending) {
      if (wr) {
        for (int i = 0; i < 3; i++) tl->mt.xafter[i] = x[i];
        tl->ds_imin1 = ds;
      }
      after_pending = false;
    }
    ds2 = ds1;
    ds1 = ds;
    if (np == 0 && wr)
      for (int i = 0; i < 3; i++) tl->xfirst[i] = x[i];
    for (int i = 0; i < 3; i++) xprev[i] = x[i];
    zprev = z;
    np++;

    if (M.refrac && z <= zrefrac) n = 1 + 7.753e-05 * p / t;
    else n = 1;

    for (int i = 0; i < 3; i++) ex1[i] = ex0[i] * n;

    if (RL == 4 && M.refrac && z <= zrefrac) {
      // lane sub's probe: the midpoint (0) or the midpoint displaced by h
      // along axis sub - 1, after the displacements of the axes before it
      // were made and undone as in the serial order below
      for (int i = 0; i < 3; i++) xh[i] = x[i] + 0.5 * ds * ex0[i];
      for (int i = 0; i + 1 < sub; i++) {
        xh[i] += h;
        xh[i] -= h;
      }
      if (sub > 0) xh[sub - 1] += h;
      double pq, tq;
      int zg2 = zg;
      intpol_pt(atm, anp, zdir, zg2, lc, cart2z(xh), &pq, &tq);
      const double rq = 7.753e-05 * pq / tq;
      n = __shfl(rq, 0, RL);
      for (int i = 0; i < 3; i++) ngr[i] = (__shfl(rq, i + 1, RL) - n) * (1 / h);
    } else if (M.refrac && z <= zrefrac) {
      // the refractivity at the midpoint and at the three probes displaced
      // by h along each axis, in the reference's order (each displacement
      // is undone by subtracting h before the next axis moves)
      double zq[4], pq[4], tq[4];
      for (int i = 0; i < 3; i++) xh[i] = x[i] + 0.5 * ds * ex0[i];
      zq[0] = cart2z(xh);
      for (int i = 0; i < 3; i++) {
        xh[i] += h;
        zq[i + 1] = cart2z(xh);
        xh[i] -= h;
      }
      int zg2 = zg;
      // all four between the probes' cached levels (intpol_pt's walk would
      // return them for each): the cached formula for each, side by side;
      // otherwise intpol_pt probe by probe (not unrolled)
      bool all = lc.ip >= 0 && zg2 == lc.ip && zdir != 0;
      for (int k = 0; k < 4; k++) {
        if (zdir > 0) all = all && (zg2 == 0 || !(lc.z0 > zq[k])) && (zg2 == anp - 2 || lc.z1 > zq[k]);
        else all = all && (zg2 == 0 || !(lc.z0 <= zq[k])) && (zg2 == anp - 2 || lc.z1 <= zq[k]);
      }
      if (all) {
        for (int k = 0; k < 4; k++) {
          pq[k] = lc.plog ? lc.p0 * exp(lc.pq * (zq[k] - lc.z0)) : lc.p0 + lc.pq * (zq[k] - lc.z0);
          tq[k] = lc.t0 + lc.tq * (zq[k] - lc.z0);
        }
      } else {
#pragma unroll 1
        for (int k = 0; k < 4; k++) intpol_pt(atm, anp, zdir, zg2, lc, zq[k], &pq[k], &tq[k]);
      }
      n = 7.753e-05 * pq[0] / tq[0];
      for (int i = 0; i < 3; i++) ngr[i] = (7.753e-05 * pq[i + 1] / tq[i + 1] - n) * (1 / h);
    } else
      for (int i = 0; i < 3; i++) ngr[i] = 0;

    for (int i = 0; i < 3; i++) ex1[i] += ds * ngr[i];
    norm = sqrt(ex1[0] * ex1[0] + ex1[1] * ex1[1] + ex1[2] * ex1[2]);
    const double rn1 = 1 / norm;
    for (int i = 0; i < 3; i++) ex1[i] *= rn1;
    for (int i = 0; i < 3; i++) x[i] += 0.5 * ds * (ex0[i] + ex1[i]);
    for (int i = 0; i < 3; i++) ex0[i] = ex1[i];
  }
  R.np = np;
  if (!wr) return;
  if (np > 0) {
    for (int i = 0; i < 3; i++) {
      tl->xprev[i] = xprev[i];
      tl->xcut[i] = x[i];
      tl->ecut[i] = ex0[i];
    }
    tl->zprev = zprev;
    tl->zcut = z;
    tl->dscut = dscut;
    tl->zmin = zmin;
    tl->zmax = zmax;
    tl->ds1 = ds1;
    tl->ds2 = ds2;
    tl->mt.zlow = zlow;
    tl->mt.imin = imin;
    tl->mt.after_pending = after_pending;
    tl->mt.nan_seen = nan_seen;
    tl->zdir = zdir;
  }
  info[ir] = R;
}

// The end of a ray's march (k_raytrace): the step that left the altitude
// range is clipped to it and becomes the last LOS point, as raytrace does
// (src/jurassic.c:4290-4306), then the tangent point, the solar-incidence
// factor and the trapezoid step lengths and column densities.  Same
// operations as in the march, so the same bits.
__global__ __launch_bounds__(WAVE) void k_ray_finish(ModelView M, BatchView B, LosView L, RayInfo *info,
                                                     const RayTail *tail, int blk0) {
  JUR_FP_EXACT
  const int lane = threadIdx.x;
  const int ir = (blk0 + blockIdx.x) * WAVE + lane;
  if (ir >= B.nray) return;
  constexpr bool wr = true;
  RayInfo R = info[ir];
  if (R.status != JUR_OK || R.np == 0) return;
  double *base = L.los + (size_t)blockIdx.x * L.cap * L.nfield * WAVE + lane;
  float *basef = L.losf ? L.losf + (size_t)blockIdx.x * L.cap * L.nfield * WAVE + lane : nullptr;
  const atm_t *atm = &B.atm[B.ray_prof[ir]];
  const RayTail T = tail[ir];
  MinTrack mt = T.mt;
  int np = R.np;

  // the last step, clipped to the altitude range
  const int stop = (T.zcut < T.zmin ? 2 : 1);
  const double frac = ((T.zcut < T.zmin ? T.zmin : T.zmax) - T.zprev) / (T.zcut - T.zprev);
  double zd, lon, lat, xh[3], x[3];
  cart2geo(T.xprev, &zd, &lon, &lat);
  geo2cart(T.zprev, lon, lat, xh);
  for (int i = 0; i < 3; i++) x[i] = xh[i] + frac * (T.xcut[i] - xh[i]);
  const double z = cart2z(x);
  // los->ds[np-1] = ds * frac: the raw step length of the last point written
  const int npm = np;
  const double clip = T.dscut * frac;

  // intpol_atm at the clipped point with the march's level pair formulas
  int g = -1;
  const int ipa = atm_level(atm, atm->np, T.zdir, z, g);
  const double x0 = atm->z[ipa], x1 = atm->z[ipa + 1];
  const double y0 = atm->p[ipa], y1 = atm->p[ipa + 1];
  const bool lplog = (y1 / y0 > 0);
  const double pq = lplog ? log(y1 / y0) / (x1 - x0) : (y1 - y0) / (x1 - x0);
  const double tq = (atm->t[ipa + 1] - atm->t[ipa]) / (x1 - x0);
  const double dzl = z - x0;
  const double p = lplog ? y0 * exp(pq * dzl) : y0 + pq * dzl;
  double t = atm->t[ipa] + tq * dzl;
  if (np >= L.cap) {
    R.status = (L.cap >= NLOS) ? JUR_ERR_LOS : 100;
    info[ir] = R;
    return;
  }
  // trapezoid segment lengths (src/jurassic.c:4410-4411) of the last point
  // written, from its clipped raw length, and of this one (raw length 0);
  // the march wrote those of the others
  const double dtr0 = npm == 1 ? clip * 0.5 : 0.5 * (T.ds2 + clip);
  const double dtr1 = 0.5 * (clip + 0.0);
  LOSW2(F_DS, npm - 1, dtr0);
  {
    const double p0 = LOSW(F_P, npm - 1), t0 = LOSW(F_T, npm - 1);
    for (int ig = 0; ig < M.ng; ig++)
      LOSW3(L.f_u + ig, npm - 1, 10 * LOSW(F_Q0 + ig, npm - 1) * p0 / (JUR_KB * t0) * dtr0,
            ig == M.ig_co2 || ig == M.ig_h2o);
  }
  LOSW(F_Z, np) = z;
  LOSW2(F_P, np, p);
  LOSW2(F_T, np, t);
  for (int ig = 0; ig < M.ng; ig++) {
    const double q0 = atm->q[ig][ipa];
    const double qs = (atm->q[ig][ipa + 1] - atm->q[ig][ipa]) / (x1 - x0);
    const double q = q0 + qs * dzl;
    LOSW3(F_Q0 + ig, np, q, ig == M.ig_h2o);
    LOSW3(L.f_u + ig, np, 10 * q * p / (JUR_KB * t) * dtr1, ig == M.ig_co2 || ig == M.ig_h2o);
  }
  for (int iw = 0; iw < M.nw; iw++) {
    const double k0 = atm->k[iw][ipa];
    const double ks = (atm->k[iw][ipa + 1] - atm->k[iw][ipa]) / (x1 - x0);
    LOSW2(L.f_k + iw, np, k0 + ks * dzl);
  }
  LOSW2(F_DS, np, dtr1);
  if (L.f_cl >= 0) {
    const double aux = (M.ncl > 0 && atm->cldz > 0)
                           ? exp(-0.5 * (((z - atm->clz) / atm->cldz) * ((z - atm->clz) / atm->cldz)))
                           : 0.0;
    LOSW2(L.f_cl, np, aux);
  }
  if (L.f_xe >= 0)
    for (int i = 0; i < 3; i++) {
      LOSW(L.f_xe + i, np) = x[i];
      LOSW(L.f_xe + 3 + i, np) = 0.0;
    }
  mt.add(np, z, x, T.xprev);
  np++;
  if (M.nsf > 0 && atm->sft > 0) t = atm->sft;
  R.sft = (stop == 2 ? t : -999);
  R.np = np;

  // tangent_point at the first minimum imin found during the march
  double zl, lonl, latl;
  cart2geo(x, &zl, &lonl, &latl);
  if (mt.imin <= 0 || mt.imin >= np - 1) {
    R.tpz = LOSW(F_Z, np - 1);
    R.tplon = lonl;
    R.tplat = latl;
  } else {
    const double yy0 = LOSW(F_Z, mt.imin - 1);
    const double yy1 = LOSW(F_Z, mt.imin);
    const double yy2 = LOSW(F_Z, mt.imin + 1);
    // the raw step lengths at imin and imin + 1 (tangent_point runs
    // before the trapezoid rule in raytrace)
    auto rawds = [&](int i) {
      return i == npm ? 0.0 : (i == npm - 1 ? clip : (i == mt.imin ? T.ds_imin : T.ds_imin1));
    };
    const double d1 = rawds(mt.imin), d2 = rawds(mt.imin + 1);
    const double x1 = sqrt(d1 * d1 - (yy1 - yy0) * (yy1 - yy0));
    const double x2 = x1 + sqrt(d2 * d2 - (yy2 - yy1) * (yy2 - yy1));
    const double a = 1 / (x1 - x2) * (-(yy0 - yy1) / x1 + (yy0 - yy2) / x2);
    const double b = -(yy0 - yy1) / x1 - a * x1;
    const double c = yy0;
    const double xt = -b / (2 * a);
    double v[3], v0[3], v2[3], dummy;
    R.tpz = a * xt * xt + b * xt + c;
    double lon0, lat0, lon2, lat2;
    cart2geo(mt.xbefore, &dummy, &lon0, &lat0);
    cart2geo(mt.xafter, &dummy, &lon2, &lat2);
    geo2cart(yy0, lon0, lat0, v0);
    geo2cart(yy2, lon2, lat2, v2);
    for (int i = 0; i < 3; i++) v[i] = lin(0.0, v0[i], x2, v2[i], xt);
    cart2geo(v, &dummy, &R.tplon, &R.tplat);
  }

  // solar incidence (formod_pencil, src/jurassic.c:3297-3326)
  if (M.sftype >= 3 && R.sft > 0) {
    const double sza2 = (M.sfsza < 0) ? sza(B.ray_geom[(size_t)7 * ir], lonl, latl) : M.sfsza;
    if (sza2 < 89.999) {
      double x0[3], x1[3];
      geo2cart(LOSW(F_Z, np - 1), lonl, latl, x0);
      double lonf, latf, zf;
      cart2geo(T.xfirst, &zf, &lonf, &latf);
      geo2cart(LOSW(F_Z, 0), lonf, latf, x1);
      for (int i = 0; i < 3; i++) x1[i] -= x0[i];
      const double cosa = (x0[0] * x1[0] + x0[1] * x1[1] + x0[2] * x1[2]) /
                          sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]) /
                          sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
      R.rcos = cosa / cos(sza2 * (JUR_PI / 180.0));
      R.pad = 1.0;
    }
  }

  info[ir] = R;
}

// hydrostatic (src/jurassic.c:3520-3582), one lane per atmosphere.
__global__ __launch_bounds__(WAVE) void k_hydrostatic(atm_t *atm, int nprof, double hydz,
                                                      int ig_h2o) {
  JUR_FP_EXACT
  const int ia = blockIdx.x * WAVE + threadIdx.x;
  if (ia >= nprof) return;
  atm_t *a = &atm[ia];
  const double mmair = 28.96456e-3, mmh2o = 18.0153e-3, G0 = 9.80665, RI = 8.3144598;
  const int ipts = 20;
  double dzmin = 1e99, e = 0;
  int ipref = 0;
  for (int ip = 0; ip < a->np; ip++)
    if (fabs(a->z[ip] - hydz) < dzmin) {
      dzmin = fabs(a->z[ip] - hydz);
      ipref = ip;
    }
  for (int ip = ipref + 1; ip < a->np; ip++) {
    double mean = 0;
    for (int i = 0; i < ipts; i++) {
      if (ig_h2o >= 0) e = lin(0.0, a->q[ig_h2o][ip - 1], ipts - 1.0, a->q[ig_h2o][ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * G0 / RI /
              lin(0.0, a->t[ip - 1], ipts - 1.0, a->t[ip], (double)i) / ipts;
    }
    a->p[ip] = exp(log(a->p[ip - 1]) - mean * 1000 * (a->z[ip] - a->z[ip - 1]));
  }
  for (int ip = ipref - 1; ip >= 0; ip--) {
    double mean = 0;
    for (int i = 0; i < ipts; i++) {
      if (ig_h2o >= 0) e = lin(0.0, a->q[ig_h2o][ip + 1], ipts - 1.0, a->q[ig_h2o][ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * G0 / RI /
              lin(0.0, a->t[ip + 1], ipts - 1.0, a->t[ip], (double)i) / ipts;
    }
    a->p[ip] = exp(log(a->p[ip + 1]) - mean * 1000 * (a->z[ip] - a->z[ip + 1]));
  }
}

size_t ray_tail_bytes() { return sizeof(RayTail); }

// tail: ray_tail_bytes() per ray of the batch
void launch_raytrace(const ModelView &M, const BatchView &B, const LosView &L, RayInfo *info, void *tail,
                     int blk0, int nblk, hipStream_t s) {
  const size_t lds = (size_t)(6 + 2 * M.ng + 2 * M.nw) * WAVE * sizeof(double);
  // small batches: four lanes per ray, a quarter block per wavefront
  if (nblk <= 512)
    hipLaunchKernelGGL(k_raytrace<4>, dim3(4 * nblk), dim3(WAVE), lds, s, M, B, L, info, (RayTail *)tail, blk0);
  else
    hipLaunchKernelGGL(k_raytrace<1>, dim3(nblk), dim3(WAVE), lds, s, M, B, L, info, (RayTail *)tail, blk0);
  hipLaunchKernelGGL(k_ray_finish, dim3(nblk), dim3(WAVE), 0, s, M, B, L, info, (const RayTail *)tail, blk0);
}

void launch_hydrostatic(atm_t *atm, int nprof, double hydz, int ig_h2o, hipStream_t s) {
  hipLaunchKernelGGL(k_hydrostatic, dim3((nprof + WAVE - 1) / WAVE), dim3(WAVE), 0, s, atm, nprof,
                     hydz, ig_h2o);
}

}  // namespace jur
