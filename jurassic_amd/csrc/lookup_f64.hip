// FP64 emissivity lookups (intpol_tbl_ega, src/jurassic.c:3799-3891, and
// intpol_tbl_cga, :3710-3795), one independent chain per (ray, channel,
// emitter).
//
// The EGA recurrence couples the segments of one ray only through each
// emitter's own path transmittance tau_path[id][ig]; emitters and channels
// are independent.  So one wavefront handles one (block of 64 rays, channel,
// emitter), a lane per ray, and writes the segment emissivity eps of every
// LOS point to the emissivity buffer; the radiance kernel combines the
// emitters as the reference does (tau_seg *= 1 - eps in emitter order).
//
// A lane keeps tau_path, its table cell and, per corner, the bracket of the
// inverse lookup u(1 - tau_path).  Along a ray consecutive segments mostly
// stay in the same cell and bracket, so:
//  * cell data (pressure bracket, temperature grids, column sizes) are
//    reloaded only when the cell changes;
//  * each inverse lookup starts from the previous segment's bracket;
//  * the forward lookup eps(u + u_seg) starts from the inverse lookup's
//    bracket record, already in registers (u lies in that bracket and
//    u_seg >= 0).
// Every search returns the reference bisection's index and the arithmetic is
// the reference's (built with -ffp-contract=off).
#include "jur_device.hpp"

namespace jur {

#define LOSR(f, ip) LOS_AT(los, L, f, ip)

// PATH: store the path transmittance tau_path after each segment instead of
// the segment emissivity (the Jacobian restarts any segment from it).
template <bool CGA, bool PATH>
__global__ __launch_bounds__(WAVE) void k_lookup_f64(ModelView M, BatchView B, LosView L,
                                                     EpsView E, const RayInfo *info, int blk0,
                                                     int nblk) {
  JUR_FP_EXACT
  const int nd = M.nd, ng = M.ng;
  int bl, rem;
  if (!xcd_block_map(blockIdx.x, nd * ng, nblk, bl, rem)) return;
  const int id = rem / ng;
  const int ig = rem - id * ng;
  const int lane = threadIdx.x;
  const int ir = (blk0 + bl) * WAVE + lane;
  if (ir >= B.nray) return;
  const TabView &T = M.tab;
  const int tix = ig * nd + id;
  const int npt = T.np[tix];
  if (npt < 30) return;  // eps = 0 for every segment; the consumers know
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  const double *los = L.los + (size_t)bl * L.cap * L.nfield * WAVE + lane;
  const int npb = E.npb[bl];
  double *out = (double *)E.eps + E.off[bl] + ((size_t)id * ng + ig) * npb * WAVE + lane;
  const int mono = T.mono[tix];
  const int pdir = T.pdir[tix];
  const double *pg = T.p + (size_t)tix * T.TP;
  const int *ntp = T.nt + (size_t)tix * T.TP;

  double tp = 1, cu = 0, cp = 0, ct = 0;
  int gi[4];
#pragma unroll
  for (int c = 0; c < 4; c++) gi[c] = T.TU / 2;

  // cached cell
  int ipr = -1, it0 = -1, it1 = -1;
  double pa = 0, pb = 0, g00 = 0, g01 = 0, g10 = 0, g11 = 0;   // p[ipr..], T-grid starts
  double ta0 = 0, tb0 = 0, ta1 = 0, tb1 = 0;
  int nt0 = 0, nt1 = 0, n00 = 0, n01 = 0, n10 = 0, n11 = 0;
  const double *tg0 = nullptr;

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
      // pressure bracket (locate_irr): keep the cached one while it holds
      bool keep = false;
      if (ipr >= 0 && pdir != 0) {
        if (pdir > 0)
          keep = (ipr == 0 || !(pa > pq)) && (ipr == npt - 2 || pb > pq);
        else
          keep = (ipr == 0 || !(pa <= pq)) && (ipr == npt - 2 || pb <= pq);
      }
      if (!keep) {
        ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
        pa = pg[ipr];
        pb = pg[ipr + 1];
        nt0 = ntp[ipr];
        nt1 = ntp[ipr + 1];
        tg0 = T.t + ((size_t)tix * T.TP + ipr) * T.TT;
        g00 = tg0[0];
        g01 = tg0[1];
        g10 = tg0[T.TT];
        g11 = tg0[T.TT + 1];
        it0 = it1 = -1;
      }
      if (nt0 < 2 || nt1 < 2) {
        eps = 0;
      } else {
        // locate_reg on the two temperature rows
        int j0 = (int)((tq - g00) / (g01 - g00));
        j0 = j0 < 0 ? 0 : (j0 > nt0 - 2 ? nt0 - 2 : j0);
        int j1 = (int)((tq - g10) / (g11 - g10));
        j1 = j1 < 0 ? 0 : (j1 > nt1 - 2 ? nt1 - 2 : j1);
        const size_t c0 = (size_t)tix * T.TP + ipr;
        if (j0 != it0) {
          it0 = j0;
          ta0 = tg0[it0];
          tb0 = tg0[it0 + 1];
          n00 = T.nu[c0 * T.TT + it0];
          n01 = T.nu[c0 * T.TT + it0 + 1];
        }
        if (j1 != it1) {
          it1 = j1;
          ta1 = tg0[T.TT + it1];
          tb1 = tg0[T.TT + it1 + 1];
          n10 = T.nu[(c0 + 1) * T.TT + it1];
          n11 = T.nu[(c0 + 1) * T.TT + it1 + 1];
        }
        if (n00 < 2 || n01 < 2 || n10 < 2 || n11 < 2) {
          eps = 0;
        } else {
          const float4 *kc[4];
          kc[0] = T.rec + (c0 * T.TT + it0) * T.TU;
          kc[1] = kc[0] + T.TU;
          kc[2] = T.rec + ((c0 + 1) * T.TT + it1) * T.TU;
          kc[3] = kc[2] + T.TU;
          const int nc[4] = {n00, n01, n10, n11};
          double e[4];
          if (mono) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
              float4 r;
              if (CGA) {
                gi[c] = col_walk<0>(kc[c], nc[c], cu, gi[c], r);
                e[c] = eps_from_rec(r, cu);
              } else {
                const double x = 1 - tp;
                const int i = col_walk<1>(kc[c], nc[c], x, gi[c], r);
                gi[c] = i;
                const double arg = u_from_rec(r, x) + useg;
                float4 r2;
                col_walk_r<0>(kc[c], nc[c], arg, i, r, r2);
                e[c] = eps_from_rec(r2, arg);
              }
            }
          } else {
            int g2 = 0;
#pragma unroll
            for (int c = 0; c < 4; c++)
              e[c] = CGA ? col_eps(kc[c], nc[c], cu, 0, g2)
                         : col_eps(kc[c], nc[c], col_u(kc[c], nc[c], 1 - tp, 0, g2) + useg, 0, g2);
          }
          double e00 = lin(ta0, e[0], tb0, e[1], tq);
          const double e11 = lin(ta1, e[2], tb1, e[3], tq);
          e00 = CGA ? logx(pa, e00, pb, e11, pq) : lin(pa, e00, pb, e11, pq);
          e00 = (e00 < 1) ? e00 : 1;  // MAX(MIN(eps00, 1), 0)
          e00 = (e00 > 0) ? e00 : 0;
          eps = 1 - (1 - e00) / tp;
        }
      }
    }
    tp *= (1 - eps);
    out[(size_t)ip * WAVE] = PATH ? tp : eps;
  }
}

void launch_lookup_f64(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                       const RayInfo *info, int blk0, int nblk, bool path, hipStream_t s) {
  if (M.nd * M.ng == 0) return;
  const dim3 grid(padded_blocks(nblk) * M.nd * M.ng), blk(WAVE);
  if (M.formod == 0) {
    if (path) hipLaunchKernelGGL((k_lookup_f64<true, true>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
    else hipLaunchKernelGGL((k_lookup_f64<true, false>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  } else {
    if (path) hipLaunchKernelGGL((k_lookup_f64<false, true>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
    else hipLaunchKernelGGL((k_lookup_f64<false, false>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  }
}

}  // namespace jur
