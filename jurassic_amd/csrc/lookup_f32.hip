// FP32 emissivity lookups: the chains of lookup_f64.hip with FP32 table
// interpolation.
//
// Precision plan (budget: 0.01 K brightness temperature against FP64):
//  * The state of a chain is the path emissivity eps_path = 1 - tau_path.
//    The reference's update tau_path *= 1 - eps with eps = 1 - (1 - eps00) /
//    tau_path makes tau_path exactly 1 - eps00, so eps_path = eps00 carries
//    no accumulated rounding, and small emissivities keep full relative
//    precision in float where 1 - tau_path would not.
//  * The emissivity-growth step u_eq = u(eps_path), eps00 = eps(u_eq + u_seg)
//    is evaluated relative to the knot u_i of the inverse lookup's bracket:
//    off = (eps_path - eps_i) du/deps + u_seg is a small float with no
//    cancellation against u_i, and eps00 = eps_i + off deps/du while off stays
//    inside the bracket (the common case; otherwise the forward bracket is
//    searched on the absolute column density in FP64).
//  * Output is the path emissivity eps_path after each segment (the radiance
//    kernel forms the segment emissivity (eps_i - eps_{i-1}) / (1 - eps_{i-1});
//    the adjoint restarts any segment from eps_{i-1}); 1 marks an opaque
//    path (tau_path below float resolution).
//  * For the FP32 EGA Jacobian over monotone tables (PART), each step also
//    stores its partials d eps_path / d (eps_prev, p, T, u_seg) with the
//    slopes of the brackets it used (see adjoint.hip), so that the reverse
//    chains need no table access.
#include "jur_device.hpp"

namespace jur {

// LOS field f of point ip: wave-uniform block base, 32-bit lane offset
#define LOSR(f, ip) ld32(losb, ((unsigned)(ip) * nfield + (unsigned)(f)) * WAVE + ln)

// RL lanes per chain.  RL = 4 (small batches, where a wavefront's chain is
// the kernel's duration; EGA): the four lanes of a (ray, channel, emitter)
// chain each take one corner of the cell's step and exchange the corners'
// results (and bracket indices) by shuffles; every lane then does the same
// operations on the same values as one lane would, so the emissivities are
// the same to the last bit, and the first lane writes them.  A wavefront
// covers a quarter of a 64-ray block.
template <bool CGA, bool PART, int RL = 1>
__global__ __launch_bounds__(WAVE) void k_lookup_f32(ModelView M, BatchView B, LosView L,
                                                     EpsView E, const RayInfo *info, int blk0,
                                                     int nblk) {
  static_assert(RL == 1 || !CGA, "the four-lane chains are the EGA ones");
  const int nd = M.nd, ng = M.ng;
  int bl, rem;
  if (!xcd_block_map(blockIdx.x, nd * ng * RL, nblk, bl, rem)) return;
  const int qtr = rem % RL;
  rem /= RL;
  const int id = rem / ng;
  const int ig = rem - id * ng;
  const int lane = threadIdx.x;
  const int sub = lane % RL;  // the lane's corner (RL = 4)
  const int ln = RL == 1 ? lane : qtr * (WAVE / RL) + lane / RL;  // the ray's lane in its block
  const bool wr = sub == 0;
  const int ir = (blk0 + bl) * WAVE + ln;
  if (ir >= B.nray) return;
  const TabView &T = M.tab;
  const int tix = ig * nd + id;
  const int npt = T.np[tix];
  if (npt < 30) return;
  const RayInfo R = info[ir];
  if (R.status != JUR_OK) return;

  // wave-uniform bases (the ray block's LOS and emissivity chain, the
  // table's arrays) with 32-bit per-lane offsets
  const double *losb = L.los + (size_t)bl * L.cap * L.nfield * WAVE;
  const unsigned nfield = L.nfield;
  const int npb = E.npb[bl];
  float *outb = (float *)E.eps + E.off[bl] + ((size_t)id * ng + ig) * npb * WAVE;
  float4 *partb = PART ? E.part + E.off[bl] + ((size_t)id * ng + ig) * npb * WAVE : nullptr;
  const int mono = T.mono[tix];
  const int pdir = T.pdir[tix];
  const double *pg = T.p + (size_t)tix * T.TP;
  const int *ntp = T.nt + (size_t)tix * T.TP;
  const size_t tcols = (size_t)T.TP * T.TT;  // (p, T) columns per table
  const double *tgb = T.t + (size_t)tix * tcols;
  const int *nub = T.nu + (size_t)tix * tcols;
  const float4 *recb = T.rec + (size_t)tix * tcols * T.TU;

  float ep = 0.f;
  double cu = 0, cp = 0, ct = 0;
  int gi[4];
#pragma unroll
  for (int c = 0; c < 4; c++) gi[c] = T.TU / 2;

  // cached cell
  int ipr = -1, it0 = -1, it1 = -1;
  double pa = 0, pb = 0;
  float wp_scale = 0, g00 = 0, gi0 = 0, g10 = 0, gi1 = 0;
  float ta0 = 0, dt0 = 0, ta1 = 0, dt1 = 0, lpr = 0;
  int nt0 = 0, nt1 = 0, n00 = 0, n01 = 0, n10 = 0, n11 = 0;
  unsigned tq0 = 0;  // first T-grid entry of p row ipr

  // LOS values of the current point; the next point's are loaded after the
  // corner records of the current one, so that waiting for the records does
  // not wait for the (HBM) LOS loads as well
  double p = 0, t = 0, useg = 0;
  if (R.np > 0) {
    p = LOSR(F_P, 0);
    t = LOSR(F_T, 0);
    useg = LOSR(L.f_u + ig, 0);
  }
  int ip = 0;
  for (; ip < R.np; ip++) {
    const int ipn = ip + 1 < R.np ? ip + 1 : ip;
    if (CGA) {
      cu += useg;
      cp += useg * p;
      ct += useg * t;
    }
    const double pq = CGA ? cp / cu : p;
    const float tq = (float)(CGA ? ct / cu : t);
    bool keep = false;
    if (ipr >= 0 && pdir != 0) {
      if (pdir > 0)
        keep = (ipr == 0 || !(pa > pq)) && (ipr == npt - 2 || pb > pq);
      else
        keep = (ipr == 0 || !(pa <= pq)) && (ipr == npt - 2 || pb <= pq);
    }
    if (!keep) {
      const int iprev = ipr;
      ipr = (pdir != 0 && ipr >= 0) ? irr_walk(pg, npt, pq, ipr, pdir) : locate_irr(pg, npt, pq);
      // start indices for the new corner columns: a column the cell keeps
      // takes its bracket along, a new one extrapolates from the pair
      if (ipr == iprev + 1) {
        const int g0 = gi[0], g1 = gi[1];
        gi[0] = gi[2];
        gi[1] = gi[3];
        gi[2] = 2 * gi[2] - g0;
        gi[3] = 2 * gi[3] - g1;
      } else if (ipr == iprev - 1) {
        const int g2 = gi[2], g3 = gi[3];
        gi[2] = gi[0];
        gi[3] = gi[1];
        gi[0] = 2 * gi[0] - g2;
        gi[1] = 2 * gi[1] - g3;
      }
      pa = pg[ipr];
      pb = pg[ipr + 1];
      wp_scale = frcp((float)(pb - pa));
      if (CGA) lpr = (float)(1.0 / log(pb / pa));
      nt0 = ntp[ipr];
      nt1 = ntp[ipr + 1];
      tq0 = (unsigned)ipr * T.TT;
      const double t00 = ld32(tgb, tq0), t01 = ld32(tgb, tq0 + 1);
      const double t10 = ld32(tgb, tq0 + T.TT), t11 = ld32(tgb, tq0 + T.TT + 1);
      g00 = (float)t00;
      gi0 = frcp((float)(t01 - t00));
      g10 = (float)t10;
      gi1 = frcp((float)(t11 - t10));
      it0 = it1 = -1;
    }
    double pn = p, tn = t, un = useg;
    float4 dpart = make_float4(1.f, 0.f, 0.f, 0.f);  // a cell without data: e_new = e_prev
    if (nt0 >= 2 && nt1 >= 2) {
      int j0 = (int)((tq - g00) * gi0);
      j0 = j0 < 0 ? 0 : (j0 > nt0 - 2 ? nt0 - 2 : j0);
      int j1 = (int)((tq - g10) * gi1);
      j1 = j1 < 0 ? 0 : (j1 > nt1 - 2 ? nt1 - 2 : j1);
      if (j0 != it0) {
        if (j0 == it0 + 1) {
          const int g0 = gi[0];
          gi[0] = gi[1];
          gi[1] = 2 * gi[1] - g0;
        } else if (j0 == it0 - 1) {
          const int g1 = gi[1];
          gi[1] = gi[0];
          gi[0] = 2 * gi[0] - g1;
        }
        it0 = j0;
        const double a = ld32(tgb, tq0 + it0), b = ld32(tgb, tq0 + it0 + 1);
        ta0 = (float)a;
        dt0 = frcp((float)(b - a));
        n00 = ld32(nub, tq0 + it0);
        n01 = ld32(nub, tq0 + it0 + 1);
      }
      if (j1 != it1) {
        if (j1 == it1 + 1) {
          const int g2 = gi[2];
          gi[2] = gi[3];
          gi[3] = 2 * gi[3] - g2;
        } else if (j1 == it1 - 1) {
          const int g3 = gi[3];
          gi[3] = gi[2];
          gi[2] = 2 * gi[2] - g3;
        }
        it1 = j1;
        const double a = ld32(tgb, tq0 + T.TT + it1), b = ld32(tgb, tq0 + T.TT + it1 + 1);
        ta1 = (float)a;
        dt1 = frcp((float)(b - a));
        n10 = ld32(nub, tq0 + T.TT + it1);
        n11 = ld32(nub, tq0 + T.TT + it1 + 1);
      }
      if (n00 >= 2 && n01 >= 2 && n10 >= 2 && n11 >= 2) {
        // first record of each corner column in the table
        unsigned ko[4];
        ko[0] = (tq0 + it0) * T.TU;
        ko[1] = ko[0] + T.TU;
        ko[2] = (tq0 + T.TT + it1) * T.TU;
        ko[3] = ko[2] + T.TU;
        const float4 *kc[4] = {recb + ko[0], recb + ko[1], recb + ko[2], recb + ko[3]};
        const int nc[4] = {n00, n01, n10, n11};
        const float us = (float)useg;
        float ec[4], dep[4] = {0.f, 0.f, 0.f, 0.f}, deu[4] = {0.f, 0.f, 0.f, 0.f};
        if (CGA || !mono) {
          if (RL > 1 && !mono) {
            const int c = sub;
            const float4 *kcc = c == 0 ? kc[0] : (c == 1 ? kc[1] : (c == 2 ? kc[2] : kc[3]));
            const int ncc = c == 0 ? nc[0] : (c == 1 ? nc[1] : (c == 2 ? nc[2] : nc[3]));
            const float e = col_eps_bis(kcc, ncc, col_u_bis(kcc, ncc, ep) + useg);
#pragma unroll
            for (int k = 0; k < 4; k++) ec[k] = __shfl(e, k, RL);
          } else if (!mono) {
#pragma unroll 1
            for (int c = 0; c < 4; c++)
              ec[c] = CGA ? col_eps_bis(kc[c], nc[c], cu)
                          : col_eps_bis(kc[c], nc[c], col_u_bis(kc[c], nc[c], ep) + useg);
          } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
              float4 r;
              gi[c] = col_walk<0>(kc[c], nc[c], cu, gi[c], r);
              ec[c] = eps_from_rec_f(r, cu);
            }
          }
          pn = LOSR(F_P, ipn);
          tn = LOSR(F_T, ipn);
          un = LOSR(L.f_u + ig, ipn);
        } else if constexpr (RL > 1) {
          // the lane's corner: its bracket records, the step, and the
          // corners' results gathered from the quad
          const int c = sub;
          const float4 *kcc = c == 0 ? kc[0] : (c == 1 ? kc[1] : (c == 2 ? kc[2] : kc[3]));
          const int ncc = c == 0 ? nc[0] : (c == 1 ? nc[1] : (c == 2 ? nc[2] : nc[3]));
          const unsigned koc = c == 0 ? ko[0] : (c == 1 ? ko[1] : (c == 2 ? ko[2] : ko[3]));
          int g = c == 0 ? gi[0] : (c == 1 ? gi[1] : (c == 2 ? gi[2] : gi[3]));
          g = g < 0 ? 0 : (g > ncc - 2 ? ncc - 2 : g);
          const float4 r = ld32(recb, koc + g);
          const float4 rn = ld32(recb, koc + (g + 1 < ncc - 2 ? g + 1 : ncc - 2));
          __builtin_amdgcn_sched_barrier(0);
          pn = LOSR(F_P, ipn);
          tn = LOSR(F_T, ipn);
          un = LOSR(L.f_u + ig, ipn);
          __builtin_amdgcn_sched_barrier(0);
          float e, de = 0.f, du = 0.f;
          bool slow;
          if constexpr (PART) {
            ega_step_fast2_d(r, rn, ncc, ep, us, g, e, de, du, slow);
            if (slow) e = ega_corner_f32_d(kcc, ncc, ep, useg, g, de, du);
          } else {
            ega_step_fast2(r, rn, ncc, ep, us, useg, g, e, slow);
            if (slow) e = ega_step_slow(kcc, ncc, ep, useg, g, r);
          }
#pragma unroll
          for (int k = 0; k < 4; k++) {
            ec[k] = __shfl(e, k, RL);
            gi[k] = __shfl(g, k, RL);
            if constexpr (PART) {
              dep[k] = __shfl(de, k, RL);
              deu[k] = __shfl(du, k, RL);
            }
          }
        } else {
          // the bracket records at the previous segment's index and the next
          float4 r[4], rn[4];
#pragma unroll
          for (int c = 0; c < 4; c++) {
            int g = gi[c];
            g = g < 0 ? 0 : (g > nc[c] - 2 ? nc[c] - 2 : g);
            gi[c] = g;
            r[c] = ld32(recb, ko[c] + g);
            rn[c] = ld32(recb, ko[c] + (g + 1 < nc[c] - 2 ? g + 1 : nc[c] - 2));
          }
          __builtin_amdgcn_sched_barrier(0);
          pn = LOSR(F_P, ipn);
          tn = LOSR(F_T, ipn);
          un = LOSR(L.f_u + ig, ipn);
          __builtin_amdgcn_sched_barrier(0);
          // common case: the inverse and forward brackets are those records
          bool slow[4];
          if constexpr (PART) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
              int gg = gi[c];
              ega_step_fast2_d(r[c], rn[c], nc[c], ep, us, gg, ec[c], dep[c], deu[c], slow[c]);
              gi[c] = gg;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
              if (slow[c]) ec[c] = ega_corner_f32_d(kc[c], nc[c], ep, useg, gi[c], dep[c], deu[c]);
            }
          } else {
#pragma unroll
            for (int c = 0; c < 4; c++) ega_step_fast2(r[c], rn[c], nc[c], ep, us, useg, gi[c], ec[c], slow[c]);
#pragma unroll
            for (int c = 0; c < 4; c++) {
              if (slow[c]) ec[c] = ega_step_slow(kc[c], nc[c], ep, useg, gi[c], r[c]);
            }
          }
        }
        const float w0 = (tq - ta0) * dt0;
        const float w1 = (tq - ta1) * dt1;
        const float a = ec[0] + (ec[1] - ec[0]) * w0;
        const float b = ec[2] + (ec[3] - ec[2]) * w1;
        float wp;
        if (CGA)
          wp = (pq / pa > 0 && pb / pa > 0) ? __logf((float)(pq / pa)) * lpr : (float)(pq - pa) * wp_scale;
        else
          wp = (float)(pq - pa) * wp_scale;
        float en = a + (b - a) * wp;
        if constexpr (PART) {
          // the bilinear cell weights and their T and p slopes
          const float q0 = 1.f - wp, q1 = wp;
          const float cw[4] = {q0 * (1.f - w0), q0 * w0, q1 * (1.f - w1), q1 * w1};
          const float tw[4] = {-q0 * dt0, q0 * dt0, -q1 * dt1, q1 * dt1};
          const float pw[4] = {-(1.f - w0) * wp_scale, -w0 * wp_scale, (1.f - w1) * wp_scale, w1 * wp_scale};
          float sdep = 0.f, sdu = 0.f, sdt = 0.f, sdp = 0.f;
#pragma unroll
          for (int c = 0; c < 4; c++) {
            sdep += cw[c] * dep[c];
            sdu += cw[c] * deu[c];
            sdt += tw[c] * ec[c];
            sdp += pw[c] * ec[c];
          }
          // a clamped emissivity is constant
          dpart = (en < 0.f || en > 1.f) ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(sdep, sdp, sdt, sdu);
        }
        en = (en < 1.f) ? en : 1.f;
        en = (en > 0.f) ? en : 0.f;
        ep = en;
      } else {
        pn = LOSR(F_P, ipn);
        tn = LOSR(F_T, ipn);
        un = LOSR(L.f_u + ig, ipn);
      }
    } else {
      pn = LOSR(F_P, ipn);
      tn = LOSR(F_T, ipn);
      un = LOSR(L.f_u + ig, ipn);
    }
    if (wr) {
      st32(outb, (unsigned)ip * WAVE + ln, ep);
      if (PART) st32(partb, (unsigned)ip * WAVE + ln, dpart);
    }
    if (ep >= 1.f) {
      // tau_path below float resolution: every later segment is opaque too
      if (wr)
        for (ip++; ip < R.np; ip++) st32(outb, (unsigned)ip * WAVE + ln, 1.f);
      break;
    }
    p = pn;
    t = tn;
    useg = un;
  }
}

void launch_lookup_f32(const ModelView &M, const BatchView &B, const LosView &L, const EpsView &E,
                       const RayInfo *info, int blk0, int nblk, hipStream_t s) {
  if (M.nd * M.ng == 0) return;
  const dim3 grid(padded_blocks(nblk) * M.nd * M.ng), blk(WAVE);
  // small batches, EGA: four lanes per chain while that stays within the
  // five wavefronts per SIMD its registers allow
  if (M.formod != 0 && (size_t)padded_blocks(nblk) * M.nd * M.ng * 4 <= 5120) {
    const dim3 g4(padded_blocks(nblk) * M.nd * M.ng * 4);
    if (E.part)
      hipLaunchKernelGGL((k_lookup_f32<false, true, 4>), g4, blk, 0, s, M, B, L, E, info, blk0, nblk);
    else
      hipLaunchKernelGGL((k_lookup_f32<false, false, 4>), g4, blk, 0, s, M, B, L, E, info, blk0, nblk);
    return;
  }
  if (M.formod == 0)
    hipLaunchKernelGGL((k_lookup_f32<true, false>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  else if (E.part)
    hipLaunchKernelGGL((k_lookup_f32<false, true>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
  else
    hipLaunchKernelGGL((k_lookup_f32<false, false>), grid, blk, 0, s, M, B, L, E, info, blk0, nblk);
}

}  // namespace jur
