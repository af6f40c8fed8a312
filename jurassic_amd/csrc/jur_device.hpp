// Device helpers of the FP64 path.  Every function keeps the expression
// tree and the float/double promotions of the reference function it names,
// with multiply-add contraction disabled, so that the FP64 kernels agree with
// the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include "jur_internal.hpp"

namespace jur {

#define JUR_FP_EXACT _Pragma("clang fp contract(off)")

// LIN (src/jurassic.h:164-165) on doubles.
__device__ __forceinline__ double lin(double x0, double y0, double x1, double y1, double x) {
  JUR_FP_EXACT
  return y0 + (y1 - y0) / (x1 - x0) * (x - x0);
}

// LIN with float table entries: the slope is formed in float arithmetic, then
// multiplied by the double (x - x0), as C promotion does in the reference.
__device__ __forceinline__ double lin_f(float x0, float y0, float x1, float y1, double x) {
  JUR_FP_EXACT
  const float s = (y1 - y0) / (x1 - x0);
  return (double)y0 + (double)s * (x - (double)x0);
}

// LOGY (src/jurassic.h:174-177).
__device__ __forceinline__ double logy(double x0, double y0, double x1, double y1, double x) {
  JUR_FP_EXACT
  return (y1 / y0 > 0) ? (y0 * exp(log(y1 / y0) / (x1 - x0) * (x - x0))) : lin(x0, y0, x1, y1, x);
}

// LOGX (src/jurassic.h:168-171).
__device__ __forceinline__ double logx(double x0, double y0, double x1, double y1, double x) {
  JUR_FP_EXACT
  return (x / x0 > 0 && x1 / x0 > 0) ? (y0 + (y1 - y0) * log(x / x0) / log(x1 / x0))
                                      : lin(x0, y0, x1, y1, x);
}

// geo2cart (src/jurassic.c:3500-3516).
__device__ __forceinline__ void geo2cart(double z, double lon, double lat, double *x) {
  JUR_FP_EXACT
  const double radius = z + JUR_RE;
  const double latrad = lat / 180. * JUR_PI;
  const double lonrad = lon / 180. * JUR_PI;
  const double coslat = cos(latrad);
  x[0] = radius * coslat * cos(lonrad);
  x[1] = radius * coslat * sin(lonrad);
  x[2] = radius * sin(latrad);
}

// cart2geo (src/jurassic.c:108-119).
__device__ __forceinline__ void cart2geo(const double *x, double *z, double *lon, double *lat) {
  JUR_FP_EXACT
  const double radius = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  *lat = asin(x[2] / radius) * (180.0 / JUR_PI);
  *lon = atan2(x[1], x[0]) * (180.0 / JUR_PI);
  *z = radius - JUR_RE;
}

// locate_irr (src/jurassic.c:4103-4129): bisection, either grid order.
template <typename T>
__host__ __device__ __forceinline__ int locate_irr(const T *xx, int n, double x) {
  int ilo = 0, ihi = n - 1, i = (ihi + ilo) >> 1;
  if ((double)xx[i] < (double)xx[i + 1])
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if ((double)xx[i] > x) ihi = i; else ilo = i;
    }
  else
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if ((double)xx[i] <= x) ihi = i; else ilo = i;
    }
  return ilo;
}

// locate_irr by a local walk from a previous index g, for strictly
// monotone grids (dir = +1 ascending, -1 descending).  The result satisfies
// the bisection's end condition, which for a strictly monotone grid has a
// single solution, so it equals locate_irr's.
template <typename T>
__device__ __forceinline__ int irr_walk(const T *xx, int n, double x, int g, int dir) {
  g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
  if (dir > 0) {
    while (g > 0 && (double)xx[g] > x) g--;
    while (g < n - 2 && !((double)xx[g + 1] > x)) g++;
  } else {
    while (g > 0 && (double)xx[g] <= x) g--;
    while (g < n - 2 && !((double)xx[g + 1] <= x)) g++;
  }
  return g;
}

// locate_reg (src/jurassic.c:4133-4148).
__device__ __forceinline__ int locate_reg(const double *xx, int n, double x) {
  JUR_FP_EXACT
  const int i = (int)((x - xx[0]) / (xx[1] - xx[0]));
  return i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
}

// ------------------------------------------------------------------
// Column search on a table column of n >= 2 records.
// C selects the searched component: 0 = u (rec.x / rec.z), 1 = eps (rec.y / rec.w).
// Returns i in [0, n-2] with (i == 0 or v_i <= x) and (i == n-2 or v_{i+1} > x),
// which for a non-decreasing column is the index the reference bisection
// locate_tbl (src/jurassic.c:4152-4170) returns.  The walk starts at the
// previous segment's index `g`, which is usually already right.
// ------------------------------------------------------------------
template <int C>
__device__ __forceinline__ float lo_of(const float4 &r) { return C == 0 ? r.x : r.y; }
template <int C>
__device__ __forceinline__ float hi_of(const float4 &r) { return C == 0 ? r.z : r.w; }

template <int C>
__device__ __forceinline__ int col_walk(const float4 *col, int n, double x, int g, float4 &r) {
  g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
  r = col[g];
  if (g > 0 && (double)lo_of<C>(r) > x) {
    // answer < g: gallop left to a lower bound, then bisect
    int hi = g, step = 1, lo;
    float4 rl;
    for (;;) {
      int j = hi - step;
      if (j <= 0) { lo = 0; rl = col[0]; break; }
      float4 rj = col[j];
      if ((double)lo_of<C>(rj) <= x) { lo = j; rl = rj; break; }
      hi = j;
      step <<= 1;
    }
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const float4 rm = col[mid];
      if ((double)lo_of<C>(rm) > x) hi = mid; else { lo = mid; rl = rm; }
    }
    r = rl;
    return lo;
  }
  if (g < n - 2 && (double)hi_of<C>(r) <= x) {
    // answer > g: gallop right to an upper bound, then bisect
    int lo = g, step = 1, hi;
    float4 rh;
    for (;;) {
      int j = lo + step;
      if (j >= n - 2) { hi = n - 2; rh = col[n - 2]; break; }
      float4 rj = col[j];
      if ((double)hi_of<C>(rj) > x) { hi = j; rh = rj; break; }
      lo = j;
      step <<= 1;
    }
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const float4 rm = col[mid];
      if ((double)hi_of<C>(rm) > x) { hi = mid; rh = rm; } else lo = mid;
    }
    r = rh;
    return hi;
  }
  return g;
}

// col_walk with the record at the start index already in hand (rg = col[g],
// g in [0, n-2]): the common case, x still inside that bracket, loads nothing.
template <int C>
__device__ __forceinline__ int col_walk_r(const float4 *col, int n, double x, int g, const float4 &rg,
                                          float4 &r) {
  const bool left = g > 0 && (double)lo_of<C>(rg) > x;
  const bool right = g < n - 2 && (double)hi_of<C>(rg) <= x;
  if (!left && !right) {
    r = rg;
    return g;
  }
  return col_walk<C>(col, n, x, g, r);
}

// intpol_tbl_u (src/jurassic.c:3933-3967) evaluated from the bracket record
// r that the walk returned for x on a monotone column: x below the first
// knot leaves the walk at record 0 = (u_0, eps_0, ...), x above the last knot
// at record n-2 = (..., u_{n-1}, eps_{n-1}).
__device__ __forceinline__ double u_from_rec(const float4 &r, double x) {
  JUR_FP_EXACT
  if (x < (double)r.y) return lin_f(0.f, 0.f, r.y, r.x, x);
  if (x > (double)r.w) {
    const double a = log((double)(1.f - r.w)) / (double)r.z;
    return log(1 - x) / a;
  }
  return lin_f(r.y, r.x, r.w, r.z, x);
}

// intpol_tbl_eps (src/jurassic.c:3895-3929) from a bracket record.
__device__ __forceinline__ double eps_from_rec(const float4 &r, double x) {
  JUR_FP_EXACT
  if (x < (double)r.x) return lin_f(0.f, 0.f, r.x, r.y, x);
  if (x > (double)r.z) {
    const double a = log((double)(1.f - r.w)) / (double)r.z;
    return 1 - exp(a * x);
  }
  return lin_f(r.x, r.y, r.z, r.w, x);
}

// col_walk for a float key (the FP32 path): same end condition, float
// comparisons.
template <int C>
__device__ __forceinline__ int col_walk_f(const float4 *col, int n, float x, int g, float4 &r) {
  g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
  r = col[g];
  if (g > 0 && lo_of<C>(r) > x) {
    int hi = g, step = 1, lo;
    float4 rl;
    for (;;) {
      int j = hi - step;
      if (j <= 0) { lo = 0; rl = col[0]; break; }
      float4 rj = col[j];
      if (lo_of<C>(rj) <= x) { lo = j; rl = rj; break; }
      hi = j;
      step <<= 1;
    }
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const float4 rm = col[mid];
      if (lo_of<C>(rm) > x) hi = mid; else { lo = mid; rl = rm; }
    }
    r = rl;
    return lo;
  }
  if (g < n - 2 && hi_of<C>(r) <= x) {
    int lo = g, step = 1, hi;
    float4 rh;
    for (;;) {
      int j = lo + step;
      if (j >= n - 2) { hi = n - 2; rh = col[n - 2]; break; }
      float4 rj = col[j];
      if (hi_of<C>(rj) > x) { hi = j; rh = rj; break; }
      lo = j;
      step <<= 1;
    }
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const float4 rm = col[mid];
      if (hi_of<C>(rm) > x) { hi = mid; rh = rm; } else lo = mid;
    }
    r = rh;
    return hi;
  }
  return g;
}

// Fast float reciprocal and division for the FP32 path (v_rcp_f32, 1 ulp):
// the default HIP float division here is IEEE-correct, which the FP64
// path's float slopes need, but the FP32 path does not.
// base[i] with a 32-bit element index: a wave-uniform base plus a per-lane
// 32-bit byte offset, which the compiler addresses as SGPR base + VGPR offset
// without 64-bit arithmetic per lane.  For arrays far below 4 GB (one
// table, one ray block's LOS or emissivities).
template <typename V>
__device__ __forceinline__ V ld32(const V *base, unsigned i) {
  return *(const V *)((const char *)base + (size_t)(i * (unsigned)sizeof(V)));
}
template <typename V>
__device__ __forceinline__ void st32(V *base, unsigned i, V v) {
  *(V *)((char *)base + (size_t)(i * (unsigned)sizeof(V))) = v;
}

__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }

// 1/x in FP64 to about an ulp: v_rcp_f64 and two Newton steps (five
// instructions against about fifteen for a correctly rounded division), for
// the FP32 Jacobian's sweeps, which need FP64 range and accuracy but not a
// correctly rounded quotient.  x > 0 and finite.
__device__ __forceinline__ double rcp_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// tanh(x) for x >= 0 in float: (e - 1)/(e + 1) with e = exp(2x) from the
// hardware exponential (relative error about 1e-7 / x below x = 1, a few
// ulp above; 1 beyond x = 9, where tanhf rounds to 1 too).
__device__ __forceinline__ float tanh_pos_f(float x) {
  if (x > 9.f) return 1.f;
  const float e = __expf(2.f * x);
  return fdiv(e - 1.f, e + 1.f);
}

// The reference bisection itself (for columns that are not monotone).
template <int C>
__device__ __forceinline__ int col_bisect(const float4 *col, int n, double x, float4 &r) {
  int ilo = 0, ihi = n - 1, i;
  while (ihi > ilo + 1) {
    i = (ihi + ilo) >> 1;
    if ((double)lo_of<C>(col[i]) > x) ihi = i; else ilo = i;
  }
  r = col[ilo];
  return ilo;
}

// ---- FP32-path table interpolation (lookup_f32.hip, adjoint.hip) ----
// intpol_tbl_eps from a bracket record, float interpolation.
__device__ __forceinline__ float eps_from_rec_f(const float4 &r, double x) {
  if (x < (double)r.x) return fdiv(r.y, r.x) * (float)x;
  if (x > (double)r.z) return 1.f - __expf(fdiv(__logf(1.f - r.w), r.z) * (float)x);
  return r.y + fdiv(r.w - r.y, r.z - r.x) * (float)(x - (double)r.x);
}

// col_walk_f with the record at the (clamped) start index already in hand.
template <int C>
__device__ __forceinline__ int col_walk_f_r(const float4 *col, int n, float x, int g, const float4 &rg,
                                            float4 &r) {
  const bool left = g > 0 && lo_of<C>(rg) > x;
  const bool right = g < n - 2 && hi_of<C>(rg) <= x;
  if (!left && !right) {
    r = rg;
    return g;
  }
  return col_walk_f<C>(col, n, x, g, r);
}

// One EGA growth step on one table column, FP32 path: from the path
// emissivity ep, u_eq = u(ep) (intpol_tbl_u, src/jurassic.c:3933-3967), then
// eps(u_eq + u_seg) (intpol_tbl_eps, :3895-3929), both evaluated relative to
// the knot of the inverse lookup's bracket r.
//
// ega_step_fast covers the common case without branches: ep inside the eps
// bracket of the record at the previous index (no walk, not below the first
// or above the last knot) and u_eq + u_seg still inside its u bracket.
// Otherwise slow is set and ega_step_slow gives the result.
__device__ __forceinline__ void ega_step_fast(const float4 &r, float ep, float us, bool gt0, bool ltn2,
                                              float &ec, bool &slow) {
  const float du = r.z - r.x, de = r.w - r.y;
  const float off = (ep - r.y) * fdiv(du, de) + us;
  const bool walk = (gt0 && r.y > ep) || (ltn2 && r.w <= ep);
  const bool outside = ep < r.y || ep > r.w;
  slow = walk || outside || !(off < du);
  ec = r.y + off * fdiv(de, du);
}

// The fast step with the records at g and g + 1 in hand (r0 = col[g], r1 =
// col[min(g + 1, n - 2)]): it also covers the inverse bracket moving up by one
// knot and the grown column crossing into the next u bracket, which is what a
// growing path emissivity does every few segments.  Results are those of
// ega_step_slow for the same inputs (same brackets, same formulas); slow is
// set for everything else (walks of more than one knot, the extrapolation
// branches below the first and above the last knot).  g returns the inverse
// lookup's bracket index.
__device__ __forceinline__ void ega_step_fast2(const float4 &r0, const float4 &r1, int n, float ep, float us,
                                               double useg, int &g, float &ec, bool &slow) {
  const bool top = g >= n - 2;
  // inverse bracket: stay at g, or move to g + 1
  const bool stay = (g == 0 || r0.y <= ep) && (top || r0.w > ep);
  const bool up = !top && r0.w <= ep && (g + 1 >= n - 2 || r1.w > ep);
  const float4 rb = stay ? r0 : r1;
  const int b = stay ? g : g + 1;
  bool ok = (stay || up) && !(ep < rb.y) && !(ep > rb.w);
  const float du = rb.z - rb.x, de = rb.w - rb.y;
  const float off = (ep - rb.y) * fdiv(du, de) + us;
  const float e_in = rb.y + off * fdiv(de, du);
  // forward bracket when off leaves the inverse bracket: the record after b
  // (only r1 is in hand, so only from b == g), or b itself where rounding of
  // du left arg below its upper knot
  const double arg = (double)rb.x + (double)off;
  const bool cross = !(off < du);
  const bool b_top = b >= n - 2;
  const bool inb = b_top ? !(arg > (double)rb.z) : !((double)rb.z <= arg);
  const bool nxt = stay && !top && (double)r0.z <= arg && (g + 1 >= n - 2 ? !(arg > (double)r1.z) : (double)r1.z > arg);
  const float4 rf = inb ? rb : r1;
  const float e_x = rf.y + fdiv(rf.w - rf.y, rf.z - rf.x) * (float)(arg - (double)rf.x);
  ok = ok && (!cross || inb || nxt);
  ec = cross ? e_x : e_in;
  // above the last knot (paths that have grown past a column's last
  // emissivity, e.g. a limb path climbing into thin layers)
  const bool above = (stay || up) && b_top && ep > rb.w;
  if (above) ec = 1.f - (1.f - ep) * __expf(fdiv(__logf(1.f - rb.w), rb.z) * us);
  ok = ok || above;
  slow = !ok;
  g = ok ? b : g;
}

// ega_step_fast2 with the partials of the step: dep = d eps / d ep and
// deu = d eps / d u_seg, the slopes of the brackets used (as the FP32
// lookup evaluates them: reciprocal-multiply float slopes).
__device__ __forceinline__ void ega_step_fast2_d(const float4 &r0, const float4 &r1, int n, float ep, float us,
                                                 int &g, float &ec, float &dep, float &deu, bool &slow) {
  const bool top = g >= n - 2;
  const bool stay = (g == 0 || r0.y <= ep) && (top || r0.w > ep);
  const bool up = !top && r0.w <= ep && (g + 1 >= n - 2 || r1.w > ep);
  const float4 rb = stay ? r0 : r1;
  const int b = stay ? g : g + 1;
  bool ok = (stay || up) && !(ep < rb.y) && !(ep > rb.w);
  const float du = rb.z - rb.x, de = rb.w - rb.y;
  const float su = fdiv(du, de);
  const float off = (ep - rb.y) * su + us;
  const float se_in = fdiv(de, du);
  const float e_in = rb.y + off * se_in;
  const double arg = (double)rb.x + (double)off;
  const bool cross = !(off < du);
  const bool b_top = b >= n - 2;
  const bool inb = b_top ? !(arg > (double)rb.z) : !((double)rb.z <= arg);
  const bool nxt = stay && !top && (double)r0.z <= arg && (g + 1 >= n - 2 ? !(arg > (double)r1.z) : (double)r1.z > arg);
  const float4 rf = inb ? rb : r1;
  const float sf = fdiv(rf.w - rf.y, rf.z - rf.x);
  const float e_x = rf.y + sf * (float)(arg - (double)rf.x);
  ok = ok && (!cross || inb || nxt);
  ec = cross ? e_x : e_in;
  deu = cross ? sf : se_in;
  dep = deu * su;
  const bool above = (stay || up) && b_top && ep > rb.w;
  if (above) {
    const float a = fdiv(__logf(1.f - rb.w), rb.z);
    const float ex = __expf(a * us);
    ec = 1.f - (1.f - ep) * ex;
    deu = ex > 0.f ? -a * (1.f - ep) * ex : 0.f;
    dep = ex;
  }
  ok = ok || above;
  slow = !ok;
  g = ok ? b : g;
}

// The general step, starting the inverse search at g (record rg = col[g] in
// hand); g returns the inverse lookup's bracket index.
__device__ __forceinline__ float ega_step_slow(const float4 *col, int n, float ep, double useg, int &g,
                                               const float4 &rg) {
  float4 r;
  const int i = col_walk_f_r<1>(col, n, ep, g, rg, r);
  g = i;
  if (ep < r.y) {
    // below the first knot: u_eq = ep u_0 / eps_0 (small), absolute
    const double arg = (double)(fdiv(r.x, r.y) * ep) + useg;
    float4 r2;
    col_walk_r<0>(col, n, arg, i, r, r2);
    return eps_from_rec_f(r2, arg);
  }
  if (ep > r.w) {
    // above the last knot: eps(u) = 1 - exp(a u), a = log(1 - eps_last) /
    // u_last, so eps(u_eq + u_seg) = 1 - (1 - ep) exp(a u_seg)
    const float a = fdiv(__logf(1.f - r.w), r.z);
    return 1.f - (1.f - ep) * __expf(a * (float)useg);
  }
  const float du = r.z - r.x, de = r.w - r.y;
  const float off = (ep - r.y) * fdiv(du, de) + (float)useg;
  if (off < du) return r.y + off * fdiv(de, du);
  const double arg = (double)r.x + (double)off;
  float4 r2;
  col_walk_r<0>(col, n, arg, i, r, r2);
  return eps_from_rec_f(r2, arg);
}

// FP32 path: eps_from_rec_f with its slope.
__device__ __forceinline__ float eps_rec_f_vs(const float4 &r, double x, float &s) {
  if (x < (double)r.x) {
    s = fdiv(r.y, r.x);
    return s * (float)x;
  }
  if (x > (double)r.z) {
    const float a = fdiv(__logf(1.f - r.w), r.z);
    const float ex = __expf(a * (float)x);
    s = ex > 0.f ? -a * ex : 0.f;
    return 1.f - ex;
  }
  s = fdiv(r.w - r.y, r.z - r.x);
  return r.y + s * (float)(x - (double)r.x);
}

// General FP32 EGA step on one monotone column (ega_step_slow) with its
// partials de = d eps / d ep and se = d eps / d u_seg; the inverse search
// starts at g and g returns its bracket.
__device__ __forceinline__ float ega_corner_f32_d(const float4 *col, int n, float ep, double x, int &g,
                                                  float &de, float &se) {
  float4 r;
  const int i = col_walk_f<1>(col, n, ep, g, r);
  g = i;
  float e, su;
  if (ep < r.y) {
    su = fdiv(r.x, r.y);
    const double arg = (double)(su * ep) + x;
    float4 r2;
    col_walk_r<0>(col, n, arg, i, r, r2);
    e = eps_rec_f_vs(r2, arg, se);
    de = se * su;
  } else if (ep > r.w) {
    // e = 1 - (1 - ep) exp(a u_seg), as ega_step_slow evaluates it
    const float a = fdiv(__logf(1.f - r.w), r.z);
    const float ex = __expf(a * (float)x);
    e = 1.f - (1.f - ep) * ex;
    se = ex > 0.f ? -a * (1.f - ep) * ex : 0.f;
    de = ex;
  } else {
    const float dur = r.z - r.x, der = r.w - r.y;
    su = fdiv(dur, der);
    const float off = (ep - r.y) * su + (float)x;
    if (off < dur) {
      se = fdiv(der, dur);
      e = r.y + off * se;
    } else {
      const double arg = (double)r.x + (double)off;
      float4 r2;
      col_walk_r<0>(col, n, arg, i, r, r2);
      e = eps_rec_f_vs(r2, arg, se);
    }
    de = se * su;
  }
  return e;
}

// Reference-order lookups for columns that are not monotone (bisection).
__device__ __forceinline__ float col_eps_bis(const float4 *col, int n, double x) {
  const float4 r0 = col[0];
  if (x < (double)r0.x) return fdiv(r0.y, r0.x) * (float)x;
  const float4 rl = col[n - 1];
  if (x > (double)rl.x) return 1.f - __expf(fdiv(__logf(1.f - rl.y), rl.x) * (float)x);
  float4 r;
  col_bisect<0>(col, n, x, r);
  return r.y + fdiv(r.w - r.y, r.z - r.x) * (float)(x - (double)r.x);
}

__device__ __forceinline__ double col_u_bis(const float4 *col, int n, float x) {
  const float4 r0 = col[0];
  if (x < r0.y) return (double)(fdiv(r0.x, r0.y) * x);
  const float4 rl = col[n - 1];
  if (x > rl.y) return (double)(log1pf(-x) * fdiv(rl.x, __logf(1.f - rl.y)));
  float4 r;
  col_bisect<1>(col, n, (double)x, r);
  return (double)r.x + (double)(fdiv(r.z - r.x, r.w - r.y) * (x - r.y));
}

// intpol_tbl_eps on one column (src/jurassic.c:3895-3929).
// For a monotone column the local walk runs first and the two boundary
// branches are read off the bracket record it returns: x below the first
// knot can only leave the walk at record 0, which holds (u_0, eps_0), and x
// above the last knot only at record n-2, which holds (u_{n-1}, eps_{n-1}).
// Branch conditions and arithmetic are those of the reference.
__device__ __forceinline__ double col_eps(const float4 *col, int n, double x, int mono, int &g) {
  JUR_FP_EXACT
  float4 r;
  if (mono) {
    g = col_walk<0>(col, n, x, g, r);
    if (x < (double)r.x) return lin_f(0.f, 0.f, r.x, r.y, x);
    if (x > (double)r.z) {
      const double a = log((double)(1.f - r.w)) / (double)r.z;
      return 1 - exp(a * x);
    }
    return lin_f(r.x, r.y, r.z, r.w, x);
  }
  const float4 r0 = col[0];
  if (x < (double)r0.x) return lin_f(0.f, 0.f, r0.x, r0.y, x);
  const float4 rl = col[n - 1];
  if (x > (double)rl.x) {
    const double a = log((double)(1.f - rl.y)) / (double)rl.x;
    return 1 - exp(a * x);
  }
  g = col_bisect<0>(col, n, x, r);
  return lin_f(r.x, r.y, r.z, r.w, x);
}

// intpol_tbl_u on one column (src/jurassic.c:3933-3967), same scheme.
__device__ __forceinline__ double col_u(const float4 *col, int n, double x, int mono, int &g) {
  JUR_FP_EXACT
  float4 r;
  if (mono) {
    g = col_walk<1>(col, n, x, g, r);
    if (x < (double)r.y) return lin_f(0.f, 0.f, r.y, r.x, x);
    if (x > (double)r.w) {
      const double a = log((double)(1.f - r.w)) / (double)r.z;
      return log(1 - x) / a;
    }
    return lin_f(r.y, r.x, r.w, r.z, x);
  }
  const float4 r0 = col[0];
  if (x < (double)r0.y) return lin_f(0.f, 0.f, r0.y, r0.x, x);
  const float4 rl = col[n - 1];
  if (x > (double)rl.y) {
    const double a = log((double)(1.f - rl.y)) / (double)rl.x;
    return log(1 - x) / a;
  }
  g = col_bisect<1>(col, n, x, r);
  return lin_f(r.y, r.x, r.w, r.z, x);
}

// Table cell of a (p, T) point: pressure bracket ipr and the temperature
// brackets it0, it1 of the two pressure levels (intpol_tbl_ega,
// src/jurassic.c:3826-3832).  ok is false when a temperature row has fewer
// than two entries (the reference then uses eps = 0).
struct Cell {
  int ipr, it0, it1;
  bool ok;
  double pa, pb;     // p[ipr], p[ipr+1]
  double ta0, tb0;   // t[ipr][it0], t[ipr][it0+1]
  double ta1, tb1;   // t[ipr+1][it1], t[ipr+1][it1+1]
};

__device__ __forceinline__ Cell locate_cell(const TabView &T, int tix, int npt, double p, double t) {
  Cell c;
  const double *pg = T.p + (size_t)tix * T.TP;
  c.ipr = locate_irr(pg, npt, p);
  c.pa = pg[c.ipr];
  c.pb = pg[c.ipr + 1];
  const int *ntp = T.nt + (size_t)tix * T.TP;
  const int nt0 = ntp[c.ipr], nt1 = ntp[c.ipr + 1];
  c.ok = (nt0 >= 2 && nt1 >= 2);
  c.it0 = c.it1 = 0;
  c.ta0 = c.tb0 = c.ta1 = c.tb1 = 0;
  if (c.ok) {
    const double *tg0 = T.t + ((size_t)tix * T.TP + c.ipr) * T.TT;
    const double *tg1 = tg0 + T.TT;
    c.it0 = locate_reg(tg0, nt0, t);
    c.it1 = locate_reg(tg1, nt1, t);
    c.ta0 = tg0[c.it0];
    c.tb0 = tg0[c.it0 + 1];
    c.ta1 = tg1[c.it1];
    c.tb1 = tg1[c.it1 + 1];
  }
  return c;
}

// Per-(channel, emitter) table constants, loaded once per wavefront.
struct GasConst {
  int tix, npt, mono;
};

// ------------------------------------------------------------------
// Continua (formod_continua, src/jurassic.c:3075-3117, with ctmco2
// :886-1744, ctmh2o :1746-2794, ctmn2 :2796-2859, ctmo2 :2861-2918), FP64.
// ------------------------------------------------------------------
// The continua split into a per-channel part (everything that depends only
// on the wavenumber) and a per-segment part, with the reference's
// operations: a radiance kernel evaluates the channel part once per lane
// instead of once per segment.  Two evaluations differ from the reference's
// expressions by a few ulp (radiances within 5e-16 relative), which halves
// the cost of the H2O term:
//  - the self-continuum power pow(cw260/cw296, (296 - T)/36) is
//    exp((296 - T)/36 * ln(cw260/cw296)), the logarithm taken once per
//    channel; for a ratio of 0, infinity or NaN this is what pow gives too,
//    except at exponent 0 (T = 296 K), where pow gives 1 and so does the
//    code;
//  - tanh(x), x = 0.7193876 nu / T > 0, is em / (em + 2) with em =
//    expm1(2x), which loses no digits for small x (1 beyond x = 20, where
//    tanh(x) rounds to 1).
struct CtmChan {
  bool co2, h2o, n2, o2;
  double co2_296, co2_260, co2_230;
  double h2o_slf, h2o_ratio, h2o_frn;  // sfac*cw296, cw260/cw296, cwfrn*fscal
  double h2o_lr;                       // ln(h2o_ratio)
  double n2_b, n2_beta, o2_b, o2_beta;
};

// h2o_ratio^e (the H2O self-continuum temperature power).
__device__ __forceinline__ double ctm_h2o_pow(const CtmChan &k, double e) {
  JUR_FP_EXACT
  return e == 0 ? 1.0 : exp(e * k.h2o_lr);
}

// tanh(x) for x >= 0.
__device__ __forceinline__ double tanh_pos(double x) {
  JUR_FP_EXACT
  if (x > 20) return 1.0;
  const double em = expm1(2 * x);
  return em / (em + 2);
}

__device__ __forceinline__ CtmChan ctm_chan(const CtmView &c, double nu) {
  JUR_FP_EXACT
  CtmChan k;
  k.co2 = k.h2o = k.n2 = k.o2 = false;
  k.co2_296 = k.co2_260 = k.co2_230 = k.h2o_slf = k.h2o_ratio = k.h2o_frn = 0;
  k.h2o_lr = 0;
  k.n2_b = k.n2_beta = k.o2_b = k.o2_beta = 0;
  {
    const double xw = nu / 2 + 1;
    if (xw >= 1 && xw < 2001) {
      const int iw = (int)xw;
      const double dw = xw - iw;
      const double ew = 1 - dw;
      k.co2 = true;
      k.co2_296 = ew * c.co2_296[iw - 1] + dw * c.co2_296[iw];
      k.co2_260 = ew * c.co2_260[iw - 1] + dw * c.co2_260[iw];
      k.co2_230 = ew * c.co2_230[iw - 1] + dw * c.co2_230[iw];
    }
  }
  {
    const double xw = nu / 10 + 1;
    if (xw >= 1 && xw < 2001) {
      const int iw = (int)xw;
      const double dw = xw - iw;
      const double ew = 1 - dw;
      const double cw296 = ew * c.h2o_296[iw - 1] + dw * c.h2o_296[iw];
      const double cw260 = ew * c.h2o_260[iw - 1] + dw * c.h2o_260[iw];
      const double cwfrn = ew * c.h2o_frn[iw - 1] + dw * c.h2o_frn[iw];
      double sfac;
      if (nu <= 820 || nu >= 960) {
        sfac = 1;
      } else {
        const double xx = (nu - 820) / 10;
        const int ix = (int)xx;
        const double dx = xx - ix;
        sfac = (1 - dx) * c.h2o_xfcrev[ix] + dx * c.h2o_xfcrev[ix + 1];
      }
      const double vf2 = (nu - 370) * (nu - 370);
      const double vf6 = vf2 * vf2 * vf2;
      const double fscal = 36100 / (vf2 + vf6 * 1e-8 + 36100) * -.25 + 1;
      k.h2o = true;
      k.h2o_slf = sfac * cw296;
      k.h2o_ratio = cw260 / cw296;
      k.h2o_lr = log(k.h2o_ratio);
      k.h2o_frn = cwfrn * fscal;
    }
  }
  if (!(nu < c.n2_nu[0] || nu > c.n2_nu[97])) {
    const int idx = locate_reg(c.n2_nu, 98, nu);
    k.n2 = true;
    k.n2_b = lin(c.n2_nu[idx], c.n2_b[idx], c.n2_nu[idx + 1], c.n2_b[idx + 1], nu);
    k.n2_beta = lin(c.n2_nu[idx], c.n2_beta[idx], c.n2_nu[idx + 1], c.n2_beta[idx + 1], nu);
  }
  if (!(nu < c.o2_nu[0] || nu > c.o2_nu[89])) {
    const int idx = locate_reg(c.o2_nu, 90, nu);
    k.o2 = true;
    k.o2_b = lin(c.o2_nu[idx], c.o2_b[idx], c.o2_nu[idx + 1], c.o2_b[idx + 1], nu);
    k.o2_beta = lin(c.o2_nu[idx], c.o2_beta[idx], c.o2_nu[idx + 1], c.o2_beta[idx + 1], nu);
  }
  return k;
}

// The per-segment parts take 1/T (rt) where the reference divides by T, and
// divide by constants as products with their reciprocals, folded at compile
// time: an FP64 division costs about fifteen instructions, and a product by
// a reciprocal differs from the quotient by at most an ulp or so.
__device__ __forceinline__ double ctmco2_seg(const CtmChan &k, double p, double t, double u) {
  JUR_FP_EXACT
  if (!k.co2) return 0;
  const double dt230 = t - 230;
  const double dt260 = t - 260;
  const double dt296 = t - 296;
  const double ctw = dt260 * 5.050505e-4 * dt296 * k.co2_230 - dt230 * 9.259259e-4 * dt296 * k.co2_260 +
                     dt230 * 4.208754e-4 * dt260 * k.co2_296;
  return u * (1 / JUR_NA / 1000 / JUR_P0) * p * ctw;
}

__device__ __forceinline__ double ctmh2o_seg(const CtmChan &k, double nu, double p, double t, double rt,
                                             double q, double u) {
  JUR_FP_EXACT
  if (!k.h2o) return 0;
  const double ctwslf = k.h2o_slf * ctm_h2o_pow(k, (296 - t) * (1.0 / (296 - 260)));
  const double a1 = nu * u * tanh_pos(.7193876 * nu * rt);
  const double a2 = 296 * rt;
  const double a3 = p * (1 / JUR_P0) * (q * ctwslf + (1 - q) * k.h2o_frn) * 1e-20;
  return a1 * a2 * a3;
}

__device__ __forceinline__ double ctmn2_seg(const CtmChan &k, double p, double t, double rt) {
  JUR_FP_EXACT
  if (!k.n2) return 0;
  const double q_n2 = 0.79, t0 = 273.0, tr = 296.0;
  const double r = p * (t0 / JUR_P0) * rt;
  return 0.1 * (r * r) * exp(k.n2_beta * (1 / tr - rt)) * q_n2 * k.n2_b *
         (q_n2 + (1 - q_n2) * (1.294 - t * (0.4545 / tr)));
}

__device__ __forceinline__ double ctmo2_seg(const CtmChan &k, double p, double rt) {
  JUR_FP_EXACT
  if (!k.o2) return 0;
  const double q_o2 = 0.21, t0 = 273, tr = 296;
  const double r = p * (t0 / JUR_P0) * rt;
  return 0.1 * (r * r) * exp(k.o2_beta * (1 / tr - rt)) * q_o2 * k.o2_b;
}

// formod_srcfunc (src/jurassic.c:3483-3496) for one channel column of the
// source-function table (sr_col[TBLNS]).
__device__ __forceinline__ double srcfunc(const double *st, const double *sr_col, double t) {
  JUR_FP_EXACT
  const int it = locate_reg(st, TBLNS, t);
  return lin(st[it], sr_col[it], st[it + 1], sr_col[it + 1], t);
}

// srcfunc on the uniform temperature grid with its inverse spacing inv_dst
// in hand (the per-segment form: products instead of the two divisions of
// locate_reg and LIN).
__device__ __forceinline__ double srcfunc_u(const double *st, double inv_dst, const double *sr_col, double t) {
  JUR_FP_EXACT
  int it = (int)((t - st[0]) * inv_dst);
  it = it < 0 ? 0 : (it > TBLNS - 2 ? TBLNS - 2 : it);
  return sr_col[it] + (sr_col[it + 1] - sr_col[it]) * inv_dst * (t - st[it]);
}

// PLANCK (src/jurassic.h:183-184).
__device__ __forceinline__ double planck(double t, double nu) {
  JUR_FP_EXACT
  return JUR_C1 * (nu * nu * nu) / expm1(JUR_C2 * nu / t);
}

// BRIGHT (src/jurassic.h:126-127).
__device__ __forceinline__ double bright(double rad, double nu) {
  JUR_FP_EXACT
  return JUR_C2 * nu / log1p(JUR_C1 * (nu * nu * nu) / rad);
}

// XCD-aware work mapping.  The hardware hands workgroup ids to the 8 XCDs
// round-robin (id % 8).  A ray block's G work items (its channels, or its
// channel-emitter chains) should share one XCD's L2, since they read the
// same LOS data; consecutive ray blocks should go to different XCDs, since
// block cost varies smoothly with geometry (limb ray length grows with
// depth) and contiguous ranges would unbalance the XCDs.  With the grid padded
// to nblk8 = 8 ceil(nblk/8) blocks, workgroup wg maps to block
// 8 (k / G) + x and item k % G, x = wg % 8, k = wg / 8: a bijection in which
// XCD x owns blocks x, x+8, x+16, ...  Returns false for padding groups.
__device__ __forceinline__ bool xcd_block_map(int wg, int G, int nblk, int &blk, int &item) {
  const int x = wg & 7, k = wg >> 3;
  const int q = k / G;
  item = k - q * G;
  blk = q * 8 + x;
  return blk < nblk;
}

__host__ __device__ __forceinline__ int padded_blocks(int nblk) { return (nblk + 7) & ~7; }

// Contiguous-range remap (kept for kernels without a per-block structure).
__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
  constexpr int NX = 8;
  const int per = nwg / NX, rem = nwg % NX;
  const int x = wg % NX, k = wg / NX;
  // XCD x owns logical range [x*per + min(x, rem), ...) of length per + (x < rem)
  if (k < per) return x * per + (x < rem ? x : rem) + k;
  // tail ids (k == per) exist only for x < rem
  return x * per + x + k;
}

// ------------------------------------------------------------------
// Ray state handed from the march (raytrace.hip) to its finish and to the
// adjoint of the refracted path (adjoint.hip).
// ------------------------------------------------------------------
// First minimum of z along the LOS (gsl_stats_min_index semantics: a NaN
// ends the search) and the positions of its two neighbours, for
// tangent_point.
struct MinTrack {
  double zlow = 0;
  double xbefore[3] = {0, 0, 0}, xafter[3] = {0, 0, 0};
  int imin = 0;
  bool after_pending = false, nan_seen = false;
  // point np at z, position x; xprev is the position of point np - 1
  __device__ __forceinline__ void add(int np, double z, const double *x, const double *xprev) {
    if (!nan_seen) {
      const bool lower = (np == 0) || z < zlow;
      if (lower) zlow = z;
      if (lower || isnan(z)) {
        imin = np;
        if (np > 0)
          for (int i = 0; i < 3; i++) xbefore[i] = xprev[i];
        after_pending = true;
        nan_seen = isnan(z);
      } else if (after_pending) {
        for (int i = 0; i < 3; i++) xafter[i] = x[i];
        after_pending = false;
      }
    } else if (after_pending) {
      for (int i = 0; i < 3; i++) xafter[i] = x[i];
      after_pending = false;
    }
  }
};

// What the march (k_raytrace) hands to k_ray_finish for a ray with LOS
// points: the step that left the altitude range, which is clipped to it
// there, and the state for the tangent point.
struct RayTail {
  double xprev[3];  // the last point written
  double xcut[3];   // the step beyond the altitude range
  double xfirst[3];
  double zprev, zcut, dscut;  // altitudes of both, step length
  double zmin, zmax;
  double ds1, ds2;            // raw step lengths of the last two points written
  double ds_imin, ds_imin1;   // raw step lengths at the first minimum and after it
  double ecut[3];             // the march direction at the step beyond the range
  MinTrack mt;
  int zdir;
};

}  // namespace jur
