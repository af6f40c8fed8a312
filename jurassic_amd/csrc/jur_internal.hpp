// Internal data layout shared by the host code and the HIP kernels of
// libjurassic_gpu.so.  Plain structs only; device pointers are raw.
#pragma once

#include <stdint.h>

#include "../../include/jurassic_gpu.h"

namespace jur {

// Physical constants (same values as src/jurassic.h:252-340).
#define JUR_C1 1.19104259e-8
#define JUR_C2 1.43877506
#define JUR_KB 1.3806504e-23
#define JUR_NA 6.02214199e23
#define JUR_RE 6367.421
#define JUR_P0 1013.25
#define JUR_TMIN 100.
#define JUR_TMAX 400.
#define JUR_TSUN 5780.
#define JUR_PI 3.14159265358979323846

// Rays are processed in blocks of WAVE consecutive rays: one lane per ray.
constexpr int WAVE = 64;

// ------------------------------------------------------------------
// Emissivity tables in HBM, laid out [gas][channel][p][T][u].
// Each (p, T) column of nu knots is stored as nu records
//   rec[i] = (u_i, eps_i, u_{i+1}, eps_{i+1})   (last: (u_n-1, e_n-1, same))
// so that one aligned 16-byte load holds a whole interpolation bracket.
// ------------------------------------------------------------------
struct TabView {
  int nd, ng;
  int TP, TT, TU;          // padded capacities of p, T, u
  const int *np;           // [ng][nd]           (-1: missing table)
  const int *nt;           // [ng][nd][TP]
  const int *nu;           // [ng][nd][TP][TT]
  const double *p;         // [ng][nd][TP]
  const double *t;         // [ng][nd][TP][TT]
  const float4 *rec;       // [ng][nd][TP][TT][TU]
  const int *mono;         // [ng][nd]: 1 if every column is non-decreasing
                           // in u and eps (search by local walk), 0 otherwise
                           // (search by the reference bisection)
  const int *pdir;         // [ng][nd]: +1/-1 strictly ascending/descending p grid,
                           // 0 otherwise (search by bisection)
  int all_mono;            // 1 if mono is 1 for every table with np >= 30
  int shared;              // 1 if every table with np >= 30 has the same p grid,
                           // T-grid sizes and T grids (one cell search per
                           // segment serves all of them)
  int ref_tix;             // a table holding the shared grids
};

// ------------------------------------------------------------------
// Line-of-sight buffer written by the ray tracer, read by the lookup and
// radiance kernels.  Per block of WAVE rays: point-major, then field, lane
// fastest:
//   los[((blk * cap + ip) * nfield + f) * WAVE + lane]
// so that the 64 lanes of a wavefront read one field of one point with a
// single coalesced 512-byte access, and all fields of a point are adjacent.
// ------------------------------------------------------------------
// Longitudes and latitudes of the points are used only by the ray tracer
// itself (tangent point, solar geometry), which computes them where needed.
enum {
  F_Z = 0,
  F_P = 1,
  F_T = 2,
  F_DS = 3,   // step length: raw while tracing, trapezoid rule afterwards
  F_Q0 = 4    // then q[ng], k[nw], cloud factor (if clouds), u[ng]
};

struct LosView {
  double *los;
  int cap;      // points per ray
  int nfield;
  int f_k;      // first window-extinction field
  int f_cl;     // cloud factor field, -1 if none
  int f_u;      // first column-density field
  float *losf;  // FP32 forward pass: the fields rounded to float, same layout
                // (the FP32 radiance kernel rounds every field it reads);
                // null otherwise
  int f_xe;     // Jacobian pass with refraction: first of six fields holding
                // the point's Cartesian position and the march direction
                // leaving it (x0 x1 x2 e0 e1 e2); -1 otherwise
};

// Per-segment emissivities of every (channel, emitter), written by the
// lookup kernels and read by the radiance kernels.  Block blk holds npb[blk]
// points per chain (the longest LOS of its rays), packed from off[blk]:
//   eps[off[blk] + (((id * ng + ig) * npb[blk] + ip) * WAVE + lane]
// FP64 path: the reference's eps of intpol_tbl_ega/cga (double), or, for
// the Jacobian, the path transmittance tau_path after each segment (double).
// FP32 path: the path emissivity 1 - tau_path after each segment (float);
// 1 marks an opaque path.
struct EpsView {
  void *eps;
  const size_t *off;   // [nblk] element offset of each block (chunk-relative)
  const int *npb;      // [nblk] points per chain of each block
  float4 *part;        // FP32 EGA Jacobian: per chain step (d/d e_prev, d/dp, d/dT,
                       // d/du) at the emissivity's index; null otherwise
};

// Jacobian pass (adjoint.hip).  Per chunk of ray blocks:
//   kt  [blk][nd][ncol][WAVE]  K tile of each (ray block, channel), lane fastest
//   scr [soff[blk] + ((id * npb[blk] + ip) * nscr + k) * WAVE + lane]
//       forward-sweep values per point: tau_i (k = 0), D_i (k = 1, reflection)
struct JacView {
  const int *xmap;      // [nmap][JUR_XMAP_LEN] column of each state element, -1 if absent
  int nmap;
  const int *prof_map;  // [nprof] state map of each atmosphere
  int ncol;             // row length of K
  double *kt;
  double *scr;
  const size_t *soff;   // [nblk]
  int nscr;             // values per point: tau_i/H_i, D_i with a reflecting
                        // surface, then d rad / d z_i and d rad / d ds_i when
                        // the path is differentiated
  int *nef;             // [blk][nd][WAVE] points before the first opaque segment
  int kz;               // scratch index of d rad / d z_i (d rad / d ds_i at
                        // kz + 1): the refracted path is differentiated; -1: not
  double *pco;          // kz >= 0: per march step the JUR_NPCO coefficients of
                        // the reverse step, block b at soff[b] / (nd nscr) * JUR_NPCO,
                        // [step][coefficient][lane]
};

// Coefficients per march step of the adjoint of the refracted path: 53
// used and one spare, so that a step's block (JUR_NPCO x 64 lanes x 8 B)
// is a whole number of 1 KiB wavefront loads.
constexpr int JUR_NPCO = 54;

#define LOS_AT(base, L, f, ip) (base)[((size_t)(ip) * (L).nfield + (size_t)(f)) * WAVE]

// Per-ray results of the ray tracer.
struct RayInfo {
  int np;          // LOS points
  int status;      // JUR_OK, JUR_ERR_LOS, JUR_ERR_OBSERVER
  double sft;      // surface temperature if the LOS hit the ground, else -999
  double tpz, tplon, tplat;
  double rcos;     // solar incidence factor (SFTYPE 3), 0 if sun below horizon
  double pad;
};

// Per-channel constants.
struct ChanInfo {
  double nu;
  int window;
  int icl;         // cloud spectral index (locate_irr on clnu), 0 if none
  int isf;         // surface spectral index (locate_irr on sfnu), 0 if none
  int pad;
  double planck_sun;   // PLANCK(TSUN, nu)
};

// Continuum coefficient tables on the device (see ctm_coeffs.h).
struct CtmView {
  const double *co2_296, *co2_260, *co2_230;
  const double *h2o_296, *h2o_260, *h2o_frn, *h2o_xfcrev;
  const double *n2_b, *n2_beta, *n2_nu;
  const double *o2_b, *o2_beta, *o2_nu;
};

// Forward-model settings that the kernels read.
struct ModelView {
  int nd, ng, nw, ncl, nsf;
  int sftype;
  double sfsza;
  int formod;          // 0 CGA, 1 EGA
  int refrac;
  double rayds, raydz;
  int ig_co2, ig_h2o;  // emitter indices for the continua, -1 if off/absent
  int ctm_n2, ctm_o2;
  int write_bbt;
  const ChanInfo *chan;   // [nd]
  const double *clnu;     // [NCL]
  const double *sfnu;     // [NSF]
  const double *st;       // [TBLNS] source-function temperatures
  const double *sr;       // [nd][TBLNS] source-function radiances
  CtmView ctm;
  TabView tab;
};

// One batch of rays (device pointers).
struct BatchView {
  int nprof;
  atm_t *atm;
  int nray;
  const int *ray_prof;
  const double *ray_geom;   // [nray][7]
};

// One retrieval of the batched Levenberg-Marquardt engine (lm_kernels.hip,
// retrieval.cpp): sizes and device buffers.
struct LmDesc {
  int n, m;                  // state size, measurement rows (rays x channels)
  int ray0, nray;            // rays of the retrieval in the current pass
  double *sainv, *cmat, *a;  // n x n: S_a^-1, K^T S_eps^-1 K, work
  double *k;                 // m x n: K scaled by 1/sigma_eps
  double *sig, *dy;          // m: 1/sigma_eps (0: masked), y_meas - y
  double *dx, *b, *step;     // n
  double *g;                 // n x m gain matrix (error analysis only)
};

}  // namespace jur
