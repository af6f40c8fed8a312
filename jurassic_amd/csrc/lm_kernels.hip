// Dense algebra of the Levenberg-Marquardt optimal estimation
// (optimal_estimation, src/retrieval.c:448-666) for a batch of independent
// retrievals.  Retrieval r has a state of n_r elements and m_r measurement
// rows; LmDesc holds its sizes and device buffers.  The kernels take a list
// of the retrievals that need the step, so retrievals at different stages
// of their iteration share launches.
//
//  k_lm_gather   K rows of a retrieval from a pass's K [ray][channel][ncol]
//                into a compact m x n matrix, rows scaled by 1/sigma_eps
//                (rows of masked measurements are zero)
//  k_lm_normal   C = (S^-1/2 K)^T (S^-1/2 K), lower triangle by 64x64 LDS
//                tiles on the FP64 matrix cores, mirrored
//  k_lm_cholinv  S_a^-1 (matrix_invert, src/retrieval.c:373-399): Cholesky,
//                then one column of the inverse per thread
//  k_lm_rhs      b = K^T S_eps^-1 dy - S_a^-1 dx
//  k_lm_solve    A = S_a^-1 (1 + lambda) + C, Cholesky A = L L^T, L L^T s = b
//  k_lm_quad     dx^T S_a^-1 dx (the a priori part of cost_function,
//                src/retrieval.c:340-369)
//  k_ana_cov     error analysis (src/retrieval.c:668-726): retrieval
//                covariance (C + S_a^-1)^-1 into the S_a^-1 buffer
//  k_ana_gain    gain matrix G = cov K^T S_eps^-1 = cov Kc^T diag(1/sigma)
//  k_ana_avk     averaging kernels A = G K = cov C into the work buffer
//
// Cholesky is right-looking by 16-column panels: the panel factored in LDS,
// the trailing matrix (in global memory, which L2 holds: n <= 324 is 0.8 MB)
// updated by 16 x 16 tiles on the FP64 matrix cores.
#include "jur_device.hpp"

namespace jur {

constexpr int LM_T = 256;

// Compact scaled K: kc[i][j] = K[row(i)][j] * sig_inv[i]; rows of masked
// measurements (sig_inv = 0) are zero.  Rows are (ray, channel) in order.
__global__ __launch_bounds__(LM_T) void k_lm_gather(const LmDesc *D, const int *list, const double *kpass,
                                                    int nd, int ncol) {
  const LmDesc d = D[list[blockIdx.y]];
  const int i = blockIdx.x;  // measurement row (ray-major, channel-minor)
  if (i >= d.nray * nd) return;
  const double s = d.sig[i];
  const double *src = kpass + ((size_t)d.ray0 * nd + i) * ncol;
  double *dst = d.k + (size_t)i * d.n;
  for (int j = threadIdx.x; j < d.n; j += LM_T) dst[j] = (s != 0) ? src[j] * s : 0.0;
}

// FP64 matrix-core tile: D (16 x 16) += A (16 x 4) B (4 x 16) with
// v_mfma_f64_16x16x4_f64.  Lane l holds A[l % 16][l / 16], B[l / 16][l % 16]
// and D[l / 16 + 4 v][l % 16] in its v-th accumulator.
typedef double dbl4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ dbl4 mfma_f64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// C = Kc^T Kc: one workgroup per 64x64 tile (ti >= tj), rows of Kc streamed
// through LDS 32 at a time; each of the four wavefronts forms a 32 x 32
// quadrant as 2 x 2 tiles of 16 x 16 on the FP64 matrix cores.
__global__ __launch_bounds__(LM_T) void k_lm_normal(const LmDesc *D, const int *list) {
  __shared__ double a[32][65], b[32][65];
  const LmDesc d = D[list[blockIdx.y]];
  const int nt = (d.n + 63) / 64;
  // lower-triangular tile index -> (ti, tj)
  int t = blockIdx.x, ti = 0;
  while (t > ti) {
    t -= ti + 1;
    ti++;
  }
  const int tj = t;
  if (ti >= nt) return;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const int ii = lane & 15, kq = lane >> 4;
  const int qi = 32 * (wave >> 1), qj = 32 * (wave & 1);
  dbl4 acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++) acc[p][q] = dbl4{0.0, 0.0, 0.0, 0.0};
  const double *K = d.k;
  const int i0 = ti * 64, j0 = tj * 64;
  for (int r0 = 0; r0 < d.m; r0 += 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 64; e += LM_T) {
      const int rr = e >> 6, cc = e & 63;
      const int r = r0 + rr;
      a[rr][cc] = (r < d.m && i0 + cc < d.n) ? K[(size_t)r * d.n + i0 + cc] : 0.0;
      b[rr][cc] = (r < d.m && j0 + cc < d.n) ? K[(size_t)r * d.n + j0 + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll 2
    for (int k4 = 0; k4 < 32; k4 += 4) {
      const double av0 = a[k4 + kq][qi + ii], av1 = a[k4 + kq][qi + 16 + ii];
      const double bv0 = b[k4 + kq][qj + ii], bv1 = b[k4 + kq][qj + 16 + ii];
      acc[0][0] = mfma_f64(av0, bv0, acc[0][0]);
      acc[0][1] = mfma_f64(av0, bv1, acc[0][1]);
      acc[1][0] = mfma_f64(av1, bv0, acc[1][0]);
      acc[1][1] = mfma_f64(av1, bv1, acc[1][1]);
    }
  }
  double *C = d.cmat;
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = i0 + qi + 16 * p + kq + 4 * v, j = j0 + qj + 16 * q + ii;
        if (i < d.n && j < d.n) {
          C[(size_t)i * d.n + j] = acc[p][q][v];
          C[(size_t)j * d.n + i] = acc[p][q][v];
        }
      }
}

// Largest n whose Cholesky and triangular solves use the staged paths below
// (the LDS panel, the per-lane rows of one wavefront).
constexpr int CHOL_NMAX = 512;

// In-place Cholesky A = L L^T of an n x n row-major matrix (lower triangle
// of a holds L; of the upper triangle only the diagonal 16 x 16 blocks may
// change).  Returns false if a pivot is not positive.  Called by all
// threads of the workgroup.
//
// For n <= CHOL_NMAX: right-looking by panels of 16 columns.  The panel
// (rows j0..n-1) is staged in LDS column-major and factored there by the
// first wavefront alone (no workgroup barrier per column); then the four
// wavefronts update the trailing matrix by 16 x 16 tiles on the FP64 matrix
// cores: A22 -= P P^T as four 16 x 16 x 4 products per tile, read and written
// once per panel.  The sums of the 16 panel products per element are formed
// in the matrix core's order, so L agrees with the column-by-column
// algorithm to rounding.
constexpr int CHOL_B = 16;

__device__ bool chol_wg(double *a, int n, int *bad) {
  if (n > CHOL_NMAX) {
    for (int j = 0; j < n; j++) {
      if (threadIdx.x == 0) {
        const double s = a[(size_t)j * n + j];
        if (!(s > 0)) *bad = 1;
        a[(size_t)j * n + j] = sqrt(s);
      }
      __syncthreads();
      if (*bad) return false;
      const double l = a[(size_t)j * n + j];
      for (int i = j + 1 + threadIdx.x; i < n; i += LM_T) a[(size_t)i * n + j] /= l;
      __syncthreads();
      // trailing update of the lower triangle, rows i > j, columns j < k <= i
      const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
      for (int i = j + 1 + ty; i < n; i += 16) {
        const double lij = a[(size_t)i * n + j];
        for (int k = j + 1 + tx; k <= i; k += 16) a[(size_t)i * n + k] -= lij * a[(size_t)k * n + j];
      }
      __syncthreads();
    }
    return true;
  }
  // panel P[c][r] = a[j0 + r][j0 + c], rows padded with zeros to whole tiles
  __shared__ double P[CHOL_B][CHOL_NMAX];
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  for (int j0 = 0; j0 < n; j0 += CHOL_B) {
    const int jb = n - j0 < CHOL_B ? n - j0 : CHOL_B;
    const int m = n - j0;
    const int mp = (m + 15) & ~15;
    for (int e = threadIdx.x; e < CHOL_B * mp; e += LM_T) {
      const int c = e / mp, r = e - c * mp;
      P[c][r] = (c < jb && r < m) ? a[(size_t)(j0 + r) * n + j0 + c] : 0.0;
    }
    __syncthreads();
    if (wave == 0) {
      for (int c = 0; c < jb; c++) {
        const double s = P[c][c];
        if (!(s > 0)) {
          if (lane == 0) *bad = 1;
          break;
        }
        const double l = sqrt(s);
        for (int r = c + 1 + lane; r < m; r += WAVE) P[c][r] /= l;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) P[c][c] = l;
        // the panel's later columns, rows r >= k
        for (int k = c + 1; k < jb; k++) {
          const double pk = P[c][k];
          for (int r = k + lane; r < m; r += WAVE) P[k][r] -= P[c][r] * pk;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    __syncthreads();
    if (*bad) return false;
    // the panel's lower part back to a
    for (int e = threadIdx.x; e < m * jb; e += LM_T) {
      const int r = e / jb, c = e - r * jb;
      if (r >= c) a[(size_t)(j0 + r) * n + j0 + c] = P[c][r];
    }
    // trailing update of the tiles (bi, bk), bi >= bk, of rows and columns
    // t0 + 16 bi .., t0 + 16 bk ..
    const int t0 = j0 + jb;
    if (t0 < n) {
      const int nb = (n - t0 + 15) / 16;
      const int ntile = nb * (nb + 1) / 2;
      const int ii = lane & 15, kq = lane >> 4;
      for (int t = wave; t < ntile; t += LM_T / WAVE) {
        int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
        while (bi * (bi + 1) / 2 > t) bi--;
        const int bk = t - bi * (bi + 1) / 2;
        const int ri = t0 + 16 * bi, ck = t0 + 16 * bk;
        dbl4 acc;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = ri + kq + 4 * v, k = ck + ii;
          acc[v] = (i < n && k < n) ? a[(size_t)i * n + k] : 0.0;
        }
#pragma unroll
        for (int s4 = 0; s4 < CHOL_B / 4; s4++) {
          const int c = 4 * s4 + kq;
          acc = mfma_f64(-P[c][ri - j0 + ii], P[c][ck - j0 + ii], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = ri + kq + 4 * v, k = ck + ii;
          if (i < n && k < n) a[(size_t)i * n + k] = acc[v];
        }
      }
    }
    __syncthreads();
  }
  return true;
}

// Solve L L^T x = b in place in x (x holds b on entry); all threads.  For
// n <= CHOL_NMAX the first wavefront runs both substitutions alone: lane l
// keeps rows l, l + 64, ... in registers, each step's solved element is
// broadcast by a lane shuffle, and the next column of L is loaded while the
// current one is applied, so no step waits for a workgroup barrier or a
// store.
__device__ void chol_solve_wg(const double *l, int n, double *x) {
  if (n > CHOL_NMAX) {
    for (int j = 0; j < n; j++) {
      if (threadIdx.x == 0) x[j] /= l[(size_t)j * n + j];
      __syncthreads();
      const double xj = x[j];
      for (int i = j + 1 + threadIdx.x; i < n; i += LM_T) x[i] -= l[(size_t)i * n + j] * xj;
      __syncthreads();
    }
    for (int j = n - 1; j >= 0; j--) {
      if (threadIdx.x == 0) x[j] /= l[(size_t)j * n + j];
      __syncthreads();
      const double xj = x[j];
      for (int i = threadIdx.x; i < j; i += LM_T) x[i] -= l[(size_t)j * n + i] * xj;
      __syncthreads();
    }
    return;
  }
  constexpr int R = CHOL_NMAX / WAVE;
  if (threadIdx.x < WAVE) {
    const int lane = threadIdx.x;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + WAVE * r;
      acc[r] = i < n ? x[i] : 0.0;
    }
    // forward: L y = b, column j of L applied to the rows below j
    double lc[R], dg = n > 0 ? l[0] : 1.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + WAVE * r;
      lc[r] = (i > 0 && i < n) ? l[(size_t)i * n] : 0.0;
    }
    for (int j = 0; j < n; j++) {
      const int rj = j / WAVE;
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < R; r++) v = r == rj ? acc[r] : v;
      const double xj = __shfl(v / dg, j % WAVE, WAVE);
      // the next column and diagonal element
      double ln[R];
      const int jn = j + 1;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = lane + WAVE * r;
        ln[r] = (jn < n && i > jn && i < n) ? l[(size_t)i * n + jn] : 0.0;
      }
      const double dn = jn < n ? l[(size_t)jn * n + jn] : 1.0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = lane + WAVE * r;
        if (i == j) acc[r] = xj;
        else if (i > j) acc[r] -= lc[r] * xj;
      }
#pragma unroll
      for (int r = 0; r < R; r++) lc[r] = ln[r];
      dg = dn;
    }
    // backward: L^T x = y, row j of L applied to the rows above j
    dg = n > 0 ? l[(size_t)(n - 1) * n + (n - 1)] : 1.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + WAVE * r;
      lc[r] = (n > 0 && i < n - 1) ? l[(size_t)(n - 1) * n + i] : 0.0;
    }
    for (int j = n - 1; j >= 0; j--) {
      const int rj = j / WAVE;
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < R; r++) v = r == rj ? acc[r] : v;
      const double xj = __shfl(v / dg, j % WAVE, WAVE);
      double ln[R];
      const int jn = j - 1;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = lane + WAVE * r;
        ln[r] = (jn >= 0 && i < jn) ? l[(size_t)jn * n + i] : 0.0;
      }
      const double dn = jn >= 0 ? l[(size_t)jn * n + jn] : 1.0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = lane + WAVE * r;
        if (i == j) acc[r] = xj;
        else if (i < j) acc[r] -= lc[r] * xj;
      }
#pragma unroll
      for (int r = 0; r < R; r++) lc[r] = ln[r];
      dg = dn;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + WAVE * r;
      if (i < n) x[i] = acc[r];
    }
  }
  __syncthreads();
}

__device__ void chol_inverse_wg(const double *a, int n, double *inv);

// S_a^-1: a holds S_a on entry and is overwritten by its Cholesky factor;
// inv receives the inverse.  A diagonal S_a is inverted elementwise, as the
// reference does.  status[r] = 1 if S_a is not positive definite.
__global__ __launch_bounds__(LM_T) void k_lm_cholinv(const LmDesc *D, const int *list, int *status) {
  __shared__ int bad, offdiag;
  const int r = list[blockIdx.x];
  const LmDesc d = D[r];
  double *a = d.a;
  double *inv = d.sainv;
  const int n = d.n;
  if (threadIdx.x == 0) bad = offdiag = 0;
  __syncthreads();
  for (long e = threadIdx.x; e < (long)n * n; e += LM_T)
    if (e / n != e % n && a[e] != 0) offdiag = 1;
  __syncthreads();
  if (!offdiag) {
    for (long e = threadIdx.x; e < (long)n * n; e += LM_T) inv[e] = (e / n == e % n) ? 1 / a[e] : 0.0;
    if (threadIdx.x == 0) status[r] = 0;
    return;
  }
  if (!chol_wg(a, n, &bad)) {
    if (threadIdx.x == 0) status[r] = 1;
    return;
  }
  chol_inverse_wg(a, n, inv);
  if (threadIdx.x == 0) status[r] = 0;
}

// (L L^T)^-1 from the Cholesky factor in the lower triangle of a: column c
// of the inverse by L y = e_c, L^T z = y, one thread per column.
__device__ void chol_inverse_wg(const double *a, int n, double *inv) {
  for (int c = threadIdx.x; c < n; c += LM_T) {
    // y (forward), stored in column c of inv, rows >= c
    for (int i = 0; i < n; i++) {
      double v = (i == c) ? 1.0 : 0.0;
      if (i > c)
        for (int k = c; k < i; k++) v -= a[(size_t)i * n + k] * inv[(size_t)k * n + c];
      else if (i < c)
        v = 0.0;
      inv[(size_t)i * n + c] = (i >= c) ? v / a[(size_t)i * n + i] : 0.0;
    }
    for (int i = n - 1; i >= 0; i--) {
      double v = inv[(size_t)i * n + c];
      for (int k = i + 1; k < n; k++) v -= a[(size_t)k * n + i] * inv[(size_t)k * n + c];
      inv[(size_t)i * n + c] = v / a[(size_t)i * n + i];
    }
  }
}

// b = Kc^T (sig_inv * dy) - S_a^-1 dx  (Kc is already scaled by sig_inv).
__global__ __launch_bounds__(LM_T) void k_lm_rhs(const LmDesc *D, const int *list) {
  const LmDesc d = D[list[blockIdx.y]];
  const int j = blockIdx.x * LM_T + threadIdx.x;
  if (j >= d.n) return;
  double v = 0;
  for (int i = 0; i < d.m; i++) v += d.k[(size_t)i * d.n + j] * (d.sig[i] * d.dy[i]);
  double u = 0;
  for (int k = 0; k < d.n; k++) u += d.sainv[(size_t)j * d.n + k] * d.dx[k];
  d.b[j] = -1.0 * u + v;
}

// A = S_a^-1 (1 + lmpar) + C; A = L L^T; step = A^-1 b.
__global__ __launch_bounds__(LM_T) void k_lm_solve(const LmDesc *D, const int *list, const double *lmpar,
                                                   int *status) {
  __shared__ int bad;
  const int r = list[blockIdx.x];
  const LmDesc d = D[r];
  const int n = d.n;
  const double lp = lmpar[r];
  const double *S = d.sainv, *Cm = d.cmat;
  double *A = d.a;
  if (threadIdx.x == 0) bad = 0;
  for (long e = threadIdx.x; e < (long)n * n; e += LM_T) A[e] = S[e] * (1 + lp) + Cm[e];
  double *x = d.step;
  for (int i = threadIdx.x; i < n; i += LM_T) x[i] = d.b[i];
  __syncthreads();
  if (!chol_wg(A, n, &bad)) {
    if (threadIdx.x == 0) status[r] = 1;
    return;
  }
  chol_solve_wg(A, n, x);
  if (threadIdx.x == 0) status[r] = 0;
}

// q[r] = dx^T S_a^-1 dx.
__global__ __launch_bounds__(LM_T) void k_lm_quad(const LmDesc *D, const int *list, double *q) {
  __shared__ double part[LM_T];
  const int r = list[blockIdx.x];
  const LmDesc d = D[r];
  const double *S = d.sainv;
  const double *x = d.dx;
  double acc = 0;
  for (int i = threadIdx.x; i < d.n; i += LM_T) {
    double v = 0;
    for (int k = 0; k < d.n; k++) v += S[(size_t)i * d.n + k] * x[k];
    acc += x[i] * v;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = LM_T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) q[r] = part[0];
}

// Per-retrieval vectors to and from one packed staging buffer (one copy per
// phase instead of one per retrieval): which = 0 puts dx (n) then dy (m),
// which = 3 puts sig (m), which = 1 gets b (n), which = 2 gets step (n).
__global__ __launch_bounds__(LM_T) void k_lm_stage(const LmDesc *D, const int *list, const size_t *off,
                                                   double *stage, int which) {
  const LmDesc d = D[list[blockIdx.x]];
  double *s = stage + off[blockIdx.x];
  if (which == 0) {
    for (int i = threadIdx.x; i < d.n; i += LM_T) d.dx[i] = s[i];
    for (int i = threadIdx.x; i < d.m; i += LM_T) d.dy[i] = s[d.n + i];
  } else if (which == 3) {
    for (int i = threadIdx.x; i < d.m; i += LM_T) d.sig[i] = s[i];
  } else {
    const double *v = which == 1 ? d.b : d.step;
    for (int i = threadIdx.x; i < d.n; i += LM_T) s[i] = v[i];
  }
}

// S_a of set_cov_apr (src/retrieval.c:808-920) into d.a from packed
// per-element inputs in stage + off[b]: [sigma | z | x | y | zc | cz | ch |
// quantity], 8 n doubles (x, y, zc: geo2cart(0, lon, lat)).
__global__ __launch_bounds__(LM_T) void k_lm_cov_apr(const LmDesc *D, const int *list, const size_t *off,
                                                     const double *stage) {
  const LmDesc d = D[list[blockIdx.x]];
  const int n = d.n;
  const double *s = stage + off[blockIdx.x];
  const double *sg = s, *z = s + n, *px = s + 2 * n, *py = s + 3 * n, *pz = s + 4 * n;
  const double *cz = s + 5 * n, *ch = s + 6 * n, *q = s + 7 * n;
  for (long e = threadIdx.x + (long)blockIdx.y * LM_T; e < (long)n * n; e += (long)LM_T * gridDim.y) {
    const int i = (int)(e / n), j = (int)(e - (long)i * n);
    double v = 0;
    if (i == j) {
      v = sg[i] * sg[i];
    } else if (q[i] == q[j] && cz[i] > 0 && ch[i] > 0) {
      const double dx = px[i] - px[j], dy = py[i] - py[j], dz = pz[i] - pz[j];
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      v = sg[i] * sg[j] * exp(-dist / ch[i] - fabs(z[i] - z[j]) / cz[i]);
    }
    d.a[e] = v;
  }
}

// Retrieval covariance of the error analysis: a = C + S_a^-1 (C =
// Kc^T Kc of the last Jacobian), a = L L^T, inverse into the S_a^-1 buffer
// (matrix_invert, src/retrieval.c:373-399).  status[r] = 1 if a is not
// positive definite.
__global__ __launch_bounds__(LM_T) void k_ana_cov(const LmDesc *D, const int *list, int *status) {
  __shared__ int bad;
  const int r = list[blockIdx.x];
  const LmDesc d = D[r];
  const int n = d.n;
  if (threadIdx.x == 0) bad = 0;
  for (long e = threadIdx.x; e < (long)n * n; e += LM_T) d.a[e] = d.cmat[e] + d.sainv[e];
  __syncthreads();
  if (!chol_wg(d.a, n, &bad)) {
    if (threadIdx.x == 0) status[r] = 1;
    return;
  }
  chol_inverse_wg(d.a, n, d.sainv);
  if (threadIdx.x == 0) status[r] = 0;
}

// G[i][j] = sum_k cov[i][k] Kc[j][k] sig[j] (cov in the S_a^-1 buffer).
__global__ __launch_bounds__(LM_T) void k_ana_gain(const LmDesc *D, const int *list) {
  const LmDesc d = D[list[blockIdx.y]];
  const long tot = (long)d.n * d.m;
  for (long e = (long)blockIdx.x * LM_T + threadIdx.x; e < tot; e += (long)gridDim.x * LM_T) {
    const int i = (int)(e / d.m), j = (int)(e - (long)i * d.m);
    const double *ci = d.sainv + (size_t)i * d.n, *kj = d.k + (size_t)j * d.n;
    double v = 0;
    for (int k = 0; k < d.n; k++) v += ci[k] * kj[k];
    d.g[e] = v * d.sig[j];
  }
}

// A[i][j] = sum_k cov[i][k] C[k][j] into the work buffer.
__global__ __launch_bounds__(LM_T) void k_ana_avk(const LmDesc *D, const int *list) {
  const LmDesc d = D[list[blockIdx.y]];
  const long tot = (long)d.n * d.n;
  for (long e = (long)blockIdx.x * LM_T + threadIdx.x; e < tot; e += (long)gridDim.x * LM_T) {
    const int i = (int)(e / d.n), j = (int)(e - (long)i * d.n);
    const double *ci = d.sainv + (size_t)i * d.n;
    double v = 0;
    for (int k = 0; k < d.n; k++) v += ci[k] * d.cmat[(size_t)k * d.n + j];
    d.a[e] = v;
  }
}

// ---------------- launchers ----------------
void launch_lm_cov_apr(const LmDesc *D, const int *list, int cnt, const size_t *off, const double *stage,
                       hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_cov_apr, dim3(cnt, 16), dim3(LM_T), 0, s, D, list, off, stage);
}
void launch_lm_stage(const LmDesc *D, const int *list, int cnt, const size_t *off, double *stage, int which,
                     hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_stage, dim3(cnt), dim3(LM_T), 0, s, D, list, off, stage, which);
}
void launch_lm_gather(const LmDesc *D, const int *list, int cnt, int max_rows, const double *kpass, int nd,
                      int ncol, hipStream_t s) {
  if (cnt == 0 || max_rows == 0) return;
  hipLaunchKernelGGL(k_lm_gather, dim3(max_rows, cnt), dim3(LM_T), 0, s, D, list, kpass, nd, ncol);
}
void launch_lm_normal(const LmDesc *D, const int *list, int cnt, int max_n, hipStream_t s) {
  if (cnt == 0) return;
  const int nt = (max_n + 63) / 64;
  hipLaunchKernelGGL(k_lm_normal, dim3(nt * (nt + 1) / 2, cnt), dim3(LM_T), 0, s, D, list);
}
void launch_lm_cholinv(const LmDesc *D, const int *list, int cnt, int *status, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_cholinv, dim3(cnt), dim3(LM_T), 0, s, D, list, status);
}
void launch_lm_rhs(const LmDesc *D, const int *list, int cnt, int max_n, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_rhs, dim3((max_n + LM_T - 1) / LM_T, cnt), dim3(LM_T), 0, s, D, list);
}
void launch_lm_solve(const LmDesc *D, const int *list, int cnt, const double *lmpar, int *status, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_solve, dim3(cnt), dim3(LM_T), 0, s, D, list, lmpar, status);
}
void launch_lm_quad(const LmDesc *D, const int *list, int cnt, double *q, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lm_quad, dim3(cnt), dim3(LM_T), 0, s, D, list, q);
}

void launch_ana(const LmDesc *D, const int *list, int cnt, int *status, hipStream_t s, int phase) {
  if (cnt == 0) return;
  if (phase == 0) hipLaunchKernelGGL(k_ana_cov, dim3(cnt), dim3(LM_T), 0, s, D, list, status);
  if (phase == 1) hipLaunchKernelGGL(k_ana_gain, dim3(64, cnt), dim3(LM_T), 0, s, D, list);
  if (phase == 2) hipLaunchKernelGGL(k_ana_avk, dim3(64, cnt), dim3(LM_T), 0, s, D, list);
}

}  // namespace jur
