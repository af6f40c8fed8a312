// Host-side internals of libjurassic_gpu.so shared by its C++ sources
// (jurassic_gpu.cpp: forward model and Jacobians; retrieval.cpp:
// optimal estimation and the request server).
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <string>
#include <vector>

#include "jur_internal.hpp"

namespace jur {

// ------------------------------------------------------------------
// Errors.
// ------------------------------------------------------------------
extern thread_local char g_err[1024];
int set_err(int code, const char *fmt, ...);

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      (void)hipGetLastError(); /* the runtime keeps the error as pending too */     \
      return set_err(JUR_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_),      \
                     __FILE__, __LINE__);                                            \
    }                                                                                \
  } while (0)

// Allocation calls of the library, process-wide: how many, the time spent
// in them, and the part of it spent in calls that freed a smaller buffer
// first (a free waits for the whole device).
struct AllocStats {
  std::atomic<long> calls{0};
  std::atomic<long long> ns{0};
  std::atomic<long long> regrow_ns{0};  // of calls that freed a smaller buffer first
};
inline AllocStats g_alloc;
struct AllocTimer {
  bool regrow = false;  // set when the call freed a smaller buffer
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  ~AllocTimer() {
    const long long d = std::chrono::duration_cast<std::chrono::nanoseconds>(
                            std::chrono::steady_clock::now() - t0).count();
    g_alloc.calls++;
    g_alloc.ns += d;
    if (regrow) g_alloc.regrow_ns += d;
  }
};
// Size of a buffer of elements of elt bytes that has to grow from n to at
// least m elements: half as much again as before, at most 256 MiB more, so
// that sizes that creep upwards reallocate rarely and a large pass, which
// sizes its buffers to a share of the free HBM, still gets what it asked for.
inline size_t grow_to(size_t n, size_t m, size_t elt) {
  const size_t g = std::min(n / 2, ((size_t)256 << 20) / elt);
  return m < n + g ? n + g : m;
}

// A buffer of at most this size that grows keeps its old allocation until it
// is released instead of freeing it at once: hipFree and hipHostFree wait
// for the whole device, which would stall every stream of a serving process
// each time one of its work buffers grows.  With the growth rule above the
// kept allocations add up to less than three times the final size (twice
// for a buffer that stays under 512 MiB).  Larger buffers, which only
// large passes use, are freed as before.
constexpr size_t KEEP_ON_GROW_BYTES = (size_t)1 << 30;

// Pinned host memory that grows on demand (contents are not kept when it
// grows): the source or target of asynchronous copies, which the runtime
// then runs on the copy engines without staging them through a pageable
// bounce buffer.
template <typename T>
struct HostBuf {
  T *p = nullptr;
  size_t n = 0;
  std::vector<T *> kept;  // smaller allocations replaced by growth
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    for (T *q : kept) (void)hipHostFree(q);
    kept.clear();
    p = nullptr;
    n = 0;
  }
  int reserve(size_t m) {
    if (m <= n && p) return JUR_OK;
    AllocTimer t;
    m = grow_to(n, m, sizeof(T));
    if (p && n * sizeof(T) <= KEEP_ON_GROW_BYTES) {
      kept.push_back(p);
      p = nullptr;
      n = 0;
    } else {
      t.regrow = p != nullptr;
      release();
    }
    if (m == 0) m = 1;
    if (hipHostMalloc((void **)&p, m * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return set_err(JUR_ERR_MEM, "pinned host allocation of %zu bytes failed", m * sizeof(T));
    }
    n = m;
    return JUR_OK;
  }
  T *data() { return p; }
  T &operator[](size_t i) { return p[i]; }
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  std::vector<T *> kept;  // smaller allocations replaced by growth
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    for (T *q : kept) (void)hipFree(q);
    kept.clear();
    p = nullptr;
    n = 0;
  }
  int reserve(size_t m) {
    if (m <= n && p) return JUR_OK;
    AllocTimer t;
    m = grow_to(n, m, sizeof(T));
    if (p && n * sizeof(T) <= KEEP_ON_GROW_BYTES) {
      kept.push_back(p);
      p = nullptr;
      n = 0;
    } else {
      t.regrow = p != nullptr;
      release();
    }
    if (m == 0) m = 1;
    if (hipMalloc((void **)&p, m * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return set_err(JUR_ERR_MEM, "device allocation of %zu bytes failed", m * sizeof(T));
    }
    n = m;
    return JUR_OK;
  }
  int upload(const T *h, size_t m) {
    int rc = reserve(m);
    if (rc) return rc;
    if (m) HIPCHK(hipMemcpy(p, h, m * sizeof(T), hipMemcpyHostToDevice));
    return JUR_OK;
  }
  // Ordered on stream s, not waited for: h must stay valid until the
  // stream is next synchronized.
  int upload_async(const T *h, size_t m, hipStream_t s) {
    int rc = reserve(m);
    if (rc) return rc;
    if (m) HIPCHK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
    return JUR_OK;
  }
  // Ordered on stream s (complete on return).
  int upload(const T *h, size_t m, hipStream_t s) {
    int rc = reserve(m);
    if (rc) return rc;
    if (m) {
      HIPCHK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    return JUR_OK;
  }
};

// Device work memory of one pass: the context's grow-only buffers (base ==
// nullptr), or 256-byte aligned slices of a fixed scratch array (the
// no-allocation entry point), handed out in order and released together by
// resetting `used` to an earlier value.
struct PassMem {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  template <typename T>
  int get(DevBuf<T> &b, size_t n, T **out) {
    if (!base) {
      int rc = b.reserve(n);
      if (rc) return rc;
      *out = b.p;
      return JUR_OK;
    }
    const size_t bytes = ((n ? n : 1) * sizeof(T) + 255) & ~(size_t)255;
    if (bytes > cap - used)
      return set_err(JUR_ERR_MEM,
                     "the pass needs more than the %zu-byte scratch array of the no-allocation entry point "
                     "(%zu bytes more at this step): use fewer rays per call or jur_ctx_formod_device",
                     cap, bytes - (cap - used));
    *out = (T *)(base + used);
    used += bytes;
    return JUR_OK;
  }
};

// ------------------------------------------------------------------
// Settings copied out of ctl_t (plus a channel list of any length).
// ------------------------------------------------------------------
struct Setup {
  int ng = 0, nd = 0, nw = 1, ncl = 0, nsf = 0, sftype = 2, tblfmt = 1, formod = 1;
  int ctm_co2 = 1, ctm_h2o = 1, ctm_n2 = 1, ctm_o2 = 1, refrac = 1, write_bbt = 0;
  double sfsza = -999, hydz = -999, rayds = 10, raydz = 0.1;
  std::vector<std::string> emitter;
  std::vector<double> nu;
  std::vector<int> window;
  std::vector<double> clnu, sfnu;
  std::string tblbase, fov;

  std::string key() const {
    std::string k;
    char b[64];
    k += tblbase + "|" + fov + "|";
    for (auto &e : emitter) k += e + ",";
    k += "|";
    for (size_t i = 0; i < nu.size(); i++) {
      snprintf(b, sizeof b, "%.17g:%d,", nu[i], window[i]);
      k += b;
    }
    for (double v : clnu) { snprintf(b, sizeof b, "c%.17g,", v); k += b; }
    for (double v : sfnu) { snprintf(b, sizeof b, "s%.17g,", v); k += b; }
    snprintf(b, sizeof b, "|%d %d %d %d %d %d %d %d", ng, nd, nw, ncl, nsf, sftype, tblfmt, formod);
    k += b;
    snprintf(b, sizeof b, "|%d%d%d%d %d %d", ctm_co2, ctm_h2o, ctm_n2, ctm_o2, refrac, write_bbt);
    k += b;
    snprintf(b, sizeof b, "|%.17g %.17g %.17g %.17g", sfsza, hydz, rayds, raydz);
    k += b;
    return k;
  }
};

}  // namespace jur

// ------------------------------------------------------------------
// The device context.
// ------------------------------------------------------------------
struct jur_ctx {
  jur::Setup s;
  int device = 0;
  // tables
  jur::DevBuf<int> d_np, d_nt, d_nu, d_mono, d_pdir;
  jur::DevBuf<double> d_p, d_t;
  jur::DevBuf<float4> d_rec;
  int TP = 2, TT = 2, TU = 2;
  int shared = 0, ref_tix = 0, all_mono = 1;
  size_t table_bytes = 0;
  // model constants
  jur::DevBuf<double> d_st, d_sr, d_clnu, d_sfnu, d_ctm;
  jur::DevBuf<jur::ChanInfo> d_chan;
  jur::CtmView ctm{};
  // FOV shape
  std::vector<double> fov_dz, fov_w;
  jur::DevBuf<double> d_fov_dz, d_fov_w;
  // work buffers
  jur::DevBuf<double> d_los;
  jur::DevBuf<float> d_losf;  // FP32 forward pass: LOS fields rounded to float
  jur::DevBuf<char> d_eps;
  jur::DevBuf<float4> d_part;  // FP32 EGA Jacobian: per-step lookup partials
  jur::DevBuf<int> d_npb;
  jur::DevBuf<size_t> d_eoff;
  jur::DevBuf<jur::RayInfo> d_info;
  jur::DevBuf<char> d_tail;  // per-ray state from the march to the ray finish
  jur::DevBuf<double> d_zr;
  jur::DevBuf<unsigned long long> d_red;
  jur::DevBuf<double> d_rad0, d_tau0;
  jur::DevBuf<int> d_err;
  // Jacobian work buffers
  jur::DevBuf<double> d_kt, d_scr, d_pco;
  jur::DevBuf<size_t> d_soff;
  jur::DevBuf<int> d_xmap, d_nef;
  size_t los_budget = 0;  // bytes for LOS + emissivity buffers (0: from free HBM)
  int pass_kind = -1;     // precision, Jacobian and path flags of the last pass
                          // that sized the work buffers from free HBM
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_ms[3] = {0, 0, 0};
  double last_mean_np = 0;

  ~jur_ctx() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace jur {

struct FovPkg {
  const int *lo = nullptr, *hi = nullptr;  // device, per ray
};

// Jacobian request of a pass: state maps (device), K (device).
struct JacArgs {
  const int *xmap = nullptr;
  int nmap = 0;
  const int *prof_map = nullptr;
  int ncol = 0;
  double *k = nullptr;
};

// One pass over a device batch: the forward model, or with jac the forward
// model and its Jacobian (adjoint sweep).
// mem: where the pass's work buffers live (nullptr: the context's buffers).
int forward_pass(jur_ctx *c, const BatchView &B, int precision, const unsigned char *d_mask, double *d_rad,
                 double *d_tau, double *d_tp, const FovPkg *fov, hipStream_t st, const JacArgs *jac = nullptr,
                 PassMem *mem = nullptr);

// The cached device context of a ctl (tables read on first use).
int ctx_for(const ctl_t *ctl, jur_ctx **out);

// State and measurement vectors (atm2x, x2atm, obs2y: src/jurassic.c:29-104,
// 5896-5946, 4174-4197), the state map of the device Jacobian, kernel()'s
// step sizes (src/jurassic.c:4046-4068).
enum { IDXP = 0, IDXT = 1 };
size_t atm2x_h(const ctl_t *ctl, const atm_t *atm, double *x, int *iqa, int *ipa);
void x2atm_h(const ctl_t *ctl, const double *x, atm_t *atm);
size_t obs2y_h(const ctl_t *ctl, const obs_t *obs, double *y);
size_t state_map_h(const ctl_t *ctl, const atm_t *atm, int *xmap);
int fd_step(const ctl_t *ctl, int iq, double x0, double *h);
int check_jac_ctl(const jur_ctx *c);

}  // namespace jur
