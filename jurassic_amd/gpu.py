"""Python bindings of libjurassic_gpu.so (the MI355X forward model).

Mirrors the reference operator interface for this path: formod(ctl, atm,
obs) (src/jurassic.h:890-893) and kernel(ctl, atm, obs, k)
(src/jurassic.h:1008-1012), on the same ctl_t/atm_t/obs_t structs, plus the
device-resident batch interface used by the benchmark.  Errors raise
JurassicError carrying the library's status code and message (the
reference prints an error and exits instead).

The library must be built (jurassic_amd/build.py); there is no CPU
fallback: if libjurassic_gpu.so is missing this module raises on import of
the library.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libjurassic_gpu.so")
_LIB = None


class JurassicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (abi.ERR_NAMES.get(code, code), msg))
        self.code = code


class jur_batch_t(C.Structure):
    _fields_ = [
        ("nprof", C.c_int), ("atm", C.c_void_p), ("nray", C.c_int),
        ("ray_prof", C.c_void_p), ("ray_geom", C.c_void_p), ("mask", C.c_void_p),
        ("rad", C.c_void_p), ("tau", C.c_void_p), ("tp", C.c_void_p),
    ]


EXPORTS = [
    "jur_gpu_last_error", "jur_gpu_abi_check", "formod_gpu", "jur_gpu_formod",
    "jur_gpu_formod_batch", "jur_gpu_kernel", "kernel_gpu", "jur_ctx_create",
    "jur_ctx_create_wide", "jur_ctx_destroy", "jur_ctx_nd", "jur_ctx_formod_device",
    "jur_ctx_last_timing", "jur_ctx_last_mean_np", "jur_state_map", "jur_ctx_jacobian_device",
    "jur_gpu_optimal_estimation", "jur_server_create", "jur_server_destroy", "jur_server_submit",
    "jur_server_wait", "jur_server_fetch", "jur_server_stats", "jur_server_profile",
    "jur_ctx_create_device_tables", "jur_gpu_optimal_estimation_ana", "jur_tblgen", "jur_tblgen_batch",
]

ANA_MATS = ("cov", "corr", "gain", "avk")
ANA_VECS = ("err_total", "err_noise", "err_formod", "cont", "res")
ANA_ATMS = ("atm_err_total", "atm_err_noise", "atm_err_formod", "atm_cont", "atm_res")


class jur_ana_t(C.Structure):
    """Error-analysis outputs of one retrieval (jurassic_gpu.h)."""
    _fields_ = [(k, C.c_void_p) for k in ANA_MATS + ANA_VECS + ANA_ATMS]


class jur_dtables_t(C.Structure):
    """Emissivity tables in device memory, tbl_t order (jurassic_gpu.h)."""
    _fields_ = [
        ("tp", C.c_int), ("tt", C.c_int), ("tu", C.c_int),
        ("np", C.c_void_p), ("nt", C.c_void_p), ("nu", C.c_void_p),
        ("p", C.c_void_p), ("t", C.c_void_p), ("u", C.c_void_p), ("eps", C.c_void_p),
        ("filt_off", C.c_void_p), ("filt_nu", C.c_void_p), ("filt_w", C.c_void_p),
    ]


class jur_jac_t(C.Structure):
    _fields_ = [
        ("nmap", C.c_int), ("xmap", C.c_void_p), ("prof_map", C.c_void_p), ("ncol", C.c_int),
        ("k", C.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise OSError("libjurassic_gpu.so is not built (run jurassic_amd/build.py)")
        L = C.CDLL(LIBPATH)
        P, I, S, D = C.c_void_p, C.c_int, C.c_size_t, C.c_double
        L.jur_gpu_last_error.restype = C.c_char_p
        L.jur_gpu_abi_check.argtypes = [S, S, S]
        L.jur_gpu_abi_check.restype = I
        L.jur_gpu_formod.argtypes = [P, P, P, I]
        L.jur_gpu_formod.restype = I
        L.jur_gpu_formod_batch.argtypes = [P, P, P, I, I]
        L.jur_gpu_formod_batch.restype = I
        L.jur_gpu_kernel.argtypes = [P, P, P, C.POINTER(D), S, S, I, I]
        L.jur_gpu_kernel.restype = I
        L.jur_ctx_create.argtypes = [P, C.POINTER(P)]
        L.jur_ctx_create.restype = I
        L.jur_ctx_create_wide.argtypes = [P, I, C.POINTER(D), C.POINTER(I), C.POINTER(P)]
        L.jur_ctx_create_wide.restype = I
        L.jur_ctx_create_device_tables.argtypes = [P, I, C.POINTER(D), C.POINTER(I), P, C.POINTER(P)]
        L.jur_ctx_create_device_tables.restype = I
        L.jur_ctx_destroy.argtypes = [P]
        L.jur_ctx_nd.argtypes = [P]
        L.jur_ctx_nd.restype = I
        L.jur_ctx_formod_device.argtypes = [P, C.POINTER(jur_batch_t), I, P]
        L.jur_ctx_formod_device.restype = I
        L.jur_ctx_formod_device_noalloc.argtypes = [P, C.POINTER(jur_batch_t), I, P]
        L.jur_ctx_formod_device_noalloc.restype = I
        L.jur_ctx_last_timing.argtypes = [P, C.POINTER(D), I]
        L.jur_ctx_last_timing.restype = I
        L.jur_ctx_last_mean_np.argtypes = [P]
        L.jur_ctx_last_mean_np.restype = D
        L.jur_state_map.argtypes = [P, P, C.POINTER(I)]
        L.jur_state_map.restype = I
        L.jur_ctx_jacobian_device.argtypes = [P, C.POINTER(jur_batch_t), C.POINTER(jur_jac_t), I, P]
        L.jur_ctx_jacobian_device.restype = I
        L.jur_gpu_optimal_estimation.argtypes = [P, P, I, P, P, P, P, C.POINTER(D), C.POINTER(I), I, I]
        L.jur_gpu_optimal_estimation.restype = I
        L.jur_gpu_optimal_estimation_ana.argtypes = [P, P, I, P, P, P, P, C.POINTER(D), C.POINTER(I), I, I, P]
        L.jur_gpu_optimal_estimation_ana.restype = I
        L.jur_server_create.argtypes = [P, P, C.POINTER(abi.server_opts_t), C.POINTER(P)]
        L.jur_server_create.restype = I
        L.jur_server_destroy.argtypes = [P]
        L.jur_server_submit.argtypes = [P, C.POINTER(abi.request_t), C.POINTER(C.c_uint64)]
        L.jur_server_submit.restype = I
        L.jur_server_wait.argtypes = [P, C.c_uint64, D, C.POINTER(abi.result_t)]
        L.jur_server_wait.restype = I
        L.jur_server_fetch.argtypes = [P, C.c_uint64, P, P, C.POINTER(D), C.POINTER(D)]
        L.jur_server_fetch.restype = I
        L.jur_server_stats.argtypes = [P, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.jur_server_stats.restype = I
        L.jur_server_profile.argtypes = [P, C.POINTER(D), I]
        L.jur_server_profile.restype = I
        L.jur_tblgen.argtypes = [D, D, C.POINTER(D), I, D, D, C.POINTER(D), C.POINTER(D), I,
                                 C.POINTER(D), C.POINTER(D), I, C.POINTER(I)]
        L.jur_tblgen.restype = I
        L.jur_tblgen_batch.argtypes = [I, C.POINTER(D), C.POINTER(D), C.POINTER(D), I, D, D, C.POINTER(D),
                                       C.POINTER(D), I, C.POINTER(D), C.POINTER(D), I, C.POINTER(I)]
        L.jur_tblgen_batch.restype = I
        if L.jur_gpu_abi_check(C.sizeof(abi.ctl_t), C.sizeof(abi.atm_t), C.sizeof(abi.obs_t)):
            raise OSError("ABI mismatch between jurassic_amd.abi and libjurassic_gpu.so")
        _LIB = L
    return _LIB


def _check(rc):
    if rc:
        raise JurassicError(rc, lib().jur_gpu_last_error().decode())


def formod(ctl, atm, obs, precision=abi.JUR_FP64):
    """formod() on the GPU: fills obs.rad/tau/tp* in place."""
    _check(lib().jur_gpu_formod(C.byref(ctl), C.byref(atm), C.byref(obs), precision))


def formod_batch(ctl, atms, obss, precision=abi.JUR_FP64):
    """n packages (ctypes arrays of atm_t and obs_t) in one device pass."""
    n = len(atms)
    _check(lib().jur_gpu_formod_batch(C.byref(ctl), C.cast(atms, C.c_void_p),
                                      C.cast(obss, C.c_void_p), n, precision))


def kernel(ctl, atm, obs, method=abi.JUR_JAC_FD, precision=abi.JUR_FP64, m=None, n=None):
    """kernel() on the GPU: returns the m x n Jacobian."""
    if m is None or n is None:
        raise ValueError("give m and n (sizes of obs2y/atm2x)")
    k = np.zeros((m, n))
    _check(lib().jur_gpu_kernel(C.byref(ctl), C.byref(atm), C.byref(obs),
                                k.ctypes.data_as(C.POINTER(C.c_double)), m, n, method, precision))
    return k


def optimal_estimation(ret, ctl, obs_meas, atm_apr, method=abi.JUR_JAC_ADJOINT,
                       precision=abi.JUR_FP64, analysis=False):
    """optimal_estimation on the GPU for one retrieval (obs_t, atm_t) or a
    list of them.  Returns (atm_i, obs_i, chisq_hist, niter), each a list
    when lists were given; with analysis=True a fifth item, the error
    analysis (dict of cov, corr, gain, avk, err_total, err_noise,
    err_formod, cont, res as numpy arrays and atm_err_total, atm_err_noise,
    atm_err_formod, atm_cont, atm_res as atm_t)."""
    single = not isinstance(obs_meas, (list, tuple))
    obss = [obs_meas] if single else list(obs_meas)
    atms = [atm_apr] if single else list(atm_apr)
    n = len(obss)
    om = (abi.obs_t * n)(*obss)
    aa = (abi.atm_t * n)(*atms)
    oi = (abi.obs_t * n)()
    ai = (abi.atm_t * n)()
    hist = np.zeros(n * (ret.conv_itmax + 1))
    nit = (C.c_int * n)()
    ana = None
    res = []
    if analysis:
        ana = (jur_ana_t * n)()
        for i in range(n):
            _, nx = state_map(ctl, atms[i])
            r, _ = abi.obs_arrays(obss[i], ctl)
            my = int(np.isfinite(r).sum())
            d = {}
            for k in ANA_MATS:
                d[k] = np.zeros((nx, my) if k == "gain" else (nx, nx))
            for k in ANA_VECS:
                d[k] = np.zeros(nx)
            for k in ANA_ATMS:
                d[k] = abi.atm_t()
            for k in ANA_MATS + ANA_VECS:
                setattr(ana[i], k, d[k].ctypes.data)
            for k in ANA_ATMS:
                setattr(ana[i], k, C.addressof(d[k]))
            res.append(d)
    _check(lib().jur_gpu_optimal_estimation_ana(
        C.byref(ret), C.byref(ctl), n, C.cast(om, C.c_void_p), C.cast(oi, C.c_void_p), C.cast(aa, C.c_void_p),
        C.cast(ai, C.c_void_p), hist.ctypes.data_as(C.POINTER(C.c_double)), nit, method, precision,
        C.cast(ana, C.c_void_p) if ana is not None else None))
    hist = hist.reshape(n, ret.conv_itmax + 1)
    out_a = [abi.copy_struct(ai[i]) for i in range(n)]
    out_o = [abi.copy_struct(oi[i]) for i in range(n)]
    out_h = [hist[i, :nit[i] + 1].copy() for i in range(n)]
    out_n = [nit[i] for i in range(n)]
    if single:
        if analysis:
            return out_a[0], out_o[0], out_h[0], out_n[0], res[0]
        return out_a[0], out_o[0], out_h[0], out_n[0]
    if analysis:
        return out_a, out_o, out_h, out_n, res
    return out_a, out_o, out_h, out_n


def tblgen(press, temp, tau, nu0, nu1, nuf, filt, maxrow=4096):
    """The emissivity rows (u, eps) of tblgen (src/tblgen.c:44-140) for one
    pressure [hPa] and temperature [K], from the optical-depth spectrum tau
    on the regular grid nu0..nu1 and the filter function (nuf, filt); the
    spectral sums run on the GPU."""
    def arr(x):
        return np.ascontiguousarray(x, dtype=np.float64)
    tau, nuf, filt = arr(tau), arr(nuf), arr(filt)
    u = np.zeros(maxrow)
    e = np.zeros(maxrow)
    nrow = C.c_int(0)
    dp = C.POINTER(C.c_double)
    _check(lib().jur_tblgen(press, temp, tau.ctypes.data_as(dp), len(tau), nu0, nu1, nuf.ctypes.data_as(dp),
                            filt.ctypes.data_as(dp), len(nuf), u.ctypes.data_as(dp), e.ctypes.data_as(dp),
                            maxrow, C.byref(nrow)))
    if nrow.value > maxrow:
        return tblgen(press, temp, tau, nu0, nu1, nuf, filt, maxrow=nrow.value)
    return u[:nrow.value], e[:nrow.value]


def tblgen_batch(press, temp, tau, nu0, nu1, nuf, filt, maxrow=1024):
    """tblgen for several spectra on one grid (tau[nspec][npts]) with their
    pressures and temperatures: a list of (u, eps) per spectrum."""
    def arr(x):
        return np.ascontiguousarray(x, dtype=np.float64)
    press, temp, tau, nuf, filt = arr(press), arr(temp), arr(tau), arr(nuf), arr(filt)
    ns, npts = tau.shape
    u = np.zeros((ns, maxrow))
    e = np.zeros((ns, maxrow))
    nrow = np.zeros(ns, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    _check(lib().jur_tblgen_batch(ns, press.ctypes.data_as(dp), temp.ctypes.data_as(dp), tau.ctypes.data_as(dp),
                                  npts, nu0, nu1, nuf.ctypes.data_as(dp), filt.ctypes.data_as(dp), len(nuf),
                                  u.ctypes.data_as(dp), e.ctypes.data_as(dp), maxrow,
                                  nrow.ctypes.data_as(C.POINTER(C.c_int))))
    if ns and nrow.max() > maxrow:
        return tblgen_batch(press, temp, tau, nu0, nu1, nuf, filt, maxrow=int(nrow.max()))
    return [(u[s, :nrow[s]], e[s, :nrow[s]]) for s in range(ns)]


def state_map(ctl, atm):
    """atm2x's state map (JUR_XMAP_LEN ints) and the state size n."""
    xm = np.zeros(abi.XMAP_LEN, dtype=np.int32)
    n = lib().jur_state_map(C.byref(ctl), C.byref(atm), xm.ctypes.data_as(C.POINTER(C.c_int)))
    if n < 0:
        _check(abi.JUR_ERR_ARG)
    return xm, n


class Context:
    """A device context (tables and source function resident in HBM)."""

    def __init__(self, ctl, nu=None, window=None):
        self.ptr = C.c_void_p()
        if nu is None:
            _check(lib().jur_ctx_create(C.byref(ctl), C.byref(self.ptr)))
        else:
            nu = np.ascontiguousarray(nu, dtype=np.float64)
            w = None if window is None else np.ascontiguousarray(window, dtype=np.int32)
            _check(lib().jur_ctx_create_wide(
                C.byref(ctl), len(nu), nu.ctypes.data_as(C.POINTER(C.c_double)),
                None if w is None else w.ctypes.data_as(C.POINTER(C.c_int)), C.byref(self.ptr)))
        self.nd = lib().jur_ctx_nd(self.ptr)

    @classmethod
    def from_device_tables(cls, ctl, nu, tables, window=None, filters=None):
        """Context whose tables come from device memory.  tables: dict of
        device tensors (torch) np [nd][ng] int32, nt [nd][ng][tp] int32,
        nu [nd][ng][tp][tt] int32, p [nd][ng][tp] float64, t [nd][ng][tp][tt]
        float64, u and eps [nd][ng][tp][tt][tu] float32.  filters: None
        (files) or (off [nd+1] int32, nu, w) host arrays."""
        self = cls.__new__(cls)
        self.ptr = C.c_void_p()
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        w = None if window is None else np.ascontiguousarray(window, dtype=np.int32)
        u = tables["u"]
        tp, tt, tu = u.shape[2], u.shape[3], u.shape[4]
        for k in ("np", "nt", "nu", "p", "t", "u", "eps"):
            assert tables[k].is_contiguous(), k
        d = jur_dtables_t(tp, tt, tu, tables["np"].data_ptr(), tables["nt"].data_ptr(),
                          tables["nu"].data_ptr(), tables["p"].data_ptr(), tables["t"].data_ptr(),
                          u.data_ptr(), tables["eps"].data_ptr(), None, None, None)
        keep = None
        if filters is not None:
            off = np.ascontiguousarray(filters[0], dtype=np.int32)
            fx = np.ascontiguousarray(filters[1], dtype=np.float64)
            fy = np.ascontiguousarray(filters[2], dtype=np.float64)
            keep = (off, fx, fy)
            d.filt_off, d.filt_nu, d.filt_w = off.ctypes.data, fx.ctypes.data, fy.ctypes.data
        _check(lib().jur_ctx_create_device_tables(
            C.byref(ctl), len(nu), nu.ctypes.data_as(C.POINTER(C.c_double)),
            None if w is None else w.ctypes.data_as(C.POINTER(C.c_int)), C.byref(d), C.byref(self.ptr)))
        del keep
        self.nd = lib().jur_ctx_nd(self.ptr)
        return self

    def formod_device(self, nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, rad_ptr, tau_ptr,
                      tp_ptr=None, mask_ptr=None, precision=abi.JUR_FP64, stream=None):
        b = jur_batch_t(nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, mask_ptr, rad_ptr,
                        tau_ptr, tp_ptr)
        _check(lib().jur_ctx_formod_device(self.ptr, C.byref(b), precision, stream))

    def formod_device_noalloc(self, nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, rad_ptr, tau_ptr,
                              tp_ptr=None, mask_ptr=None, precision=abi.JUR_FP64, stream=None):
        """formod_device with the work buffers in the library's fixed
        scratch array (no device allocation in the call)."""
        b = jur_batch_t(nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, mask_ptr, rad_ptr,
                        tau_ptr, tp_ptr)
        _check(lib().jur_ctx_formod_device_noalloc(self.ptr, C.byref(b), precision, stream))

    def jacobian_device(self, nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, rad_ptr, tau_ptr,
                        xmaps, prof_map_ptr, ncol, k_ptr, tp_ptr=None, mask_ptr=None,
                        precision=abi.JUR_FP32, stream=None):
        """Forward model and adjoint Jacobian of a device batch; xmaps is a
        host int32 array [nmap][XMAP_LEN]."""
        xmaps = np.ascontiguousarray(xmaps, dtype=np.int32).reshape(-1, abi.XMAP_LEN)
        b = jur_batch_t(nprof, atm_ptr, nray, ray_prof_ptr, ray_geom_ptr, mask_ptr, rad_ptr,
                        tau_ptr, tp_ptr)
        j = jur_jac_t(xmaps.shape[0], xmaps.ctypes.data, prof_map_ptr, ncol, k_ptr)
        _check(lib().jur_ctx_jacobian_device(self.ptr, C.byref(b), C.byref(j), precision, stream))

    def last_timing(self):
        """(ray tracing, table lookups, radiance accumulation) in ms."""
        ms = (C.c_double * 3)()
        _check(lib().jur_ctx_last_timing(self.ptr, ms, 3))
        return ms[0], ms[1], ms[2]

    def last_mean_np(self):
        return lib().jur_ctx_last_mean_np(self.ptr)

    def close(self):
        if self.ptr:
            lib().jur_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The request server (jur_server_*): forward models, Jacobians and
    retrievals submitted at any time, batched continuously on the GPU."""

    def __init__(self, ctl, ret, method=abi.JUR_JAC_ADJOINT, precision=abi.JUR_FP32, max_bytes=0,
                 max_batch=0, lanes=0, forward_lane=0, path_split=0):
        self.ptr = C.c_void_p()
        self.ret = ret
        opt = abi.server_opts_t(method, precision, max_bytes, max_batch, lanes, forward_lane, path_split)
        _check(lib().jur_server_create(C.byref(ctl), C.byref(ret), C.byref(opt), C.byref(self.ptr)))

    @staticmethod
    def request(kind, atm, obs, rett=(-999.0, -999.0), retq=None, retp=(-999.0, -999.0)):
        """A request_t; retq maps emitter index -> (zmin, zmax)."""
        r = abi.request_t()
        r.kind = kind
        r.atm = C.addressof(atm)
        r.obs = C.addressof(obs)
        r.retp_zmin, r.retp_zmax = retp
        r.rett_zmin, r.rett_zmax = rett
        for i in range(abi.NG):
            r.retq_zmin[i] = r.retq_zmax[i] = -999.0
        for i in range(abi.NW):
            r.retk_zmin[i] = r.retk_zmax[i] = -999.0
        for ig, (a, b) in (retq or {}).items():
            r.retq_zmin[ig], r.retq_zmax[ig] = a, b
        return r

    def submit(self, req):
        i = C.c_uint64()
        _check(lib().jur_server_submit(self.ptr, C.byref(req), C.byref(i)))
        return i.value

    def wait(self, rid, timeout=-1.0):
        res = abi.result_t()
        _check(lib().jur_server_wait(self.ptr, rid, timeout, C.byref(res)))
        return res

    def fetch(self, rid, res):
        """(obs, atm, K or None, chisq history or None) of a finished request;
        raises if the request failed."""
        obs, atm = abi.obs_t(), abi.atm_t()
        k = np.zeros((max(res.m, 1), max(res.n, 1))) if res.kind == abi.JUR_REQ_JACOBIAN else None
        h = np.zeros(self.ret.conv_itmax + 1) if res.kind == abi.JUR_REQ_RETRIEVAL else None
        _check(lib().jur_server_fetch(self.ptr, rid, C.byref(obs), C.byref(atm),
                                      k.ctypes.data_as(C.POINTER(C.c_double)) if k is not None else None,
                                      h.ctypes.data_as(C.POINTER(C.c_double)) if h is not None else None))
        if h is not None:
            h = h[:res.niter + 1]
        return obs, atm, k, h

    def release(self, rid):
        """Drop a finished request without copying its outputs; returns its status."""
        return lib().jur_server_fetch(self.ptr, rid, None, None, None, None)

    def stats(self):
        a, b = C.c_long(), C.c_long()
        _check(lib().jur_server_stats(self.ptr, C.byref(a), C.byref(b)))
        return a.value, b.value

    def profile(self):
        """Worker seconds: admission, Jacobian passes, forward passes, dense algebra;
        then the process's allocation seconds, allocation count and the seconds of
        the allocations that replaced a smaller buffer."""
        t = (C.c_double * 7)()
        _check(lib().jur_server_profile(self.ptr, t, 7))
        return list(t)

    def close(self):
        if self.ptr:
            lib().jur_server_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
