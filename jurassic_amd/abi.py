"""ctypes mirror of the C-ABI data model (include/jurassic_gpu.h).

ctl_t, atm_t and obs_t have the field order and capacities of the
reference header (src/jurassic.h:488-538, 541-675, 734-775).  The helpers
here fill them from Python values and read the reference's text formats
(atm/obs files, src/jurassic.c:4442-4545, 4697-4794) so that tests and the
benchmark can drive the library exactly as a C caller would.
"""
import ctypes as C
import math
import os

import numpy as np

NCL, ND, NG, NP, NR, NSF, NW, LEN = 8, 128, 8, 256, 256, 8, 4, 10000
NLOS, TBLNS = 4096, 1200
TBLNP, TBLNT, TBLNU = 41, 30, 320

JUR_OK = 0
JUR_FP64 = 0
JUR_FP32 = 1
JUR_JAC_FD = 0
JUR_JAC_ADJOINT = 1
JUR_ERR_ARG = 1
JUR_ERR_IO = 2
JUR_ERR_MEM = 7

# State map layout of the device Jacobian (JUR_XMAP_* in the C header).
XMAP_NQ = 2 + NG + NW
XMAP_CLZ = XMAP_NQ * NP
XMAP_CLDZ = XMAP_CLZ + 1
XMAP_SFZ = XMAP_CLZ + 2 + NCL
XMAP_SFP = XMAP_SFZ + 1
XMAP_SFT = XMAP_SFZ + 2
XMAP_LEN = XMAP_SFZ + 3 + NSF


def xmap_lev(iq, ip):
    """Slot of level quantity iq (0 p, 1 T, 2+ig q, 2+NG+iw k) at level ip."""
    return iq * NP + ip


def xmap_clk(icl):
    return XMAP_CLZ + 2 + icl


def xmap_sfeps(isf):
    return XMAP_SFZ + 3 + isf

class ret_t(C.Structure):
    """Retrieval settings (jur_ret_t; the part of the reference's ret_t,
    src/retrieval.c:31-113, that optimal_estimation reads)."""
    _fields_ = [
        ("kernel_recomp", C.c_int), ("conv_itmax", C.c_int), ("conv_dmin", C.c_double),
        ("err_formod", C.c_double * ND), ("err_noise", C.c_double * ND),
        ("err_press", C.c_double), ("err_press_cz", C.c_double), ("err_press_ch", C.c_double),
        ("err_temp", C.c_double), ("err_temp_cz", C.c_double), ("err_temp_ch", C.c_double),
        ("err_q", C.c_double * NG), ("err_q_cz", C.c_double * NG), ("err_q_ch", C.c_double * NG),
        ("err_k", C.c_double * NW), ("err_k_cz", C.c_double * NW), ("err_k_ch", C.c_double * NW),
        ("err_clz", C.c_double), ("err_cldz", C.c_double), ("err_clk", C.c_double * NCL),
        ("err_sfz", C.c_double), ("err_sfp", C.c_double), ("err_sft", C.c_double),
        ("err_sfeps", C.c_double * NSF),
    ]


JUR_REQ_FORWARD, JUR_REQ_JACOBIAN, JUR_REQ_RETRIEVAL = 0, 1, 2
JUR_ERR_TIMEOUT = 9


class server_opts_t(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("max_bytes", C.c_size_t),
                ("max_batch", C.c_int), ("lanes", C.c_int), ("forward_lane", C.c_int),
                ("path_split", C.c_int)]


class request_t(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("atm", C.c_void_p), ("obs", C.c_void_p),
        ("retp_zmin", C.c_double), ("retp_zmax", C.c_double),
        ("rett_zmin", C.c_double), ("rett_zmax", C.c_double),
        ("retq_zmin", C.c_double * NG), ("retq_zmax", C.c_double * NG),
        ("retk_zmin", C.c_double * NW), ("retk_zmax", C.c_double * NW),
    ]


class result_t(C.Structure):
    _fields_ = [
        ("id", C.c_uint64), ("kind", C.c_int), ("status", C.c_int), ("message", C.c_char * 256),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
        ("n", C.c_int), ("m", C.c_int), ("niter", C.c_int),
    ]


ERR_NAMES = {
    0: "JUR_OK", 1: "JUR_ERR_ARG", 2: "JUR_ERR_IO", 3: "JUR_ERR_LOS",
    4: "JUR_ERR_OBSERVER", 5: "JUR_ERR_HIP", 6: "JUR_ERR_UNSUPPORTED",
    7: "JUR_ERR_MEM", 8: "JUR_ERR_FOV", 9: "JUR_ERR_TIMEOUT",
}


class atm_t(C.Structure):
    _fields_ = [
        ("np", C.c_int),
        ("time", C.c_double * NP), ("z", C.c_double * NP),
        ("lon", C.c_double * NP), ("lat", C.c_double * NP),
        ("p", C.c_double * NP), ("t", C.c_double * NP),
        ("q", (C.c_double * NP) * NG), ("k", (C.c_double * NP) * NW),
        ("clz", C.c_double), ("cldz", C.c_double), ("clk", C.c_double * NCL),
        ("sfz", C.c_double), ("sfp", C.c_double), ("sft", C.c_double),
        ("sfeps", C.c_double * NSF),
    ]


class ctl_t(C.Structure):
    _fields_ = [
        ("ng", C.c_int), ("emitter", (C.c_char * LEN) * NG),
        ("nd", C.c_int), ("nu", C.c_double * ND),
        ("nw", C.c_int), ("window", C.c_int * ND),
        ("ncl", C.c_int), ("clnu", C.c_double * NCL),
        ("nsf", C.c_int), ("sfnu", C.c_double * NSF),
        ("sftype", C.c_int), ("sfsza", C.c_double),
        ("tblbase", C.c_char * LEN), ("tblfmt", C.c_int),
        ("hydz", C.c_double),
        ("ctm_co2", C.c_int), ("ctm_h2o", C.c_int), ("ctm_n2", C.c_int), ("ctm_o2", C.c_int),
        ("refrac", C.c_int), ("rayds", C.c_double), ("raydz", C.c_double),
        ("fov", C.c_char * LEN),
        ("retp_zmin", C.c_double), ("retp_zmax", C.c_double),
        ("rett_zmin", C.c_double), ("rett_zmax", C.c_double),
        ("retq_zmin", C.c_double * NG), ("retq_zmax", C.c_double * NG),
        ("retk_zmin", C.c_double * NW), ("retk_zmax", C.c_double * NW),
        ("ret_clz", C.c_int), ("ret_cldz", C.c_int), ("ret_clk", C.c_int),
        ("ret_sfz", C.c_int), ("ret_sfp", C.c_int), ("ret_sft", C.c_int),
        ("ret_sfeps", C.c_int),
        ("write_bbt", C.c_int), ("write_matrix", C.c_int), ("formod", C.c_int),
        ("rfmbin", C.c_char * LEN), ("rfmhit", C.c_char * LEN),
        ("rfmxsc", (C.c_char * LEN) * NG),
    ]


class obs_t(C.Structure):
    _fields_ = [
        ("nr", C.c_int),
        ("time", C.c_double * NR), ("obsz", C.c_double * NR),
        ("obslon", C.c_double * NR), ("obslat", C.c_double * NR),
        ("vpz", C.c_double * NR), ("vplon", C.c_double * NR),
        ("vplat", C.c_double * NR), ("tpz", C.c_double * NR),
        ("tplon", C.c_double * NR), ("tplat", C.c_double * NR),
        ("tau", (C.c_double * NR) * ND), ("rad", (C.c_double * NR) * ND),
    ]


def np_view(arr):
    """numpy view (no copy) of a ctypes array field."""
    return np.ctypeslib.as_array(arr)


def make_ctl(emitters, nu, tblbase="-", tblfmt=2, **kw):
    """ctl_t with the defaults of read_ctl (src/jurassic.c:4547-4653)."""
    c = ctl_t()
    c.ng = len(emitters)
    for i, e in enumerate(emitters):
        c.emitter[i].value = e.encode()
    c.nd = len(nu)
    for i, v in enumerate(nu):
        c.nu[i] = float(v)
    c.nw = 1
    c.ncl = 0
    c.nsf = 0
    c.sftype = 2
    c.sfsza = -999
    c.tblbase = tblbase.encode()
    c.tblfmt = tblfmt
    c.hydz = -999
    c.ctm_co2 = c.ctm_h2o = c.ctm_n2 = c.ctm_o2 = 1
    c.refrac = 1
    c.rayds = 10
    c.raydz = 0.1
    c.fov = b"-"
    for f in ("retp_zmin", "retp_zmax", "rett_zmin", "rett_zmax"):
        setattr(c, f, -999)
    for i in range(NG):
        c.retq_zmin[i] = -999
        c.retq_zmax[i] = -999
    for i in range(NW):
        c.retk_zmin[i] = -999
        c.retk_zmax[i] = -999
    c.write_bbt = 0
    c.write_matrix = 0
    c.formod = 1
    c.rfmbin = b"-"
    c.rfmhit = b"-"
    for k, v in kw.items():
        if k in ("emitter",):
            raise ValueError(k)
        fld = getattr(c, k)
        if isinstance(fld, C.Array):
            for i, x in enumerate(v):
                fld[i] = x
        else:
            setattr(c, k, v)
    return c


def make_atm(z, p, t, q, k=None, lon=0.0, lat=0.0, time=0.0):
    a = atm_t()
    n = len(z)
    if n > NP:
        raise ValueError("too many levels")
    a.np = n
    av = np_view
    av(a.z)[:n] = z
    av(a.p)[:n] = p
    av(a.t)[:n] = t
    av(a.lon)[:n] = lon
    av(a.lat)[:n] = lat
    av(a.time)[:n] = time
    q = np.atleast_2d(q)
    for ig in range(q.shape[0]):
        av(a.q[ig])[:n] = q[ig]
    if k is not None:
        k = np.atleast_2d(k)
        for iw in range(k.shape[0]):
            av(a.k[iw])[:n] = k[iw]
    for i in range(NSF):
        a.sfeps[i] = 1.0
    return a


def make_obs(obsz, obslon, obslat, vpz, vplon, vplat, time=None):
    o = obs_t()
    n = len(obsz)
    if n > NR:
        raise ValueError("too many rays")
    o.nr = n
    av = np_view
    av(o.time)[:n] = 0.0 if time is None else time
    av(o.obsz)[:n] = obsz
    av(o.obslon)[:n] = obslon
    av(o.obslat)[:n] = obslat
    av(o.vpz)[:n] = vpz
    av(o.vplon)[:n] = vplon
    av(o.vplat)[:n] = vplat
    return o


def _data_rows(path):
    rows = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if not s or s[0].startswith("#"):
                continue
            try:
                rows.append([float(x) for x in s])
            except ValueError:
                continue
    return rows


def read_atm(path, ctl):
    """read_atm (src/jurassic.c:4442-4545) for files without clouds/surface."""
    rows = _data_rows(path)
    a = atm_t()
    a.np = len(rows)
    for i, r in enumerate(rows):
        a.time[i], a.z[i], a.lon[i], a.lat[i], a.p[i], a.t[i] = r[:6]
        for ig in range(ctl.ng):
            a.q[ig][i] = r[6 + ig]
        for iw in range(ctl.nw):
            a.k[iw][i] = r[6 + ctl.ng + iw]
    for i in range(NSF):
        a.sfeps[i] = 1.0
    return a


def read_obs(path, ctl):
    """read_obs (src/jurassic.c:4697-4794)."""
    rows = _data_rows(path)
    o = obs_t()
    o.nr = len(rows)
    for i, r in enumerate(rows):
        (o.time[i], o.obsz[i], o.obslon[i], o.obslat[i], o.vpz[i], o.vplon[i],
         o.vplat[i], o.tpz[i], o.tplon[i], o.tplat[i]) = r[:10]
        for id_ in range(ctl.nd):
            o.rad[id_][i] = r[10 + id_]
            o.tau[id_][i] = r[10 + ctl.nd + id_]
    return o


def obs_arrays(obs, ctl):
    """(rad[nd][nr], tau[nd][nr]) as numpy copies."""
    nr, nd = obs.nr, ctl.nd
    rad = np.array([np_view(obs.rad[i])[:nr] for i in range(nd)])
    tau = np.array([np_view(obs.tau[i])[:nr] for i in range(nd)])
    return rad, tau


def copy_struct(s):
    d = type(s)()
    C.memmove(C.byref(d), C.byref(s), C.sizeof(s))
    return d


def fmt_g(x):
    """C printf %g of a double."""
    return "%g" % x
