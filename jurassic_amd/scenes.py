"""Synthetic scenes for the forward model: atmospheres, geometries,
emissivity tables and filter functions.

Every file written here uses the reference's own formats so that the same
inputs drive the reference-compatible CPU path and the MI355X library:
binary emissivity tables (TBLFMT = 2, read_tbl/write_tbl,
src/jurassic.c:5037-5060 and 5857-5882), and two-column filter functions
named <tblbase>_<nu %.4f>.filt (read_shape, src/jurassic.c:4907-4946).

Workloads (sizes as named by the benchmark configurations):
  nadir_co2      81-level standard profile, CO2 only, 2 channels at 667.8 and
                 668.5 cm^-1, 10 nadir rays from 705 km at view angles 0..45 deg.
  limb_3gas      4096 profiles (standard T plus a seeded Gaussian random field,
                 sigma 3 K, exponential correlation with 5 km length), CO2, H2O
                 and O3, 12 channels 700..1400 cm^-1, observer at 800 km, 32
                 tangent heights 5..67 km.
  nadir_jac      the limb_3gas atmospheres, 8 nadir view angles 0..50 deg, 96
                 channels 650..1600 cm^-1, state = T plus three vmr profiles.
"""
import math
import os
import struct

import numpy as np

from . import abi

RE = 6367.421

GAS_SCALE = {"CO2": 1e-2, "H2O": 1e-3, "O3": 5e-3}


# ----------------------------------------------------------------------
# Atmospheres.
# ----------------------------------------------------------------------

def std_temperature(z):
    """Piecewise-linear temperature [K] at altitude z [km]:
    288 K at the surface, -6.5 K/km to 11 km, isothermal to 20 km,
    +1 K/km to 32 km, +2.8 K/km to 47 km, then -2.5 K/km."""
    z = np.asarray(z, dtype=float)
    t = np.empty_like(z)
    t11 = 288.0 - 6.5 * 11.0
    t32 = t11 + 1.0 * 12.0
    t47 = t32 + 2.8 * 15.0
    for i, zz in enumerate(z.flat):
        if zz <= 11:
            v = 288.0 - 6.5 * zz
        elif zz <= 20:
            v = t11
        elif zz <= 32:
            v = t11 + 1.0 * (zz - 20)
        elif zz <= 47:
            v = t32 + 2.8 * (zz - 32)
        else:
            v = t47 - 2.5 * (zz - 47)
        t.flat[i] = v
    return t


def std_pressure(z):
    return 1013.25 * np.exp(-np.asarray(z, dtype=float) / 7.0)


def vmr_profile(gas, z):
    z = np.asarray(z, dtype=float)
    if gas == "CO2":
        return np.full_like(z, 400e-6)
    if gas == "H2O":
        return np.maximum(1e-2 * np.exp(-z / 2.0), 4e-6)
    if gas == "O3":
        return 8e-6 * np.exp(-0.5 * ((z - 30.0) / 7.0) ** 2)
    if gas in ("CH4",):
        return np.full_like(z, 1.8e-6)
    if gas in ("N2O",):
        return np.full_like(z, 3.2e-7)
    if gas in ("CO",):
        return 1e-7 * np.exp(-z / 10.0)
    if gas in ("HNO3",):
        return 5e-9 * np.exp(-0.5 * ((z - 25.0) / 6.0) ** 2)
    if gas in ("F11", "CFC-11"):
        return np.full_like(z, 2.3e-10)
    raise ValueError(gas)


_GRF_FACTOR = {}


def _grf_factor(z, sigma, corr_len):
    key = (z.tobytes(), sigma, corr_len)
    if key not in _GRF_FACTOR:
        cov = sigma ** 2 * np.exp(-np.abs(z[:, None] - z[None, :]) / corr_len)
        _GRF_FACTOR[key] = np.linalg.cholesky(cov + 1e-12 * np.eye(len(z)))
    return _GRF_FACTOR[key]


def grf_perturbation(iprof, z, sigma=3.0, corr_len=5.0, seed=1234):
    """Seeded Gaussian random field with covariance
    sigma^2 exp(-|dz|/corr_len) on the levels z; one stream per profile."""
    z = np.asarray(z, dtype=float)
    l = _grf_factor(z, sigma, corr_len)
    rng = np.random.default_rng([seed, iprof])
    return l @ rng.standard_normal(len(z))


def make_profile_atm(gases, iprof=None, z=None, perturb=True, seed=1234):
    """atm_t on 81 levels 0..80 km; T perturbed by the random field if
    iprof is given."""
    if z is None:
        z = np.arange(81, dtype=float)
    t = std_temperature(z)
    if iprof is not None and perturb:
        t = t + grf_perturbation(iprof, z, seed=seed)
    p = std_pressure(z)
    q = np.array([vmr_profile(g, z) for g in gases]) if gases else np.zeros((0, len(z)))
    k = np.zeros((1, len(z)))
    return abi.make_atm(z, p, t, q, k)


# ----------------------------------------------------------------------
# Geometries.
# ----------------------------------------------------------------------

def nadir_geometry(obsz, angles_deg):
    """View points on the ground for nadir view angles theta from an
    observer at (0, 0): lat = asin((RE+obsz)/RE sin theta) - theta."""
    th = np.radians(np.asarray(angles_deg, dtype=float))
    lat = np.degrees(np.arcsin((RE + obsz) / RE * np.sin(th)) - th)
    n = len(th)
    return dict(obsz=np.full(n, obsz), obslon=np.zeros(n), obslat=np.zeros(n),
                vpz=np.zeros(n), vplon=np.zeros(n), vplat=lat)


def limb_geometry(obsz, tangent_heights):
    """The reference limb convention (src/limb.c:49-57):
    vplat = acos((RE+z)/(RE+obsz)) [deg], observer at (0, 0)."""
    zt = np.asarray(tangent_heights, dtype=float)
    n = len(zt)
    return dict(obsz=np.full(n, obsz), obslon=np.zeros(n), obslat=np.zeros(n),
                vpz=zt.copy(), vplon=np.zeros(n),
                vplat=180.0 / math.pi * np.arccos((RE + zt) / (RE + obsz)))


def make_obs(geom):
    return abi.make_obs(geom["obsz"], geom["obslon"], geom["obslat"],
                        geom["vpz"], geom["vplon"], geom["vplat"])


# ----------------------------------------------------------------------
# Tables and filters.
# ----------------------------------------------------------------------

def table_grids(np_=40, nt=20, nu=200, pmin=0.01, pmax=1100.0, tmin=160.0,
                tmax=340.0, umin=1e15, umax=1e27):
    p = np.exp(np.linspace(math.log(pmin), math.log(pmax), np_))
    t = np.linspace(tmin, tmax, nt)
    u = np.exp(np.linspace(math.log(umin), math.log(umax), nu))
    return p, t, u


def synthetic_eps(p, t, u, scale, u0=1e20, a=0.8):
    """eps = 1 - exp(-k (u/u0)^a), k = scale (p/1013)^0.7 (T/250)^-1.5,
    shape [np][nt][nu], stored as float32 (saturates to 1.0f)."""
    k = scale * (p[:, None] / 1013.0) ** 0.7 * (t[None, :] / 250.0) ** -1.5
    x = k[:, :, None] * (u[None, None, :] / u0) ** a
    return (-np.expm1(-x)).astype(np.float32)


def write_table_bin(path, p, t, u, eps):
    """Binary table layout of write_tbl (src/jurassic.c:5857-5882):
    int np; double p[np]; per p: int nt; double t[nt];
    per (p,T): int nu; float u[nu]; float eps[nu]."""
    u32 = np.asarray(u, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(p)))
        f.write(np.asarray(p, dtype="<f8").tobytes())
        for ip in range(len(p)):
            f.write(struct.pack("<i", len(t)))
            f.write(np.asarray(t, dtype="<f8").tobytes())
            for it in range(len(t)):
                f.write(struct.pack("<i", len(u32)))
                f.write(u32.astype("<f4").tobytes())
                f.write(np.asarray(eps[ip, it], dtype="<f4").tobytes())


def read_table_bin(path):
    """Read a binary table (layout of write_table_bin): returns (p[np],
    [t[nt] per p], [[u[nu], eps[nu]] per (p, T)])."""
    with open(path, "rb") as f:
        data = f.read()
    off = 0

    def take(fmt, n):
        nonlocal off
        a = np.frombuffer(data, dtype=fmt, count=n, offset=off)
        off += a.nbytes
        return a
    np_ = int(take("<i4", 1)[0])
    p = take("<f8", np_).copy()
    ts, cols = [], []
    for _ in range(np_):
        nt = int(take("<i4", 1)[0])
        ts.append(take("<f8", nt).copy())
        row = []
        for _ in range(nt):
            nu = int(take("<i4", 1)[0])
            row.append((take("<f4", nu).copy(), take("<f4", nu).copy()))
        cols.append(row)
    return p, ts, cols


def read_filter(path):
    """Filter function file (two columns, '#' comments) as read_shape reads it."""
    x, y = [], []
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and not line.lstrip().startswith("#"):
            try:
                x.append(float(parts[0]))
                y.append(float(parts[1]))
            except ValueError:
                pass
    return np.array(x), np.array(y)


def write_table_ascii(path, p, t, u, eps):
    """ASCII table rows "p T u eps" as write_tbl (src/jurassic.c:5836-5853)."""
    with open(path, "w") as f:
        f.write("# $1 = pressure [hPa]\n# $2 = temperature [K]\n"
                "# $3 = column density [molecules/cm^2]\n# $4 = emissivity [-]\n")
        for ip in range(len(p)):
            for it in range(len(t)):
                f.write("\n")
                for iu in range(len(u)):
                    f.write("%g %g %e %e\n" % (p[ip], t[it], np.float32(u[iu]),
                                                eps[ip, it, iu]))


def write_boxcar_filter(path, nu, width=1.0, step=0.0005):
    """Boxcar of the given width sampled at `step`, with one zero sample
    pair at each end (the shape of tests/limb_test/boxcar_792.0000.filt)."""
    half = int(round(width / 2 / step))
    n = 2 * half + 5
    with open(path, "w") as f:
        f.write("# $1 = wavenumber [cm^-1]\n# $2 = filter function\n\n")
        for i in range(n):
            off = (i - half - 2) * step
            val = 1 if 2 <= i < n - 3 else 0
            f.write("%.4f %d\n" % (nu + off, val))


def table_name(tblbase, nu, gas, fmt=2):
    return "%s_%.4f_%s.%s" % (tblbase, nu, gas, "tab" if fmt == 1 else "bin")


def write_scene_tables(dirpath, gases, nus, scales=None, tblbase_name="syn",
                       grids=None, fmt=2):
    """Write one table per (channel, gas) and one filter per channel.
    Returns the tblbase path."""
    os.makedirs(dirpath, exist_ok=True)
    tblbase = os.path.join(dirpath, tblbase_name)
    p, t, u = grids if grids is not None else table_grids()
    for ig, g in enumerate(gases):
        s = (scales or {}).get(g, GAS_SCALE.get(g, 1e-2))
        eps = synthetic_eps(p, t, u, s)
        for nu in nus:
            path = table_name(tblbase, nu, g, fmt)
            if not os.path.exists(path):
                (write_table_bin if fmt == 2 else write_table_ascii)(path + ".tmp", p, t, u, eps)
                os.replace(path + ".tmp", path)
    for nu in nus:
        path = "%s_%.4f.filt" % (tblbase, nu)
        if not os.path.exists(path):
            write_boxcar_filter(path + ".tmp", nu)
            os.replace(path + ".tmp", path)
    return tblbase


def scene_dir(name):
    root = os.environ.get("JURASSIC_SCENES", "/tmp/jurassic_scenes")
    return os.path.join(root, name)


# ----------------------------------------------------------------------
# Named workloads.
# ----------------------------------------------------------------------

def nadir_co2(ctm=True):
    """81-level standard atmosphere, CO2 only, channels 667.8 and 668.5
    cm^-1, 10 nadir rays from 705 km, view angles 0..45 deg."""
    gases, nus = ["CO2"], [667.8, 668.5]
    tb = write_scene_tables(scene_dir("nadir_co2"), gases, nus, scales={"CO2": 1e-2})
    c = 1 if ctm else 0
    ctl = abi.make_ctl(gases, nus, tblbase=tb, tblfmt=2, ctm_co2=c, ctm_h2o=c,
                       ctm_n2=c, ctm_o2=c)
    atm = make_profile_atm(gases)
    obs = make_obs(nadir_geometry(705.0, np.arange(10) * 5.0))
    return ctl, atm, obs


def limb_channels():
    return [700.0 + i * 700.0 / 11.0 for i in range(12)]


def limb_3gas_ctl(ctm=True):
    gases, nus = ["CO2", "H2O", "O3"], limb_channels()
    tb = write_scene_tables(scene_dir("limb_3gas"), gases, nus)
    c = 1 if ctm else 0
    return abi.make_ctl(gases, nus, tblbase=tb, tblfmt=2, ctm_co2=c, ctm_h2o=c,
                        ctm_n2=c, ctm_o2=c)


def limb_tangent_heights():
    return 5.0 + 2.0 * np.arange(32)


def limb_3gas(iprof, nrays=32, ctm=True):
    ctl = limb_3gas_ctl(ctm)
    atm = make_profile_atm(["CO2", "H2O", "O3"], iprof=iprof)
    obs = make_obs(limb_geometry(800.0, limb_tangent_heights()[:nrays]))
    return ctl, atm, obs


def jac_channels():
    return [650.0 + 10.0 * i for i in range(96)]


def nadir_jac_ctl(nd=96, ctm=True):
    gases, nus = ["CO2", "H2O", "O3"], jac_channels()[:nd]
    tb = write_scene_tables(scene_dir("nadir_jac"), gases, jac_channels())
    c = 1 if ctm else 0
    return abi.make_ctl(gases, nus, tblbase=tb, tblfmt=2, ctm_co2=c, ctm_h2o=c,
                        ctm_n2=c, ctm_o2=c, rett_zmin=-100, rett_zmax=100,
                        retq_zmin=[-100] * 3, retq_zmax=[100] * 3)


def nadir_jac(iprof, nd=96, nang=8, ctm=True):
    ctl = nadir_jac_ctl(nd, ctm)
    atm = make_profile_atm(["CO2", "H2O", "O3"], iprof=iprof)
    obs = make_obs(nadir_geometry(705.0, np.arange(nang) * 50.0 / 7.0))
    return ctl, atm, obs
