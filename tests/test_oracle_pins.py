"""Closed-form pins of the oracle's table interpolation, FOV, surface,
solar, cloud and continuum terms.

Each test builds a case whose answer mathematics fixes independently of the
oracle's code path, so that a wrong interpolation weight, a dropped term or a
wrong factor in jur_oracle.c fails here:

* EGA (src/jurassic.c:3799-3891) over tables that are linear in u with a
  slope affine in (p, T): the path emissivity is sum_i s(p_i, T_i) u_i, which
  holds only for LIN in T (per pressure, on its own T grid) then LIN in p.
* CGA (src/jurassic.c:3710-3795) over tables linear in u with a slope affine
  in (ln p, T): the path emissivity is s(pbar, Tbar) U with the
  Curtis-Godson means, which holds only for LIN in T then LOGX in p.
* formod_fov (src/jurassic.c:3121-3192) on radiances linear in the view
  point altitude: the weighted mean of the shape's offsets.
* Surface reflection and the solar term (src/jurassic.c:3273-3327) in an
  isothermal scene: rad = B (1 - (1 - e) tau^2) + S (1 - e) tau^2 cos(zeta)/
  cos(sza).
* sza (src/jurassic.c:5183-5219): the sub-solar latitude equals the
  declination at the equinox and the solstices, the sub-solar longitude
  follows UTC, and the antipode sees 180 deg - sza.
* Cloud extinction (src/jurassic.c:4324-4332): a gray Gaussian layer on a
  straight ray gives tau = exp(-integral k g ds) by quadrature.
* ctmh2o's foreign-broadened term and chi factor (src/jurassic.c:2761-2791),
  ctmn2/ctmo2 away from the reference temperature (:2841-2917).
"""
import ctypes as C
import math
import os
import struct

import numpy as np
import pytest

from jurassic_amd import abi, scenes

KB = 1.3806504e-23
U_STEP = 2.0 ** 58          # table knots u_i = i 2^58, i = 1..256 (u_max = 2^66)
S_UNIT = 2.0 ** -74         # slope of eps = i m / 65536 on those knots is m 2^-74


def _write_affine_table(path, pgrid, tgrids, mfun):
    """Binary table (src/jurassic.c:5857-5882) whose every column is the line
    eps = m(p, T) u 2^-74 through the origin: knots u_i = i 2^58, eps_i =
    i m / 65536, both exact in float32 for integer m <= 256."""
    i = np.arange(1, 257, dtype=np.float64)
    u = (i * U_STEP).astype(np.float32)
    assert np.all(u.astype(np.float64) == i * U_STEP)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(pgrid)))
        f.write(np.asarray(pgrid, dtype="<f8").tobytes())
        for ip, p in enumerate(pgrid):
            tg = tgrids[ip]
            f.write(struct.pack("<i", len(tg)))
            f.write(np.asarray(tg, dtype="<f8").tobytes())
            for t in tg:
                m = mfun(ip, p, t)
                assert m == int(m) and 1 <= m <= 256
                e = (i * m / 65536.0).astype(np.float32)
                assert np.all(e.astype(np.float64) == i * m / 65536.0)
                f.write(struct.pack("<i", 256))
                f.write(u.astype("<f4").tobytes())
                f.write(e.astype("<f4").tobytes())


def _alternating_tgrids(n):
    """Regular T grids (locate_reg) that differ between neighbouring
    pressures: 150..340 K at even, 155..345 K at odd indices."""
    return [150.0 + 10 * np.arange(20) + (5.0 if k % 2 else 0.0) for k in range(n)]


def _ctl(tb, nu, formod, **kw):
    base = dict(tblbase=tb, ctm_co2=0, ctm_h2o=0, ctm_n2=0, ctm_o2=0, refrac=0, formod=formod)
    base.update(kw)
    return abi.make_ctl(["CO2"], nu, **base)


def _atm(q):
    z = np.arange(81.0)
    return abi.make_atm(z, scenes.std_pressure(z), scenes.std_temperature(z), [np.full(81, q)])


def _geometries():
    return [("nadir 0", scenes.nadir_geometry(705.0, [0.0])),
            ("nadir 35", scenes.nadir_geometry(705.0, [35.0])),
            ("limb 12", scenes.limb_geometry(800.0, [12.0]))]


def _scaled_q(oracle, ctl, geom, target_e, s_mean):
    """vmr for which the path emissivity is about target_e."""
    r = oracle.trace_ray(ctl, _atm(1e-6), scenes.make_obs(geom), 0)
    return 1e-6 * target_e / (s_mean * S_UNIT * float(np.sum(r["u"][:, 0])))


# ----------------------------------------------------------------------
# EGA: LIN in T on each pressure's own grid, then LIN in p
# ----------------------------------------------------------------------

def test_ega_interpolation_weights_closed_form(oracle, tmp_path):
    """With every column eps = s u (through the origin, so the lower-boundary
    branch is the same line), each EGA corner gives eps00 = E + s_c u_seg
    (E the path emissivity so far), and bilinear interpolation of a slope
    that is affine in (p, T) is exact: E_N = sum_i s(p_i, T_i) u_i.
    The float slope of the inverse lookup (2^74/m, rounded) is the only
    inexact step: relative error <= 6e-8 per point."""
    pgrid = 5.0 + 30.0 * np.arange(40)

    def mfun(ip, p, t):
        return 60 + (t - 150) / 5 + (p - 5) / 10

    nu = [700.0]
    tb = os.path.join(str(tmp_path), "aff")
    _write_affine_table(scenes.table_name(tb, nu[0], "CO2"), pgrid, _alternating_tgrids(40), mfun)
    scenes.write_boxcar_filter("%s_%.4f.filt" % (tb, nu[0]), nu[0])
    ctl = _ctl(tb, nu, 1, raydz=5.0)
    for name, geom in _geometries():
        q = _scaled_q(oracle, ctl, geom, 0.08, 130.0)
        atm = _atm(q)
        obs = scenes.make_obs(geom)
        oracle.formod(ctl, atm, obs)
        r = oracle.trace_ray(ctl, atm, obs, 0)
        u = r["u"][:, 0]
        m = 60 + (r["t"] - 150) / 5 + (r["p"] - 5) / 10
        e_exp = float(np.sum(m * S_UNIT * u))
        # the case stays on the tables' linear part: every corner's
        # equivalent column plus the segment's stays below the last knot
        assert e_exp / (float(np.min(m)) * S_UNIT) + float(np.max(u)) < 256 * U_STEP
        assert 0.02 < e_exp < 0.2, name
        tol = 1.2e-7 * r["np"]
        assert 1 - obs.tau[0][0] == pytest.approx(e_exp, rel=tol), name


# ----------------------------------------------------------------------
# CGA: LIN in T, then LOGX in p
# ----------------------------------------------------------------------

def test_cga_interpolation_weights_closed_form(oracle, tmp_path):
    """CGA looks up eps(pbar, Tbar, U) of the whole path so far; with eps =
    s u and s affine in (ln p, T), bilinear-in-(ln p, T) interpolation is
    exact, so the path (and channel) transmittance is 1 - s(pbar, Tbar) U
    with the Curtis-Godson means of the full path."""
    r_p = (1100.0 / 0.01) ** (1 / 39)
    pgrid = 0.01 * r_p ** np.arange(40)

    def mfun(ip, p, t):
        return 60 + (t - 150) / 5 + 3 * ip

    nu = [700.0]
    tb = os.path.join(str(tmp_path), "aff")
    _write_affine_table(scenes.table_name(tb, nu[0], "CO2"), pgrid, _alternating_tgrids(40), mfun)
    scenes.write_boxcar_filter("%s_%.4f.filt" % (tb, nu[0]), nu[0])
    ctl = _ctl(tb, nu, 0)
    for name, geom in _geometries():
        q = _scaled_q(oracle, ctl, geom, 0.3, 150.0)
        atm = _atm(q)
        obs = scenes.make_obs(geom)
        oracle.formod(ctl, atm, obs)
        r = oracle.trace_ray(ctl, atm, obs, 0)
        u = r["u"][:, 0]
        U = float(np.sum(u))
        pbar = float(np.sum(u * r["p"])) / U
        tbar = float(np.sum(u * r["t"])) / U
        m = 60 + (tbar - 150) / 5 + 3 * math.log(pbar / 0.01) / math.log(r_p)
        assert U < 256 * U_STEP
        e_exp = m * S_UNIT * U
        assert 0.05 < e_exp < 0.6, name
        assert 1 - obs.tau[0][0] == pytest.approx(e_exp, rel=1e-10), name


# ----------------------------------------------------------------------
# formod_fov
# ----------------------------------------------------------------------

def _fov_file(path, dz, w):
    with open(path, "w") as f:
        f.write("# $1 = altitude offset [km]\n# $2 = weight [-]\n")
        for a, b in zip(dz, w):
            f.write("%.17g %.17g\n" % (a, b))


def test_fov_of_linear_radiance_is_mean_offset(oracle, tmp_path):
    """formod_fov interpolates the pencil-beam radiances of the rays with the
    same time linearly in view point altitude and averages them over the
    shape: for radiances a + b vpz the result is a + b (vpz + sum w dz /
    sum w), also where the shape extrapolates past the outermost rays; a
    constant is unchanged.  Rays of another time do not mix in."""
    dz = np.array([-2.0, -0.7, 0.0, 1.1, 2.5, 3.0])
    w = np.array([0.1, 0.6, 1.0, 0.8, 0.3, 0.05])
    fov = os.path.join(str(tmp_path), "fov.tab")
    _fov_file(fov, dz, w)
    ctl = abi.make_ctl(["CO2"], [700.0, 800.0], fov=fov.encode())
    n = 14
    vpz = 8.0 + 1.5 * np.arange(n) + 0.1 * np.sin(np.arange(n))
    t = np.where(np.arange(n) % 2 == 0, 0.0, 3600.0)
    obs = abi.make_obs(np.full(n, 800.0), np.zeros(n), np.zeros(n), vpz, np.zeros(n), np.zeros(n), time=t)
    coef = {(0, 0.0): (1.0, 0.25), (1, 0.0): (-3.0, 0.5), (0, 3600.0): (7.0, -0.125), (1, 3600.0): (2.0, 1.0)}
    for id_ in range(2):
        for ir in range(n):
            a, b = coef[(id_, t[ir])]
            obs.rad[id_][ir] = a + b * vpz[ir]
            obs.tau[id_][ir] = 0.5
    assert oracle.lib().orc_formod_fov(C.byref(ctl), C.byref(obs)) == 0
    dbar = float(np.sum(w * dz) / np.sum(w))
    for id_ in range(2):
        for ir in range(n):
            a, b = coef[(id_, t[ir])]
            assert obs.rad[id_][ir] == pytest.approx(a + b * (vpz[ir] + dbar), rel=1e-12, abs=1e-12)
            assert obs.tau[id_][ir] == pytest.approx(0.5, rel=1e-14)


# ----------------------------------------------------------------------
# Surface emission, reflection and the solar term
# ----------------------------------------------------------------------

def _boxcar_planck(t, nu):
    from scipy.integrate import quad
    c1, c2 = 1.19104259e-8, 1.43877506
    v, _ = quad(lambda x: c1 * x ** 3 / math.expm1(c2 * x / t), nu - 0.5, nu + 0.5, epsabs=0, epsrel=1e-13)
    return v


@pytest.mark.parametrize("view", [0.0, 30.0])
def test_isothermal_surface_reflection_and_sun(oracle, view):
    """Isothermal atmosphere (T) over a surface at T with emissivity e: the
    upwelling emission is B (1 - tau), the surface adds e B tau, the reflected
    downwelling emission adds (1 - e) B (1 - tau) tau, so
      rad = B (1 - (1 - e) tau^2)           (SFTYPE 2; B for e = 1),
      rad = B (1 - tau) + e B tau           (SFTYPE 1, no reflection),
    and SFTYPE 3 adds the sun, 6.764e-5/(2 pi) Planck(5780 K, nu) seen
    through the path twice: (1 - e) tau^2 cos(zeta)/cos(sza), zeta the
    local zenith angle of the (straight) ray at the ground."""
    ctl0, _, _ = scenes.nadir_co2()
    nu = ctl0.nu[0]
    T = 250.0
    z = np.arange(81.0)
    atm = abi.make_atm(z, scenes.std_pressure(z), np.full(81, T), [np.full(81, 4e-4)])
    atm.sfeps[0], atm.sfeps[1] = 0.55, 0.95
    sfnu = [nu - 10.0, nu + 30.0]
    e = 0.55 + (0.95 - 0.55) * 10.0 / 40.0
    geom = scenes.nadir_geometry(705.0, [view])
    b = _boxcar_planck(T, nu)
    sun = 6.764e-5 / (2 * math.pi) * 1.19104259e-8 * nu ** 3 / math.expm1(1.43877506 * nu / 5780.0)
    ro = scenes.RE + 705.0
    zeta = math.asin(ro / scenes.RE * math.sin(math.radians(view)))
    res = {}
    for sftype in (1, 2, 3):
        ctl = abi.make_ctl(["CO2"], [nu], tblbase=ctl0.tblbase.decode(), refrac=0, nsf=2, sfnu=sfnu,
                           sftype=sftype, sfsza=40.0)
        obs = scenes.make_obs(geom)
        oracle.formod(ctl, abi.copy_struct(atm), obs)
        res[sftype] = (obs.rad[0][0], obs.tau[0][0])
    tau = res[1][1]
    assert 0.05 < tau < 0.95
    assert res[1][0] == pytest.approx(b * (1 - tau) + e * b * tau, rel=2e-6)
    assert res[2][0] == pytest.approx(b * (1 - (1 - e) * tau ** 2), rel=2e-6)
    exp3 = b * (1 - (1 - e) * tau ** 2) + sun * (1 - e) * tau ** 2 * math.cos(zeta) / math.cos(math.radians(40.0))
    assert res[3][0] == pytest.approx(exp3, rel=2e-6)
    # the sun's share (~1e-3 at 668 cm^-1) is far above the tolerance
    assert res[3][0] - res[2][0] > 1e-4 * res[2][0]
    # e = 1: no reflection and no sun, a blackbody cavity
    atm1 = abi.copy_struct(atm)
    atm1.sfeps[0] = atm1.sfeps[1] = 1.0
    ctl = abi.make_ctl(["CO2"], [nu], tblbase=ctl0.tblbase.decode(), refrac=0, nsf=2, sfnu=sfnu,
                       sftype=3, sfsza=40.0)
    obs = scenes.make_obs(geom)
    oracle.formod(ctl, atm1, obs)
    assert obs.rad[0][0] == pytest.approx(b, rel=2e-6)


# seconds since 2000-01-01T00:00Z (the time base of JURASSIC)
def _t2000(day_of_year, hh, mm):
    return (day_of_year - 1) * 86400.0 + hh * 3600.0 + mm * 60.0


def test_sza_equinox_and_solstices(oracle):
    """Solar geometry of 2000 (USNO): March equinox 03-20 07:35Z (declination
    0), June solstice 06-21 01:48Z (+23.44 deg), December solstice 12-21
    13:37Z (-23.44 deg).  Over all longitudes the smallest sza at latitude
    phi is |phi - dec|; the sub-solar longitude is 15 deg per hour from
    noon UTC (within the equation of time, under 4 deg); and the antipode
    sees the complement 180 deg - sza."""
    L = oracle.lib()
    lons = np.arange(-180.0, 180.0, 0.05)
    cases = [(_t2000(80, 7, 35), 0.0), (_t2000(173, 1, 48), 23.44), (_t2000(356, 13, 37), -23.44)]
    for t, dec in cases:
        for lat in (-60.0, -23.44, 0.0, 23.44, 45.0):
            s = np.array([L.orc_sza(t, lo, lat) for lo in lons])
            assert float(s.min()) == pytest.approx(abs(lat - dec), abs=0.1), (t, lat)
        s0 = np.array([L.orc_sza(t, lo, dec) for lo in lons])
        utc = (t % 86400.0) / 3600.0
        lon_ss = 15.0 * (12.0 - utc)
        d = (float(lons[int(np.argmin(s0))]) - lon_ss + 180.0) % 360.0 - 180.0
        assert abs(d) < 4.0, (t, d)
        for lo, la in ((10.0, 20.0), (-120.0, -50.0), (77.0, 5.0)):
            a = L.orc_sza(t, lo, la)
            b = L.orc_sza(t, lo + 180.0, -la)
            assert a + b == pytest.approx(180.0, abs=1e-9)


# ----------------------------------------------------------------------
# Clouds
# ----------------------------------------------------------------------

def test_gray_cloud_layer_transmittance(oracle):
    """No gas absorption (vmr 0), continua off, straight rays: the only
    extinction is the cloud layer k(nu) exp(-((z - clz)/cldz)^2 / 2) with
    k(nu) linear between the cloud wavenumbers, so tau = exp(-int k g ds)
    and, isothermal, rad = B (1 - tau)."""
    from scipy.integrate import quad
    from scipy.special import erf
    ctl0, _, _ = scenes.nadir_co2()
    nu = ctl0.nu[0]
    T = 250.0
    z = np.arange(81.0)
    atm = abi.make_atm(z, scenes.std_pressure(z), np.full(81, T), [np.zeros(81)])
    atm.clz, atm.cldz = 10.0, 2.0
    atm.clk[0], atm.clk[1] = 0.02, 0.06
    clnu = [nu - 20.0, nu + 20.0]
    kc = 0.04
    ctl = abi.make_ctl(["CO2"], [nu], tblbase=ctl0.tblbase.decode(), refrac=0, ncl=2, clnu=clnu,
                       ctm_co2=0, ctm_h2o=0, ctm_n2=0, ctm_o2=0, sftype=0)
    b = _boxcar_planck(T, nu)
    g = lambda zz: math.exp(-0.5 * ((zz - 10.0) / 2.0) ** 2)
    # straight down: int_0^80 g dz in closed form
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0]))
    oracle.formod(ctl, abi.copy_struct(atm), obs)
    od = kc * 2.0 * math.sqrt(math.pi / 2) * (erf(70.0 / (2.0 * math.sqrt(2))) - erf(-10.0 / (2.0 * math.sqrt(2))))
    assert -math.log(obs.tau[0][0]) == pytest.approx(od, rel=1e-6)
    assert obs.rad[0][0] == pytest.approx(b * (1 - obs.tau[0][0]), rel=2e-6)
    # limb through the layer (tangent 9 km): quadrature along the chord
    obs = scenes.make_obs(scenes.limb_geometry(800.0, [9.0]))
    oracle.formod(ctl, abi.copy_struct(atm), obs)
    r0 = scenes.RE + 9.0
    smax = math.sqrt((scenes.RE + 80.0) ** 2 - r0 ** 2)
    half, _ = quad(lambda s: g(math.sqrt(r0 ** 2 + s ** 2) - scenes.RE), 0, smax, epsabs=0, epsrel=1e-12,
                   limit=400)
    assert -math.log(obs.tau[0][0]) == pytest.approx(kc * 2 * half, rel=2e-3)
    assert obs.rad[0][0] == pytest.approx(b * (1 - obs.tau[0][0]), rel=2e-6)


# ----------------------------------------------------------------------
# Continua
# ----------------------------------------------------------------------

NA, P0 = 6.02214199e23, 1013.25


def test_h2o_continuum_foreign_and_chi(oracle):
    """ctmh2o (src/jurassic.c:2761-2791): q = 0 leaves only the foreign term
    (FRN table times the 1 - 0.25 * 36100/(36100 + dv^2 + 1e-8 dv^6) line
    shape, dv = nu - 370); the result is affine in q between the q = 0 and
    q = 1 values; 820 < nu < 960 scales the self term by the chi factor
    interpolated linearly in nu over its 10 cm^-1 table; the self term
    scales geometrically between the 296 K and 260 K tables."""
    from ctm_helper import ctm_tables
    tab = ctm_tables()
    L = oracle.lib()
    p, u = 500.0, 1e22

    def pref(nu, t):
        return nu * u * math.tanh(0.7193876 / t * nu) * 296.0 / t * p / P0 * 1e-20

    for nu in (400.0, 700.0, 1000.0, 1500.0, 2500.0):  # knots of the 10 cm^-1 grid
        iw = int(nu / 10 + 1)
        dv = nu - 370.0
        fscal = 1 - 0.25 * 36100.0 / (dv ** 2 + 1e-8 * dv ** 6 + 36100.0)
        for t in (296.0, 250.0):
            f0 = L.orc_ctmh2o(nu, p, t, 0.0, u)
            assert f0 == pytest.approx(pref(nu, t) * tab["JUR_CTM_H2O_FRN"][iw - 1] * fscal, rel=1e-12)
            f1 = L.orc_ctmh2o(nu, p, t, 1.0, u)
            for q in (1e-3, 0.02, 0.5):
                assert L.orc_ctmh2o(nu, p, t, q, u) == pytest.approx(q * f1 + (1 - q) * f0, rel=1e-12)
        # geometric mean of the two self tables half way between 296 K and 260 K
        f278 = L.orc_ctmh2o(nu, p, 278.0, 1.0, u)
        c296, c260 = tab["JUR_CTM_H2O_296"][iw - 1], tab["JUR_CTM_H2O_260"][iw - 1]
        assert f278 == pytest.approx(pref(nu, 278.0) * math.sqrt(c296 * c260), rel=1e-12)
    xf = tab["JUR_CTM_H2O_XFCREV"]
    for nu in (825.0, 847.5, 900.0, 933.3, 955.0):
        xw = nu / 10 + 1
        iw = int(xw)
        dw = xw - iw
        c296 = (1 - dw) * tab["JUR_CTM_H2O_296"][iw - 1] + dw * tab["JUR_CTM_H2O_296"][iw]
        x = (nu - 820.0) / 10.0
        ix = int(x)
        chi = xf[ix] + (x - ix) * (xf[ix + 1] - xf[ix])
        got = L.orc_ctmh2o(nu, p, 296.0, 1.0, u)
        assert got == pytest.approx(pref(nu, 296.0) * chi * c296, rel=1e-12)
        assert chi > 1.0
    assert L.orc_ctmh2o(20005.0, p, 296.0, 0.5, u) == 0.0


def test_n2_o2_continua_temperature_dependence(oracle):
    """ctmn2/ctmo2 (src/jurassic.c:2841-2917) away from 296 K: density
    squared (p/P0 273/T)^2, exp(beta (1/296 - 1/T)), and for N2 the
    (q + (1 - q)(1.294 - 0.4545 T/296)) mixture factor; linear in nu between
    the table's knots."""
    from ctm_helper import ctm_tables
    tab = ctm_tables()
    L = oracle.lib()
    for t in (200.0, 250.0, 320.0):
        for p in (100.0, 800.0):
            for i in (3, 40, 70):
                for gas in ("N2", "O2"):
                    nus, bs, betas = tab["JUR_CTM_%s_NU" % gas], tab["JUR_CTM_%s_B" % gas], tab["JUR_CTM_%s_BETA" % gas]
                    nu = 0.3 * nus[i] + 0.7 * nus[i + 1]
                    w = (nu - nus[i]) / (nus[i + 1] - nus[i])
                    bb = bs[i] + w * (bs[i + 1] - bs[i])
                    be = betas[i] + w * (betas[i + 1] - betas[i])
                    q = 0.79 if gas == "N2" else 0.21
                    exp = 0.1 * (p / P0 * 273.0 / t) ** 2 * math.exp(be * (1 / 296.0 - 1 / t)) * q * bb
                    if gas == "N2":
                        exp *= q + (1 - q) * (1.294 - 0.4545 * t / 296.0)
                        got = L.orc_ctmn2(nu, p, t)
                    else:
                        got = L.orc_ctmo2(nu, p, t)
                    assert got == pytest.approx(exp, rel=1e-11), (gas, t, p, i)
