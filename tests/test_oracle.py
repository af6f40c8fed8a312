"""Pins of the CPU FP64 oracle.

The oracle is the checker for the MI355X library, so it is itself checked
here against (1) the reference's own golden data, tests/{limb,nadir}_test
rad.org tangent-point columns, and (2) properties that mathematics fixes
(brute force, closed forms, invariants), chosen so that a dropped term, a
wrong sign or index would fail one of them.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from jurassic_amd import abi, scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden")

LIMB_EMITTERS = ["CO2", "H2O", "O3", "F11", "CCl4"]
LIMB_NU = [792.0, 832.0]
NADIR_EMITTERS = ["CO2"]
NADIR_NU = [667.782, 668.541, 669.811]


def golden_case(name):
    if name == "limb_test":
        ctl = abi.make_ctl(LIMB_EMITTERS, LIMB_NU)
    else:
        ctl = abi.make_ctl(NADIR_EMITTERS, NADIR_NU, write_bbt=1)
    atm = abi.read_atm(os.path.join(GOLD, name, "atm.tab"), ctl)
    obs = abi.read_obs(os.path.join(GOLD, name, "obs.tab"), ctl)
    ref = abi._data_rows(os.path.join(GOLD, name, "rad.org"))
    return ctl, atm, obs, ref


@pytest.mark.parametrize("name", ["limb_test", "nadir_test"])
def test_tangent_points_match_reference_golden(oracle, name):
    """rad.org columns 8-10 (src/jurassic.c:5722-5729 writes them with %g)."""
    ctl, atm, obs, ref = golden_case(name)
    assert obs.nr == len(ref)
    for ir in range(obs.nr):
        oracle.trace_ray(ctl, atm, obs, ir)
        got = ["%g" % obs.tpz[ir], "%g" % obs.tplon[ir], "%g" % obs.tplat[ir]]
        exp = ["%g" % ref[ir][7], "%g" % ref[ir][8], "%g" % ref[ir][9]]
        assert got == exp, (name, ir)


def test_locate_matches_brute_force(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for n in (2, 3, 7, 40, 81):
        xa = np.sort(rng.uniform(-5, 5, n))
        for order in (1, -1):
            xx = np.ascontiguousarray(xa[::order])
            xs = np.concatenate([rng.uniform(-6, 6, 50), xx])
            for x in xs:
                i = L.orc_locate_irr(xx.ctypes.data_as(C.POINTER(C.c_double)), n, x)
                if order == 1:
                    k = int(np.sum(xx <= x))
                else:
                    k = int(np.sum(xx > x))
                assert i == min(max(k - 1, 0), n - 2)
        f = xa.astype(np.float32)
        f = np.ascontiguousarray(np.sort(f))
        for x in np.concatenate([rng.uniform(-6, 6, 50), f.astype(float)]):
            i = L.orc_locate_tbl(f.ctypes.data_as(C.POINTER(C.c_float)), n, x)
            k = int(np.sum(f.astype(float) <= x))
            assert i == min(max(k - 1, 0), n - 2)
    g = np.linspace(160, 340, 20)
    for x in (100.0, 160.0, 170.0, 339.9, 340.0, 400.0):
        i = L.orc_locate_reg(g.ctypes.data_as(C.POINTER(C.c_double)), 20, x)
        assert i == min(max(int(math.floor((x - 160) / (180 / 19))), 0), 18)


def _col_eps_ref(u, e, x):
    """Brute-force scan.  The reference's LIN() on float table entries
    forms the slope (y1-y0)/(x1-x0) in float arithmetic (C promotion rules)
    and multiplies it by the double (x-x0); the same is done here."""
    f32 = np.float32
    if x < float(u[0]):
        return float(f32(e[0]) / f32(u[0])) * x
    if x > float(u[-1]):
        d = float(f32(1) - f32(e[-1]))
        if d == 0.0:  # saturated column: log(0) = -inf, 1 - exp(-inf) = 1
            return 1.0
        a = math.log(d) / float(u[-1])
        return 1 - math.exp(a * x)
    for i in range(len(u) - 1):
        if float(u[i]) <= x <= float(u[i + 1]):
            s = (f32(e[i + 1]) - f32(e[i])) / (f32(u[i + 1]) - f32(u[i]))
            return float(e[i]) + float(s) * (x - float(u[i]))
    raise AssertionError


def test_table_column_lookups(oracle):
    L = oracle.lib()
    fp = C.POINTER(C.c_float)
    p, t, u = scenes.table_grids()
    eps = scenes.synthetic_eps(p, t, u, 1e-2)
    uf = u.astype(np.float32)
    rng = np.random.default_rng(2)
    for ip, it in ((0, 0), (20, 10), (39, 19)):
        col = np.ascontiguousarray(eps[ip, it])
        n = len(uf)
        xs = np.exp(rng.uniform(math.log(1e13), math.log(1e28), 200))
        for x in xs:
            got = L.orc_intpol_col_eps(uf.ctypes.data_as(fp), col.ctypes.data_as(fp), n, x)
            ref = _col_eps_ref(uf, col, x)
            assert got == pytest.approx(ref, rel=1e-12, abs=1e-300)
        # inverse lookup then forward lookup returns the argument
        # (piecewise-linear inverse) wherever the column is strictly increasing
        strict = np.nonzero(np.diff(col.astype(float)) > 0)[0]
        lo, hi = float(col[0]), float(col[strict[-1] + 1]) if len(strict) else float(col[0])
        for y in rng.uniform(lo, hi, 100):
            uu = L.orc_intpol_col_u(uf.ctypes.data_as(fp), col.ctypes.data_as(fp), n, y)
            back = L.orc_intpol_col_eps(uf.ctypes.data_as(fp), col.ctypes.data_as(fp), n, uu)
            # float slopes make the two directions reciprocal only to float precision
            assert back == pytest.approx(y, rel=1e-7)
        # below the first entry both branches are linear through the origin
        y = float(col[0]) * 0.25
        uu = L.orc_intpol_col_u(uf.ctypes.data_as(fp), col.ctypes.data_as(fp), n, y)
        assert uu == pytest.approx(float(np.float32(uf[0]) / np.float32(col[0])) * y, rel=1e-15)
        assert uu == pytest.approx(0.25 * float(uf[0]), rel=1e-7)


def test_planck_bright_roundtrip(oracle):
    L = oracle.lib()
    for nu in (650.0, 700.0, 1400.0, 2700.0):
        for t in (150.0, 220.0, 288.0, 340.0):
            assert L.orc_bright(L.orc_planck(t, nu), nu) == pytest.approx(t, rel=1e-13)


def _write_linear_tables(dirpath, gases, nus, scale_exp=64):
    """Table eps(u) = u / 2^scale_exp on knots u_i = i 2^(scale_exp-8),
    i = 1..256: exact in float32, so interpolation is exact and both the
    lower-boundary branch and the knots lie on the same line."""
    os.makedirs(dirpath, exist_ok=True)
    tb = os.path.join(dirpath, "lin")
    p = np.exp(np.linspace(math.log(0.01), math.log(1100), 40))
    t = np.linspace(160, 340, 20)
    i = np.arange(1, 257, dtype=np.float64)
    u = i * 2.0 ** (scale_exp - 8)
    e = (i / 256.0).astype(np.float32)
    eps = np.broadcast_to(e, (40, 20, 256)).copy()
    for g in gases:
        for nu in nus:
            scenes.write_table_bin(scenes.table_name(tb, nu, g), p, t, u, eps)
    for nu in nus:
        scenes.write_boxcar_filter("%s_%.4f.filt" % (tb, nu), nu)
    return tb


def _boxcar_planck(t, nu):
    from scipy.integrate import quad
    c1, c2 = 1.19104259e-8, 1.43877506
    v, _ = quad(lambda x: c1 * x ** 3 / math.expm1(c2 * x / t), nu - 0.5, nu + 0.5,
                epsabs=0, epsrel=1e-13)
    return v


def test_isothermal_ega_closed_form(oracle, tmp_path):
    """Isothermal atmosphere, exponential pressure, constant vmr, linear
    table, no continua:
      limb:  tau = 1 - U/2^64 and rad = B(T) (1 - tau), U the slant column;
      nadir: rad = B(T) exactly (blackbody surface at the same T)."""
    nu = [700.0]
    tb = _write_linear_tables(str(tmp_path), ["CO2"], nu)
    ctl = abi.make_ctl(["CO2"], nu, tblbase=tb, ctm_co2=0, ctm_h2o=0, ctm_n2=0,
                       ctm_o2=0, refrac=0)
    z = np.arange(81.0)
    T = 250.0
    q = 2e-8
    atm = abi.make_atm(z, 1013.25 * np.exp(-z / 7.0), np.full(81, T), [np.full(81, q)])
    b = _boxcar_planck(T, nu[0])
    # nadir, straight down
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0]))
    oracle.formod(ctl, atm, obs)
    U = 10 * q * 1013.25 * 7.0 * (1 - math.exp(-80 / 7.0)) / (1.3806504e-23 * T)
    assert obs.tau[0][0] == pytest.approx(1 - U / 2.0 ** 64, rel=2e-5)
    assert obs.rad[0][0] == pytest.approx(b, rel=2e-6)
    # limb, 20 km tangent height: rad = B (1 - tau)
    obs = scenes.make_obs(scenes.limb_geometry(800.0, [20.0]))
    oracle.formod(ctl, atm, obs)
    tau = obs.tau[0][0]
    assert 0.05 < tau < 0.95
    assert obs.rad[0][0] == pytest.approx(b * (1 - tau), rel=2e-6)
    # limb slant column by independent quadrature of the straight line
    r0, zt = scenes.RE + 20.0, 20.0
    smax = math.sqrt((scenes.RE + 80.0) ** 2 - r0 ** 2)
    s = np.linspace(-smax, smax, 400001)
    zz = np.sqrt(r0 ** 2 + s ** 2) - scenes.RE
    n = 10 * q * 1013.25 * np.exp(-zz / 7.0) / (1.3806504e-23 * T)
    Ul = float(np.sum((n[1:] + n[:-1]) * 0.5 * np.diff(s)))
    assert 1 - tau == pytest.approx(Ul / 2.0 ** 64, rel=1e-3)


def test_two_gases_multiply(oracle, tmp_path):
    """EGA per emitter is independent; channel transmittance is the product."""
    nu = [700.0]
    tb = _write_linear_tables(str(tmp_path), ["CO2", "O3"], nu)
    kw = dict(tblbase=tb, ctm_co2=0, ctm_h2o=0, ctm_n2=0, ctm_o2=0, refrac=0)
    z = np.arange(81.0)
    p = 1013.25 * np.exp(-z / 7.0)
    t = np.full(81, 250.0)
    obs0 = scenes.make_obs(scenes.limb_geometry(800.0, [25.0]))
    taus = []
    for qs, gases in (([2e-8], ["CO2"]), ([1e-8], ["O3"]), ([2e-8, 1e-8], ["CO2", "O3"])):
        ctl = abi.make_ctl(gases, nu, **kw)
        atm = abi.make_atm(z, p, t, [np.full(81, v) for v in qs])
        obs = abi.copy_struct(obs0)
        oracle.formod(ctl, atm, obs)
        taus.append(obs.tau[0][0])
    assert taus[2] == pytest.approx(taus[0] * taus[1], rel=1e-12)


def test_srcfunc_is_filter_weighted_planck(oracle, tmp_path):
    nu = [792.0]
    tb = _write_linear_tables(str(tmp_path), ["CO2"], nu)
    ctl = abi.make_ctl(["CO2"], nu, tblbase=tb)
    st, sr = oracle.srcfunc_table(ctl)
    assert st[0] == 100.0 and st[-1] == 400.0
    for it in (0, 300, 599, 1199):
        assert sr[it, 0] == pytest.approx(_boxcar_planck(st[it], nu[0]), rel=1e-5)


def test_reference_filter_file_srcfunc(oracle, tmp_path):
    """The reference's own boxcar filter file (tests/limb_test) gives the
    boxcar mean of the Planck function."""
    import shutil
    tb = os.path.join(str(tmp_path), "boxcar")
    shutil.copy(os.path.join(GOLD, "limb_test", "boxcar_792.0000.filt"), tb + "_792.0000.filt")
    ctl = abi.make_ctl([], [792.0], tblbase=tb)
    st, sr = oracle.srcfunc_table(ctl)
    for it in (0, 500, 1199):
        assert sr[it, 0] == pytest.approx(_boxcar_planck(st[it], 792.0), rel=1e-5)


def test_straight_ray_geometry(oracle):
    """refrac = 0: limb tangent height equals the view point height, and the
    nadir path length equals the chord between 80 km and the ground."""
    ctl = abi.make_ctl(["CO2"], [700.0], refrac=0)
    atm = scenes.make_profile_atm(["CO2"])
    obs = scenes.make_obs(scenes.limb_geometry(800.0, [10.0, 30.0, 50.0]))
    for ir, zt in enumerate((10.0, 30.0, 50.0)):
        oracle.trace_ray(ctl, atm, obs, ir)
        assert obs.tpz[ir] == pytest.approx(zt, abs=2e-3)
    for th in (0.0, 20.0, 40.0):
        obs = scenes.make_obs(scenes.nadir_geometry(705.0, [th]))
        r = oracle.trace_ray(ctl, atm, obs, 0)
        g = math.radians(th)
        ro = scenes.RE + 705.0
        # distance from the observer to radius R along the view direction
        def dist(R):
            return ro * math.cos(g) - math.sqrt(R ** 2 - (ro * math.sin(g)) ** 2)
        # the entry search stops within 1 m below the top (src/jurassic.c:4261-4272)
        assert 80.0 - 1e-3 < r["z"][0] <= 80.0
        # the ground exit is clipped by a linear fraction in z, exact up to
        # the curvature of z along a slanted step, O(ds^2 / RE) ~ 1e-6 km
        assert float(np.sum(r["ds"])) == pytest.approx(dist(scenes.RE) - dist(scenes.RE + r["z"][0]), abs=1e-5)
        assert r["sft"] == pytest.approx(288.0)


def test_continua_identities(oracle):
    """Structure of the continuum formulas (src/jurassic.c:1724-1741,
    2762-2791, 2841-2917): the CO2 quadratic passes through the 230/260/296 K
    tables; H2O self-broadening reduces to the 296 K table at 296 K; N2/O2
    reduce to the tabulated B at knots and tr = 296 K."""
    from ctm_helper import ctm_tables
    tab = ctm_tables()
    L = oracle.lib()
    NA, P0 = 6.02214199e23, 1013.25
    for nu in (650.0, 700.0, 1000.0, 2000.0):  # knots of the 2 cm^-1 grid
        iw = int(nu / 2 + 1)
        for t, key in ((296.0, "JUR_CTM_CO2_296"), (260.0, "JUR_CTM_CO2_260"), (230.0, "JUR_CTM_CO2_230")):
            got = L.orc_ctmco2(nu, 500.0, t, 1e22)
            exp = 1e22 / NA / 1000 * 500.0 / P0 * tab[key][iw - 1]
            assert got == pytest.approx(exp, rel=2e-6)
    assert L.orc_ctmco2(5000.0, 500.0, 250.0, 1e22) == 0.0
    for nu in (700.0, 1000.0, 1500.0):  # outside 820..960: sfac = 1
        iw = int(nu / 10 + 1)
        q = 1.0  # pure self-broadening
        got = L.orc_ctmh2o(nu, 500.0, 296.0, q, 1e22)
        exp = nu * 1e22 * math.tanh(0.7193876 / 296.0 * nu) * 500.0 / P0 * tab["JUR_CTM_H2O_296"][iw - 1] * 1e-20
        assert got == pytest.approx(exp, rel=1e-12)
        got260 = L.orc_ctmh2o(nu, 500.0, 260.0, q, 1e22)
        exp260 = nu * 1e22 * math.tanh(0.7193876 / 260.0 * nu) * 296 / 260.0 * 500.0 / P0 * tab["JUR_CTM_H2O_260"][iw - 1] * 1e-20
        assert got260 == pytest.approx(exp260, rel=1e-12)
    for i in (0, 10, 50):
        nu = tab["JUR_CTM_N2_NU"][i]
        got = L.orc_ctmn2(nu, P0, 296.0)
        q = 0.79
        exp = 0.1 * (273.0 / 296.0) ** 2 * q * tab["JUR_CTM_N2_B"][i] * (q + (1 - q) * (1.294 - 0.4545))
        assert got == pytest.approx(exp, rel=1e-12)
        nu = tab["JUR_CTM_O2_NU"][i]
        got = L.orc_ctmo2(nu, P0, 296.0)
        exp = 0.1 * (273.0 / 296.0) ** 2 * 0.21 * tab["JUR_CTM_O2_B"][i]
        assert got == pytest.approx(exp, rel=1e-12)
    assert L.orc_ctmn2(1000.0, P0, 296.0) == 0.0
    assert L.orc_ctmo2(1000.0, P0, 296.0) == 0.0


def test_hydrostatic_isothermal(oracle):
    """Dry isothermal atmosphere: p = p_ref exp(-(z - z_ref)/H), H = R T/(M g)."""
    ctl = abi.make_ctl(["CO2"], [700.0], hydz=10.0)
    z = np.arange(0, 81, 1.0)
    T = 240.0
    atm = abi.make_atm(z, np.full(81, 1.0), np.full(81, T), [np.full(81, 4e-4)])
    atm.p[10] = 250.0
    oracle.lib().orc_hydrostatic(C.byref(ctl), C.byref(atm))
    H = 8.3144598 * T / (28.96456e-3 * 9.80665) / 1000.0
    for i in (0, 5, 30, 80):
        assert atm.p[i] == pytest.approx(250.0 * math.exp(-(z[i] - 10.0) / H), rel=1e-12)


def test_kernel_linear_table_column(oracle, tmp_path):
    """kernel() columns for vmr on a linear table: rad = B (1 - tau),
    1 - tau = U/2^64 is linear in the level vmr, so the forward difference is
    exact: K_j = B/2^64 * sum_i 10 p_i/(kB t_i) ds_i w_j(z_i)."""
    nu = [700.0]
    tb = _write_linear_tables(str(tmp_path), ["CO2"], nu)
    ctl = abi.make_ctl(["CO2"], nu, tblbase=tb, ctm_co2=0, ctm_h2o=0, ctm_n2=0,
                       ctm_o2=0, refrac=0, retq_zmin=[-100], retq_zmax=[100])
    z = np.arange(81.0)
    T = 250.0
    atm = abi.make_atm(z, 1013.25 * np.exp(-z / 7.0), np.full(81, T), [np.full(81, 2e-8)])
    obs = scenes.make_obs(scenes.limb_geometry(800.0, [30.0]))
    k = oracle.kernel(ctl, atm, obs)
    r = oracle.trace_ray(ctl, atm, obs, 0)
    b = _boxcar_planck(T, nu[0])
    w = np.zeros(81)
    for zi, pi, ti, dsi in zip(r["z"], r["p"], r["t"], r["ds"]):
        j = min(int(math.floor(zi)), 79)
        f = zi - z[j]
        c = 10 * pi / (1.3806504e-23 * ti) * dsi
        w[j] += c * (1 - f)
        w[j + 1] += c * f
    exp = b / 2.0 ** 64 * w
    assert k.shape == (1, 81)
    np.testing.assert_allclose(k[0], exp, rtol=2e-5, atol=1e-12 * np.abs(exp).max())


def test_optimal_estimation_linear_model(oracle):
    """On a linear forward model LM converges to the linear-Gaussian optimum
    x = xa + (K^T Se^-1 K + Sa^-1)^-1 K^T Se^-1 (y - F(xa)), F(x) = K x + c."""
    ctl = abi.make_ctl(["CO2"], [700.0 + i for i in range(6)], rett_zmin=0, rett_zmax=100)
    z = np.arange(0, 20.0, 1.0)
    atm = abi.make_atm(z, 1013.25 * np.exp(-z / 7), 250 + np.zeros(20), [np.full(20, 4e-4)])
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0, 10.0, 20.0]))
    m, n = 18, 20
    rng = np.random.default_rng(5)
    K = rng.normal(size=(m, n)) * 0.05
    c = rng.normal(size=m)
    x_true = 250 + rng.normal(size=n) * 2
    y = K @ x_true + c
    for ir in range(3):
        for id_ in range(6):
            obs.rad[id_][ir] = y[ir * 6 + id_]

    def fwd(user, xp, yp, kp):
        x = np.ctypeslib.as_array(xp, shape=(n,))
        yy = np.ctypeslib.as_array(yp, shape=(m,))
        yy[:] = K @ x + c
        if kp:
            np.ctypeslib.as_array(kp, shape=(m * n,))[:] = K.ravel()

    ret = oracle.orc_ret_t()
    ret.kernel_recomp = 3
    ret.conv_itmax = 30
    ret.conv_dmin = 1e-8
    for i in range(6):
        ret.err_noise[i] = 0.1
    ret.err_temp = 2.0
    ret.err_temp_cz = 5.0
    ret.err_temp_ch = 500.0
    atm_i, obs_i, hist, nit = oracle.optimal_estimation(ret, ctl, obs, atm, fwd)
    xa = np.full(n, 250.0)
    sa = 4.0 * np.exp(-np.abs(z[:, None] - z[None, :]) / 5.0)
    se_inv = np.eye(m) / 0.01
    xhat = xa + np.linalg.solve(K.T @ se_inv @ K + np.linalg.inv(sa), K.T @ se_inv @ (y - K @ xa - c))
    xi = np.ctypeslib.as_array(atm_i.t)[:n]
    np.testing.assert_allclose(xi, xhat, rtol=0, atol=1e-6)
    assert np.all(np.diff(hist) <= 1e-12)


def test_optimal_estimation_error_analysis_linear_model(oracle):
    """Error analysis (src/retrieval.c:668-726) on a linear forward model
    against its closed forms: cov = (K^T Se^-1 K + Sa^-1)^-1 with Se = noise^2
    + (formod % y)^2, G = cov K^T Se^-1, noise and forward-model errors
    sqrt(diag(G S G^T)), A = G K, and the identity A = I - cov Sa^-1."""
    nd = 6
    ctl = abi.make_ctl(["CO2"], [700.0 + i for i in range(nd)], rett_zmin=0, rett_zmax=100)
    z = np.arange(0, 20.0, 1.0)
    atm = abi.make_atm(z, 1013.25 * np.exp(-z / 7), 250 + np.zeros(20), [np.full(20, 4e-4)])
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0, 10.0, 20.0]))
    m, n = 18, 20
    rng = np.random.default_rng(7)
    K = rng.normal(size=(m, n)) * 0.05
    c = 50 + rng.normal(size=m)
    y = K @ (250 + rng.normal(size=n) * 2) + c
    for ir in range(3):
        for id_ in range(nd):
            obs.rad[id_][ir] = y[ir * nd + id_]

    def fwd(user, xp, yp, kp):
        x = np.ctypeslib.as_array(xp, shape=(n,))
        np.ctypeslib.as_array(yp, shape=(m,))[:] = K @ x + c
        if kp:
            np.ctypeslib.as_array(kp, shape=(m * n,))[:] = K.ravel()

    ret = oracle.orc_ret_t()
    ret.kernel_recomp = 3
    ret.conv_itmax = 30
    ret.conv_dmin = 1e-8
    noise = 0.1 + 0.05 * np.arange(nd)
    formod = np.array([0.0, 1.0, 0.0, 2.0, 0.5, 0.0])  # percent
    for i in range(nd):
        ret.err_noise[i] = noise[i]
        ret.err_formod[i] = formod[i]
    ret.err_temp, ret.err_temp_cz, ret.err_temp_ch = 2.0, 5.0, 500.0
    *_, ana = oracle.optimal_estimation(ret, ctl, obs, atm, fwd, analysis=True)
    sa = 4.0 * np.exp(-np.abs(z[:, None] - z[None, :]) / 5.0)
    sn = np.tile(noise, 3)
    sf = np.abs(np.tile(formod, 3) / 100 * y)
    se_inv = np.diag(1 / (sn ** 2 + sf ** 2))
    cov = np.linalg.inv(K.T @ se_inv @ K + np.linalg.inv(sa))
    gain = cov @ K.T @ se_inv
    avk = gain @ K
    tol = dict(rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ana["cov"], cov, **tol)
    d = np.sqrt(np.diag(cov))
    np.testing.assert_allclose(ana["corr"], cov / d[:, None] / d[None, :], **tol)
    np.testing.assert_allclose(ana["gain"], gain, **tol)
    np.testing.assert_allclose(ana["avk"], avk, **tol)
    np.testing.assert_allclose(ana["err_total"], d, **tol)
    np.testing.assert_allclose(ana["err_noise"], np.sqrt(np.diag(gain @ np.diag(sn ** 2) @ gain.T)), **tol)
    np.testing.assert_allclose(ana["err_formod"], np.sqrt(np.diag(gain @ np.diag(sf ** 2) @ gain.T)), **tol)
    np.testing.assert_allclose(ana["cont"], avk.sum(axis=1), **tol)
    np.testing.assert_allclose(ana["res"], 1 / np.diag(avk), **tol)
    np.testing.assert_allclose(ana["avk"], np.eye(n) - cov @ np.linalg.inv(sa), rtol=0, atol=1e-9)
