"""GPU adjoint Jacobian against finite differences of the FP64 oracle.

The adjoint differentiates the forward model.  With refraction off the path
does not depend on the state; with refraction on it bends with p and T and
the adjoint follows it.  Two checks:

* Per state block (T, each emitter, ...), the relative L2 error against
  central differences of the oracle with kernel()'s step sizes
  (src/jurassic.c:4046-4068) scaled by 1/100 and 1/1000, the better of the
  two: larger steps cross cells of the piecewise-linear tables, smaller steps
  leave rounding noise where a column barely changes the radiance (e.g. a
  few ppt of O3 at the ground).
* Directional derivatives: K v for v = kernel()'s steps on all elements of
  one block against the oracle's central difference along v.  Moving all
  elements of a block at once changes the radiance enough that rounding
  stays far below the bound, so this check is the tight one (measured:
  ~1e-8 for T, CO2 and H2O, ~1e-5 for the trace gas O3).

With refraction on, the adjoint also differentiates the path (its bending
by the refractivity of p and T below 60 km), and the same checks apply
against differences that move the path; against kernel()'s one-sided 1 K
differences the bound is kernel()'s own truncation error.
All GPU calls go through the C-ABI.
"""
import ctypes as C

import numpy as np
import pytest

from jurassic_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from jurassic_amd import gpu as G
    G.lib()
    return G


def _blocks(oracle, ctl, atm):
    x = np.zeros(abi.XMAP_LEN)
    iqa = np.zeros(abi.XMAP_LEN, dtype=np.int32)
    n = oracle.lib().orc_atm2x(C.byref(ctl), C.byref(atm), oracle._ptr(x),
                               iqa.ctypes.data_as(C.POINTER(C.c_int)), None)
    return iqa[:n]


def _block_errors(oracle, ctl, atm, k1, k0):
    iq = _blocks(oracle, ctl, atm)
    err = {}
    for q in np.unique(iq):
        sel = iq == q
        err[int(q)] = np.linalg.norm(k1[:, sel] - k0[:, sel]) / np.linalg.norm(k0[:, sel])
    return err


def _adjoint(gpu, oracle, ctl, atm, obs, precision=abi.JUR_FP64):
    m, n = oracle.sizes(ctl, atm, obs)
    o = abi.copy_struct(obs)
    k = gpu.kernel(ctl, abi.copy_struct(atm), o, method=abi.JUR_JAC_ADJOINT, precision=precision, m=m, n=n)
    return k, o


def _fd(oracle, ctl, atm, obs, scale=1e-3, central=True):
    return oracle.kernel(ctl, abi.copy_struct(atm), abi.copy_struct(obs), central=central, scale=scale)


def _best_block_errors(oracle, ctl, atm, obs, k1, scales=(1e-2, 1e-3)):
    errs = [_block_errors(oracle, ctl, atm, k1, _fd(oracle, ctl, atm, obs, scale=s)) for s in scales]
    return {q: min(e[q] for e in errs) for q in errs[0]}


def _y(oracle, ctl, atm, obs):
    o = abi.copy_struct(obs)
    oracle.formod(ctl, abi.copy_struct(atm), o)
    m, _ = oracle.sizes(ctl, atm, obs)
    y = np.zeros(m + 1)
    oracle.lib().orc_obs2y(C.byref(ctl), C.byref(o), oracle._ptr(y), None, None)
    return y[:m]


def _directional(oracle, ctl, atm, obs, k1, rel_steps=(1e-4, 1e-3)):
    """Per block, |K v - (F(x + h v) - F(x - h v)) / 2h| / |F'v| for v = the
    kernel() step sizes on every element of the block (one sign, so that the
    response does not cancel), the better of two step sizes h: the smaller
    keeps table-cell boundaries out of the step, the larger lifts the
    difference above the rounding of levels that barely change the
    radiance."""
    n = k1.shape[1]
    x = np.zeros(n + 1)
    iqa = np.zeros(n + 1, dtype=np.int32)
    oracle.lib().orc_atm2x(C.byref(ctl), C.byref(atm), oracle._ptr(x),
                           iqa.ctypes.data_as(C.POINTER(C.c_int)), None)
    x, iqa = x[:n], iqa[:n]
    out = {}
    for q in np.unique(iqa):
        sel = iqa == q
        # kernel()'s step sizes set the scale of each element
        step = np.where(q == 1, 1.0, np.maximum(np.abs(0.01 * x), 1e-15))
        v = np.where(sel, step, 0.0)
        best = np.inf
        for h in rel_steps:
            ys = []
            for sgn in (1, -1):
                a = abi.copy_struct(atm)
                xs = x + sgn * h * v
                oracle.lib().orc_x2atm(C.byref(ctl), oracle._ptr(xs), C.byref(a))
                ys.append(_y(oracle, ctl, a, obs))
            fd = (ys[0] - ys[1]) / (2 * h)
            best = min(best, np.linalg.norm(k1 @ v - fd) / np.linalg.norm(fd))
        out[int(q)] = best
    return out


def _nadir(refrac=0, **kw):
    ctl, atm, obs = scenes.nadir_jac(1, nd=8, nang=4)
    ctl.refrac = refrac
    for k, v in kw.items():
        setattr(ctl, k, v)
    return ctl, atm, obs


def _check(err, tol, what):
    bad = {q: e for q, e in err.items() if not e <= tol}
    assert not bad, "%s: block errors above %g: %s (all: %s)" % (what, tol, bad, err)


def test_adjoint_fp64_nadir(gpu, oracle):
    ctl, atm, obs = _nadir()
    k1, o = _adjoint(gpu, oracle, ctl, atm, obs)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 3e-4, "nadir FP64")
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "nadir FP64 directional")
    # the forward part of the pass is the forward model
    o0 = abi.copy_struct(obs)
    oracle.formod(ctl, abi.copy_struct(atm), o0)
    r1, t1 = abi.obs_arrays(o, ctl)
    r0, t0 = abi.obs_arrays(o0, ctl)
    assert np.allclose(r1, r0, rtol=1e-12, atol=0) and np.allclose(t1, t0, rtol=1e-12, atol=1e-300)


def test_adjoint_fp32_nadir(gpu, oracle):
    ctl, atm, obs = _nadir()
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs, precision=abi.JUR_FP32)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 1e-3, "nadir FP32")
    _check(_directional(oracle, ctl, atm, obs, k1), 1e-3, "nadir FP32 directional")


def test_adjoint_bbt(gpu, oracle):
    ctl, atm, obs = _nadir(write_bbt=1)
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 3e-4, "nadir BT")
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "nadir BT directional")


def test_adjoint_cga(gpu, oracle):
    ctl, atm, obs = _nadir(formod=0)
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 3e-4, "nadir CGA")
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "nadir CGA directional")


def _limb(refrac):
    ctl, atm, obs = scenes.limb_3gas(2, nrays=8)
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    for ig in range(3):
        ctl.retq_zmin[ig], ctl.retq_zmax[ig] = -100, 100
    ctl.refrac = refrac
    return ctl, atm, obs


def test_adjoint_limb(gpu, oracle):
    ctl, atm, obs = _limb(0)
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 3e-4, "limb FP64")
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "limb FP64 directional")
    k2, _ = _adjoint(gpu, oracle, ctl, atm, obs, precision=abi.JUR_FP32)
    _check(_directional(oracle, ctl, atm, obs, k2), 1e-3, "limb FP32 directional")


def test_adjoint_limb_refraction(gpu, oracle):
    """Refraction on: the path bends with p and T below 60 km, and the
    adjoint follows it (k_adj_path), so it matches central differences of
    the oracle, which move the path, as tightly as with refraction off.
    Pressure is in the state here too (it enters only through the
    refractivity and intpol_atm)."""
    ctl, atm, obs = _limb(1)
    ctl.retp_zmin, ctl.retp_zmax = -100, 100
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    be = _best_block_errors(oracle, ctl, atm, obs, k1)
    de = _directional(oracle, ctl, atm, obs, k1)
    print("limb refraction FP64: block", be, "directional", de)
    # the refracted limb path crosses more table knots per step; the float
    # bracket slopes make the forward model continuous only to ~1e-7 of a
    # bracket's rise there, which sets a floor near 3e-4 for central
    # differences along a whole block
    _check(be, 1e-3, "limb refraction FP64")
    _check(de, 5e-4, "limb refraction FP64 directional")
    k2, _ = _adjoint(gpu, oracle, ctl, atm, obs, precision=abi.JUR_FP32)
    de2 = _directional(oracle, ctl, atm, obs, k2)
    print("limb refraction FP32: directional", de2)
    _check(de2, 2e-3, "limb refraction FP32 directional")


def test_adjoint_limb_refraction_vs_kernel(gpu, oracle):
    """Against kernel() itself (one-sided differences with its own steps:
    1 K for T, 1% for vmr), which moves the path as the adjoint does.  What
    remains is kernel()'s truncation error, the secant of its step over
    the nonlinear emissivity growth: the adjoint is as close to kernel() as
    converged central differences (1/100 of kernel()'s steps) are, and
    within 1e-3 of those."""
    ctl, atm, obs = _limb(1)
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    k0 = _fd(oracle, ctl, atm, obs, scale=1.0, central=False)
    kc = _fd(oracle, ctl, atm, obs, scale=1e-2, central=True)
    e_adj = _block_errors(oracle, ctl, atm, k1, k0)
    e_fd = _block_errors(oracle, ctl, atm, kc, k0)
    e_c = _best_block_errors(oracle, ctl, atm, obs, k1)
    print("limb refraction vs kernel(): adjoint", e_adj, "central differences", e_fd, "adjoint vs central", e_c)
    _check(e_c, 1e-3, "limb refraction vs central differences")
    _check({q: e_adj[q] - 1.1 * e_fd[q] for q in e_adj}, 1e-3, "limb refraction vs kernel() beyond its truncation")


def test_adjoint_nadir_refraction_jacobian_workload(gpu, oracle):
    """The Jacobian workload's own setting: nadir, refraction on, FP32
    lookups, 16 channels, 8 view angles 0..50 deg, T and CO2/H2O/O3 on 81
    levels, against kernel() and against central differences."""
    ctl, atm, obs = scenes.nadir_jac(1, nd=16, nang=8)
    assert ctl.refrac == 1
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs, precision=abi.JUR_FP32)
    ek = _block_errors(oracle, ctl, atm, k1, _fd(oracle, ctl, atm, obs, scale=1.0, central=False))
    ec = _best_block_errors(oracle, ctl, atm, obs, k1)
    print("nadir refraction FP32 vs kernel():", ek, "vs central:", ec)
    _check(ek, 5e-3, "nadir refraction FP32 vs kernel()")
    _check(ec, 1e-3, "nadir refraction FP32 vs central differences")


def test_adjoint_refraction_coarse_steps(gpu, oracle):
    """Steps of RAYDS = 10 km with no RAYDZ limit over levels 0.125 km apart
    between 12 and 36 km (249 levels): half a step climbs 0.3-0.45 km there,
    so a step's point and its four refractivity probes often span more than
    three levels, where the path adjoint adds their terms point by point
    instead of through its three-level window (and the window jumps).  T
    and p in the state, refraction on, FP64."""
    z = np.concatenate([np.arange(0.0, 12.0, 1.0), np.arange(12.0, 36.0, 0.125), np.arange(36.0, 81.0, 1.0)])
    assert len(z) <= abi.NP
    ctl, _, obs = scenes.limb_3gas(2, nrays=8)
    atm = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=2, z=z)
    ctl.refrac = 1
    ctl.raydz = 0.0
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    ctl.retp_zmin, ctl.retp_zmax = -100, 100
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    assert np.isfinite(k1).all() and np.abs(k1).max() > 0
    de = _directional(oracle, ctl, atm, obs, k1)
    print("coarse refracted steps FP64: directional", de)
    _check(de, 5e-4, "coarse refracted steps FP64 directional")


def _scalar_scene():
    ctl, atm, obs = scenes.nadir_co2()
    ctl.refrac = 0
    ctl.ncl, ctl.nsf = 2, 2
    ctl.clnu[0], ctl.clnu[1] = 600.0, 700.0
    ctl.sfnu[0], ctl.sfnu[1] = 600.0, 700.0
    ctl.sftype = 3
    ctl.sfsza = 30.0
    ctl.nw = 1
    atm.clz, atm.cldz = 8.0, 1.5
    atm.clk[0], atm.clk[1] = 0.02, 0.03
    atm.sft, atm.sfeps[0], atm.sfeps[1] = 290.0, 0.9, 0.8
    for ip in range(atm.np):
        atm.k[0][ip] = 1e-3 * np.exp(-atm.z[ip] / 10)
    ctl.retp_zmin, ctl.retp_zmax = -100, 100
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    ctl.retk_zmin[0], ctl.retk_zmax[0] = -100, 100
    ctl.ret_clz = ctl.ret_cldz = ctl.ret_clk = 1
    ctl.ret_sft = ctl.ret_sfeps = 1
    return ctl, atm, obs


def test_adjoint_pressure_extinction_clouds_surface(gpu, oracle):
    ctl, atm, obs = _scalar_scene()
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    _check(_best_block_errors(oracle, ctl, atm, obs, k1), 3e-4, "p, T, k, clouds, surface")
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "p, T, k, clouds, surface directional")


def test_adjoint_path_columns_by_fd(gpu, oracle):
    # surface altitude and pressure move the path: jur_gpu_kernel covers them
    # with kernel()'s one-sided differences
    ctl, atm, obs = _scalar_scene()
    ctl.ret_sfz = ctl.ret_sfp = 1
    k1, _ = _adjoint(gpu, oracle, ctl, atm, obs)
    k0 = _fd(oracle, ctl, atm, obs, scale=1.0, central=False)
    iq = _blocks(oracle, ctl, atm)
    ng, nw, ncl = ctl.ng, ctl.nw, ctl.ncl
    for q in (4 + ng + nw + ncl, 5 + ng + nw + ncl):
        sel = iq == q
        assert sel.sum() == 1
        assert np.allclose(k1[:, sel], k0[:, sel], rtol=1e-6, atol=1e-9 * np.abs(k0[:, sel]).max())


def test_adjoint_masked_rows(gpu, oracle):
    ctl, atm, obs = _nadir()
    obs.rad[1][0] = float("nan")
    obs.rad[5][2] = float("nan")
    k1, o = _adjoint(gpu, oracle, ctl, atm, obs)
    k0 = _fd(oracle, ctl, atm, obs)
    assert k1.shape == k0.shape == (ctl.nd * obs.nr - 2, k0.shape[1])
    _check(_directional(oracle, ctl, atm, obs, k1), 3e-5, "masked rows directional")
    assert np.isnan(o.rad[1][0]) and np.isnan(o.rad[5][2])


def test_adjoint_unsupported(gpu, oracle):
    ctl, atm, obs = _nadir(hydz=0.0)
    m, n = oracle.sizes(ctl, atm, obs)
    with pytest.raises(gpu.JurassicError) as e:
        gpu.kernel(ctl, atm, obs, method=abi.JUR_JAC_ADJOINT, m=m, n=n)
    assert e.value.code == 6


def test_device_jacobian_matches_host(gpu, oracle):
    torch = pytest.importorskip("torch")
    ctl, atm0, obs = _nadir()
    atm1 = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=7)
    m, n = oracle.sizes(ctl, atm0, obs)
    nr, nd = obs.nr, ctl.nd
    ks = [_adjoint(gpu, oracle, ctl, a, obs, precision=abi.JUR_FP32)[0] for a in (atm0, atm1)]
    xm, n2 = gpu.state_map(ctl, atm0)
    assert n2 == n
    dev = torch.device("cuda:0")
    atms = (abi.atm_t * 2)(atm0, atm1)
    a_d = torch.from_numpy(np.frombuffer(bytes(atms), dtype=np.uint8).copy()).to(dev)
    # ray order: profile-minor within each geometry
    prof = np.array([i % 2 for i in range(2 * nr)], dtype=np.int32)
    geom = np.zeros((2 * nr, 7))
    for i in range(2 * nr):
        ir = i // 2
        geom[i] = [obs.time[ir], obs.obsz[ir], obs.obslon[ir], obs.obslat[ir], obs.vpz[ir],
                   obs.vplon[ir], obs.vplat[ir]]
    prof_d = torch.from_numpy(prof).to(dev)
    geom_d = torch.from_numpy(geom).to(dev)
    pmap_d = torch.zeros(2, dtype=torch.int32, device=dev)
    rad = torch.zeros(2 * nr, nd, dtype=torch.float64, device=dev)
    tau = torch.zeros_like(rad)
    kd = torch.full((2 * nr, nd, n), float("nan"), dtype=torch.float64, device=dev)
    ctx = gpu.Context(ctl)
    try:
        ctx.jacobian_device(2, a_d.data_ptr(), 2 * nr, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(),
                            tau.data_ptr(), xm[None, :], pmap_d.data_ptr(), n, kd.data_ptr(),
                            precision=abi.JUR_FP32)
        torch.cuda.synchronize()
    finally:
        ctx.close()
    kh = kd.cpu().numpy()
    for p in range(2):
        rows = kh[p::2].reshape(nr * nd, n)
        assert np.allclose(rows, ks[p], rtol=1e-12, atol=1e-12 * np.abs(ks[p]).max())


def _nadir_batch(ctl, nprof, nang, torch):
    """nprof atmospheres x nang nadir views, ray order view major."""
    atms = (abi.atm_t * nprof)()
    for k in range(nprof):
        atms[k] = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=k)
    geom = scenes.nadir_geometry(705.0, np.arange(nang) * 50.0 / 7.0)
    nray = nprof * nang
    prof = np.tile(np.arange(nprof, dtype=np.int32), nang)
    g = np.zeros((nray, 7))
    for j in range(nang):
        for c, key in enumerate(("obsz", "obslon", "obslat", "vpz", "vplon", "vplat")):
            g[j * nprof:(j + 1) * nprof, 1 + c] = geom[key][j]
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(np.frombuffer(bytes(atms), dtype=np.uint8).copy()).to(dev)
    return atms, a_d, torch.from_numpy(prof).to(dev), torch.from_numpy(g).to(dev), nray


def test_adjoint_path_batch_regimes_agree(gpu):
    """The adjoint of the refracted path runs two ways: four channels per
    lane sharing each step's coefficients for large batches, one channel per
    lane with the next step's inputs loaded ahead for small ones.  The K rows
    of the same rays agree between a batch of 688 atmospheres x 8 views with
    96 channels (the first way) and one of its first 64 atmospheres (the
    second); each lane's arithmetic is its own, so only the T and p columns'
    order of additions may differ."""
    torch = pytest.importorskip("torch")
    ctl = scenes.nadir_jac_ctl(96)
    assert ctl.refrac == 1
    nang = 8
    out = []
    for nprof in (688, 64):
        atms, a_d, prof_d, geom_d, nray = _nadir_batch(ctl, nprof, nang, torch)
        xm, n = gpu.state_map(ctl, atms[0])
        ctx = gpu.Context(ctl)
        try:
            rad = torch.zeros(nray, ctl.nd, dtype=torch.float64, device="cuda:0")
            tau = torch.zeros_like(rad)
            kd = torch.zeros(nray, ctl.nd, n, dtype=torch.float64, device="cuda:0")
            pmap = torch.zeros(nprof, dtype=torch.int32, device="cuda:0")
            ctx.jacobian_device(nprof, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(),
                                rad.data_ptr(), tau.data_ptr(), xm[None, :], pmap.data_ptr(), n,
                                kd.data_ptr(), precision=abi.JUR_FP32)
            torch.cuda.synchronize()
            rows = torch.cat([kd[j * nprof:j * nprof + 64] for j in range(nang)])
            out.append((rad.view(nang, nprof, -1)[:, :64].cpu().numpy(), rows.cpu().numpy()))
            del kd
        finally:
            ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    k0, k1 = out[0][1], out[1][1]
    assert np.isfinite(k0).all() and np.abs(k0).max() > 0
    np.testing.assert_allclose(k1, k0, rtol=1e-12, atol=1e-12 * np.abs(k0).max())
