"""The wide path: more channels than the ctl_t capacity ND (jur_ctx_create_wide)
and eight emitters, as in a hyperspectral nadir sounder.

160 channels over 650..2700 cm^-1 (CO2, H2O, N2 and O2 continua all active
somewhere in the band), emitters CO2, H2O, O3, CH4, N2O, CO, HNO3, F11, with
tables per (channel, emitter) on the 40 p x 20 T x 200 u grid.  The oracle
holds at most ND = 128 channels per call, so it runs on channel slices of
the same scene and the GPU's wide pass is compared slice by slice:
  * forward model, EGA and CGA: FP64 within 1e-10 relative, FP32 within
    0.01 K brightness temperature;
  * adjoint Jacobian of T, H2O and O3 (n = 243; refraction off, so the path
    is fixed) against the oracle's central differences along directions in
    each block, and against the Jacobian of a context of the first ND
    channels (the path the other adjoint tests cover) to rounding.
All GPU calls go through the C-ABI.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from jurassic_amd import abi, scenes

pytestmark = pytest.mark.gpu

GASES = ["CO2", "H2O", "O3", "CH4", "N2O", "CO", "HNO3", "F11"]
GAS_K = [1e-2, 2e-3, 5e-3, 3e-3, 2e-3, 5e-4, 7e-3, 2e-4]
ND_WIDE = 160
NU = [650.0 + i * 2050.0 / (ND_WIDE - 1) for i in range(ND_WIDE)]
SLICES = [(0, 128), (128, ND_WIDE)]
ANGLES = [0.0, 17.0, 33.0, 48.0]
NPROF = 2


def _variant(id_):
    """Seven table variants spread over the channels so that a channel
    reading its neighbour's table is seen: variant (3 id) mod 7."""
    return (3 * id_) % 7


@pytest.fixture(scope="module")
def scene():
    return make_scene()


def make_scene():
    """Tables (one file per variant and emitter, linked under every channel
    name) and boxcar filters in the reference's file formats."""
    d = scenes.scene_dir("wide160")
    os.makedirs(d, exist_ok=True)
    tb = os.path.join(d, "wide")
    p, t, u = scenes.table_grids()
    for v in range(7):
        for g, k in zip(GASES, GAS_K):
            src = os.path.join(d, "var%d_%s.bin" % (v, g))
            if not os.path.exists(src):
                scenes.write_table_bin(src + ".tmp", p, t, u, scenes.synthetic_eps(p, t, u, k * (0.4 + 0.15 * v)))
                os.replace(src + ".tmp", src)
    for id_, nu in enumerate(NU):
        for g in GASES:
            dst = scenes.table_name(tb, nu, g)
            if not os.path.exists(dst):
                os.link(os.path.join(d, "var%d_%s.bin" % (_variant(id_), g)), dst)
        f = "%s_%.4f.filt" % (tb, nu)
        if not os.path.exists(f):
            scenes.write_boxcar_filter(f + ".tmp", nu)
            os.replace(f + ".tmp", f)
    return tb


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from jurassic_amd import gpu as G
    G.lib()
    return G


def _ctl(tb, nus, formod=1, refrac=1):
    return abi.make_ctl(GASES, nus, tblbase=tb, formod=formod, refrac=refrac)


def _atms():
    return [scenes.make_profile_atm(GASES, iprof=k) for k in range(NPROF)]


def _geom():
    return scenes.nadir_geometry(705.0, ANGLES)


def _device_batch(torch, atms):
    """Ray-major batch: ray i views angle i // NPROF of profile i % NPROF."""
    g = _geom()
    nr = len(ANGLES)
    nray = nr * NPROF
    arr = (abi.atm_t * NPROF)(*atms)
    prof = np.array([i % NPROF for i in range(nray)], dtype=np.int32)
    geom = np.zeros((nray, 7))
    for i in range(nray):
        j = i // NPROF
        geom[i] = [0.0, g["obsz"][j], g["obslon"][j], g["obslat"][j], g["vpz"][j], g["vplon"][j], g["vplat"][j]]
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(dev)
    return a_d, torch.from_numpy(prof).to(dev), torch.from_numpy(geom).to(dev), nray


def _oracle_forward(oracle, tb, atms, formod, refrac=1, chans=None, angles=None):
    """Oracle radiance and transmittance [nray][len(chans)] of the rays of
    the view angles with indices `angles` (default: all) over the
    atmospheres atms, ray-major as the device batch, for the channels chans
    (default: all).  The channels go in runs of at most ND per call
    and every (profile, ray) is its own package, so the oracle's threads
    share the work."""
    chans = list(range(ND_WIDE)) if chans is None else list(chans)
    runs, start = [], 0
    for i in range(1, len(chans) + 1):
        if i == len(chans) or chans[i] != chans[i - 1] + 1 or i - start == abi.ND:
            runs.append(chans[start:i])
            start = i
    angles = list(range(len(ANGLES))) if angles is None else list(angles)
    g = {key: np.asarray(v)[angles] for key, v in _geom().items()}
    nray = len(angles) * len(atms)
    rad = np.zeros((nray, len(chans)))
    tau = np.zeros((nray, len(chans)))
    col = 0
    for run in runs:
        ctl = _ctl(tb, [NU[c] for c in run], formod, refrac)
        npk = nray
        obss = (abi.obs_t * npk)()
        arr = (abi.atm_t * npk)()
        for i in range(nray):
            k, j = i % len(atms), i // len(atms)
            obss[i] = scenes.make_obs({key: np.asarray(v)[j:j + 1] for key, v in g.items()})
            arr[i] = abi.copy_struct(atms[k])
        oracle.formod_batch(ctl, arr, obss)
        for i in range(nray):
            r, t = abi.obs_arrays(obss[i], ctl)
            rad[i, col:col + len(run)] = r[:, 0]
            tau[i, col:col + len(run)] = t[:, 0]
        col += len(run)
    return rad, tau


def _bt(rad):
    c1, c2 = 1.19104259e-8, 1.43877506
    nu = np.asarray(NU)[None, :]
    return c2 * nu / np.log1p(c1 * nu ** 3 / rad)


@pytest.mark.parametrize("formod", [1, 0], ids=["ega", "cga"])
def test_wide_forward_matches_oracle_slices(gpu, oracle, scene, formod):
    torch = pytest.importorskip("torch")
    atms = _atms()
    ref_rad, ref_tau = _oracle_forward(oracle, scene, atms, formod)
    assert np.isfinite(ref_rad).all() and (ref_rad > 0).all()
    a_d, prof_d, geom_d, nray = _device_batch(torch, atms)
    ctx = gpu.Context(_ctl(scene, NU[:1], formod), nu=NU)
    try:
        assert ctx.nd == ND_WIDE
        for prec in (abi.JUR_FP64, abi.JUR_FP32):
            rad = torch.zeros(nray, ND_WIDE, dtype=torch.float64, device="cuda:0")
            tau = torch.zeros_like(rad)
            ctx.formod_device(NPROF, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(),
                              tau.data_ptr(), precision=prec)
            torch.cuda.synchronize()
            r, t = rad.cpu().numpy(), tau.cpu().numpy()
            if prec == abi.JUR_FP64:
                np.testing.assert_allclose(r, ref_rad, rtol=1e-10, atol=0)
                np.testing.assert_allclose(t, ref_tau, rtol=1e-10, atol=1e-300)
            else:
                err = float(np.max(np.abs(_bt(r) - _bt(ref_rad))))
                assert err < 0.01, err
    finally:
        ctx.close()


def _state_cols(atm):
    """Columns of T, H2O and O3 in atm2x order: T levels, then H2O, then O3."""
    return {"T": np.arange(atm.np), "H2O": atm.np + np.arange(atm.np), "O3": 2 * atm.np + np.arange(atm.np)}


def test_wide_adjoint_t_h2o_o3_matches_oracle_central_fd(gpu, oracle, scene):
    """K of the wide pass (n = 243) against the oracle's central differences
    on both channel slices: along each block's all-level direction
    (T: 1 K everywhere, vmr: the profile itself) and on single columns.
    Path held fixed (refraction off), as the adjoint assumes."""
    torch = pytest.importorskip("torch")
    atms = _atms()
    ctl = _ctl(scene, NU[:1], 1, refrac=0)
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    for g in ("H2O", "O3"):
        ig = GASES.index(g)
        ctl.retq_zmin[ig], ctl.retq_zmax[ig] = -100, 100
    xm, n = gpu.state_map(ctl, atms[0])
    assert n == 243
    cols = _state_cols(atms[0])
    a_d, prof_d, geom_d, nray = _device_batch(torch, atms)
    ctx = gpu.Context(ctl, nu=NU)
    ks = {}
    try:
        for prec in (abi.JUR_FP64, abi.JUR_FP32):
            rad = torch.zeros(nray, ND_WIDE, dtype=torch.float64, device="cuda:0")
            tau = torch.zeros_like(rad)
            kd = torch.zeros(nray, ND_WIDE, n, dtype=torch.float64, device="cuda:0")
            pmap = torch.zeros(NPROF, dtype=torch.int32, device="cuda:0")
            ctx.jacobian_device(NPROF, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(),
                                tau.data_ptr(), xm[None, :], pmap.data_ptr(), n, kd.data_ptr(), precision=prec)
            torch.cuda.synchronize()
            ks[prec] = kd.cpu().numpy()
    finally:
        ctx.close()
    # central differences of the oracle on channels at both sides of the ND
    # boundary and at the top of the band, two view angles, along each
    # block's all-level direction (T: 1 K, vmr: the profile itself) and a
    # seeded random one (weights 0.5..1.5), the better of two steps.
    # The forward model is continuous only to the float slopes of its table
    # brackets (a bracket change moves the emissivity by ~1e-7 of the
    # bracket's rise), so a central difference whose step straddles a knot
    # picks up jump / 2h; the expected error does not shrink with h and
    # sits near 1e-4 of the derivative for T at 2200 cm^-1 and, for the
    # optically thin O3, near 1e-3.  The bounds are set above those floors;
    # the narrow-context comparison below pins the wide pass exactly.
    chans = list(range(118, 138)) + list(range(150, ND_WIDE))
    angles = [0, 3]
    rows = [j * NPROF + k for j in angles for k in range(NPROF)]
    rng = np.random.default_rng(11)
    dirs = []
    for block, steps, tol64, tol32 in (("T", (1e-4, 1e-5), 3e-4, 1e-3), ("H2O", (1e-4, 1e-5), 3e-5, 1e-3),
                                       ("O3", (1e-5, 1e-6), 3e-3, 3e-3)):
        for wname, wts in (("all", np.ones(81)), ("random", rng.uniform(0.5, 1.5, 81))):
            dirs.append((block, wname, wts, steps, tol64, tol32))

    def perturbed(atm, block, wts, h):
        a = abi.copy_struct(atm)
        for ip in range(81):
            if block == "T":
                a.t[ip] += h * wts[ip]
            else:
                ig = GASES.index(block)
                a.q[ig][ip] += h * wts[ip] * atm.q[ig][ip]
        return a

    pert = [perturbed(a, block, wts, sgn * h) for block, _, wts, steps, _, _ in dirs for h in steps
            for sgn in (1, -1) for a in atms]
    y = _oracle_forward(oracle, scene, pert, 1, refrac=0, chans=chans, angles=angles)[0]
    na = len(pert)
    errs = []
    ia0 = 0
    for block, wname, wts, steps, tol64, tol32 in dirs:
        v = np.zeros(n)
        for ip in range(81):
            v[cols[block][ip]] = wts[ip] * (1.0 if block == "T" else atms[0].q[GASES.index(block)][ip])
        fds = []
        for h in steps:
            sel_p = [ia * na + ia0 * NPROF + k for ia in range(len(angles)) for k in range(NPROF)]
            sel_m = [ia * na + (ia0 + 1) * NPROF + k for ia in range(len(angles)) for k in range(NPROF)]
            fds.append((y[sel_p] - y[sel_m]) / (2 * h))
            ia0 += 2
        for prec, tol in ((abi.JUR_FP64, tol64), (abi.JUR_FP32, tol32)):
            kv = ks[prec][rows][:, chans] @ v
            best = min(float(np.max(np.abs(kv - fd))) / float(np.max(np.abs(fd))) for fd in fds)
            errs.append((best < tol, block, wname, prec, best, tol))
    print("\n".join(str(e) for e in errs))
    assert all(e[0] for e in errs), [e for e in errs if not e[0]]
    # the wide pass's K equals that of a context of the first ND channels
    # (the path every other adjoint test covers) on those channels
    ctl_n = _ctl(scene, NU[:abi.ND], 1, refrac=0)
    ctl_n.rett_zmin, ctl_n.rett_zmax = -100, 100
    for g in ("H2O", "O3"):
        ig = GASES.index(g)
        ctl_n.retq_zmin[ig], ctl_n.retq_zmax[ig] = -100, 100
    ctx = gpu.Context(ctl_n)
    try:
        rad = torch.zeros(nray, abi.ND, dtype=torch.float64, device="cuda:0")
        tau = torch.zeros_like(rad)
        kd = torch.zeros(nray, abi.ND, n, dtype=torch.float64, device="cuda:0")
        pmap = torch.zeros(NPROF, dtype=torch.int32, device="cuda:0")
        ctx.jacobian_device(NPROF, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(),
                            tau.data_ptr(), xm[None, :], pmap.data_ptr(), n, kd.data_ptr(),
                            precision=abi.JUR_FP64)
        torch.cuda.synchronize()
        kn = kd.cpu().numpy()
    finally:
        ctx.close()
    kw = ks[abi.JUR_FP64][:, :abi.ND]
    # T columns are summed with atomics by the emitter chains (last-bit order)
    np.testing.assert_allclose(kw, kn, rtol=1e-12, atol=1e-12 * np.abs(kn).max())
