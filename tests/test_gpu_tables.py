"""Tables from device memory (jur_ctx_create_device_tables) against the same
tables read from files (read_tbl, src/jurassic.c:4949-5110) and against the
oracle's formod on the same atmospheres and rays.

The device path takes the arrays of the reference's tbl_t (channel-major,
capacities TBLNP x TBLNT x TBLNU) as they are, as read_tbl takes a binary
table, so a context built from them must give bit-identical radiances and
transmittances to one built from the files: FP64 and FP32 paths, a missing
table (np < 0), a table with a decreasing column (reference bisection
instead of the local walk), filter functions from memory, and the adjoint.
All calls go through the C-ABI.
"""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from jurassic_amd import abi, scenes

pytestmark = pytest.mark.gpu

TP, TT, TU = abi.TBLNP, abi.TBLNT, abi.TBLNU


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from jurassic_amd import gpu as G
    G.lib()
    return G


def _tbl_arrays(ctl, torch):
    """tbl_t arrays of the ctl's binary table files, on the device."""
    nd, ng = ctl.nd, ctl.ng
    np_ = np.full((nd, ng), -1, dtype=np.int32)
    nt = np.zeros((nd, ng, TP), dtype=np.int32)
    nu = np.zeros((nd, ng, TP, TT), dtype=np.int32)
    p = np.zeros((nd, ng, TP))
    t = np.zeros((nd, ng, TP, TT))
    u = np.zeros((nd, ng, TP, TT, TU), dtype=np.float32)
    eps = np.zeros((nd, ng, TP, TT, TU), dtype=np.float32)
    base = ctl.tblbase.decode()
    for id_ in range(nd):
        for ig in range(ng):
            fn = scenes.table_name(base, ctl.nu[id_], ctl.emitter[ig].value.decode(), 2)
            if not os.path.exists(fn):
                continue
            pg, ts, cols = scenes.read_table_bin(fn)
            np_[id_, ig] = len(pg)
            p[id_, ig, :len(pg)] = pg
            for ip in range(len(pg)):
                nt[id_, ig, ip] = len(ts[ip])
                t[id_, ig, ip, :len(ts[ip])] = ts[ip]
                for it, (uu, ee) in enumerate(cols[ip]):
                    nu[id_, ig, ip, it] = len(uu)
                    u[id_, ig, ip, it, :len(uu)] = uu
                    eps[id_, ig, ip, it, :len(uu)] = ee
    dev = torch.device("cuda:0")
    return {k: torch.from_numpy(v).to(dev) for k, v in
            dict(np=np_, nt=nt, nu=nu, p=p, t=t, u=u, eps=eps).items()}


def _batch(ctl, nprof, geom_fn, torch):
    """nprof atmospheres x the rays of geom_fn(), ray-major, on the device."""
    atms = (abi.atm_t * nprof)()
    for k in range(nprof):
        atms[k] = scenes.make_profile_atm([ctl.emitter[i].value.decode() for i in range(ctl.ng)],
                                          iprof=k)
    geom = geom_fn()
    nr = len(geom["obsz"])
    nray = nprof * nr
    prof = np.array([i % nprof for i in range(nray)], dtype=np.int32)
    g = np.zeros((nray, 7))
    for i in range(nray):
        j = i // nprof
        g[i] = [0.0, geom["obsz"][j], geom["obslon"][j], geom["obslat"][j], geom["vpz"][j],
                geom["vplon"][j], geom["vplat"][j]]
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(np.frombuffer(bytes(atms), dtype=np.uint8).copy()).to(dev)
    return a_d, torch.from_numpy(prof).to(dev), torch.from_numpy(g).to(dev), nray


def _forward(ctx, ctl, batch, precision, torch):
    a_d, prof_d, geom_d, nray = batch
    rad = torch.zeros(nray, ctx.nd, dtype=torch.float64, device="cuda:0")
    tau = torch.zeros_like(rad)
    tp = torch.zeros(nray, 3, dtype=torch.float64, device="cuda:0")
    ctx.formod_device(a_d.numel() // C.sizeof(abi.atm_t), a_d.data_ptr(), nray, prof_d.data_ptr(),
                      geom_d.data_ptr(), rad.data_ptr(), tau.data_ptr(), tp.data_ptr(), precision=precision)
    torch.cuda.synchronize()
    return rad.cpu().numpy(), tau.cpu().numpy(), tp.cpu().numpy()


def _limb_batch(ctl, torch, nprof=3):
    return _batch(ctl, nprof, lambda: scenes.limb_geometry(800.0, scenes.limb_tangent_heights()[::4]), torch)


def _oracle_limb(ctl, nprof=3):
    """The oracle's formod on the atmospheres and rays of _limb_batch, in
    the batch's ray-major order: [nray][nd] radiance and transmittance."""
    import oracle as O
    geom = scenes.limb_geometry(800.0, scenes.limb_tangent_heights()[::4])
    nr = len(geom["obsz"])
    rad = np.zeros((nr * nprof, ctl.nd))
    tau = np.zeros_like(rad)
    for k in range(nprof):
        atm = scenes.make_profile_atm([ctl.emitter[i].value.decode() for i in range(ctl.ng)], iprof=k)
        obs = scenes.make_obs(geom)
        O.formod(ctl, atm, obs)
        r, t = abi.obs_arrays(obs, ctl)
        rad[k::nprof] = r.T
        tau[k::nprof] = t.T
    return rad, tau


def _bt(rad, nu):
    c1, c2 = 1.19104259e-8, 1.43877506
    nu = np.asarray(nu)[None, :]
    return c2 * nu / np.log1p(c1 * nu ** 3 / rad)


def _same_as_files(gpu, ctl, torch, filters=None, batch=None):
    """Bit-identical to the context built from the files, and equal to the
    oracle's formod on the same rays (FP64 to 1e-10 relative, FP32 to
    0.01 K in brightness temperature)."""
    nu = [ctl.nu[i] for i in range(ctl.nd)]
    tabs = _tbl_arrays(ctl, torch)
    own_batch = batch is None
    batch = batch or _limb_batch(ctl, torch)
    c_file = gpu.Context(ctl)
    c_dev = gpu.Context.from_device_tables(ctl, nu, tabs, filters=filters)
    ref = _oracle_limb(ctl) if own_batch else None
    try:
        for prec in (abi.JUR_FP64, abi.JUR_FP32):
            r0, t0, p0 = _forward(c_file, ctl, batch, prec, torch)
            r1, t1, p1 = _forward(c_dev, ctl, batch, prec, torch)
            assert np.array_equal(r0, r1) and np.array_equal(t0, t1) and np.array_equal(p0, p1), prec
            assert np.isfinite(r0).all() and (r0 > 0).any()
            if ref is None:
                continue
            if prec == abi.JUR_FP64:
                np.testing.assert_allclose(r1, ref[0], rtol=1e-10, atol=0)
                np.testing.assert_allclose(t1, ref[1], rtol=1e-10, atol=1e-300)
            else:
                assert float(np.max(np.abs(_bt(r1, nu) - _bt(ref[0], nu)))) < 0.01
    finally:
        c_file.close()
        c_dev.close()


def test_device_tables_match_files(gpu):
    torch = pytest.importorskip("torch")
    _same_as_files(gpu, scenes.limb_3gas_ctl(), torch)


def test_device_tables_filters_from_memory(gpu):
    torch = pytest.importorskip("torch")
    ctl = scenes.limb_3gas_ctl()
    off, xs, ys = [0], [], []
    for i in range(ctl.nd):
        x, y = scenes.read_filter("%s_%.4f.filt" % (ctl.tblbase.decode(), ctl.nu[i]))
        xs.append(x)
        ys.append(y)
        off.append(off[-1] + len(x))
    _same_as_files(gpu, ctl, torch, filters=(np.array(off), np.concatenate(xs), np.concatenate(ys)))


def _scene_copy(tmp_path, ctl, name):
    """Copy the ctl's tables and filters into a fresh directory."""
    src = os.path.dirname(ctl.tblbase.decode())
    dst = os.path.join(str(tmp_path), name)
    shutil.copytree(src, dst)
    c2 = abi.copy_struct(ctl)
    c2.tblbase = os.path.join(dst, os.path.basename(ctl.tblbase.decode())).encode()
    return c2


def test_device_tables_missing_table(gpu, tmp_path):
    torch = pytest.importorskip("torch")
    ctl = _scene_copy(tmp_path, scenes.limb_3gas_ctl(), "missing")
    os.remove(scenes.table_name(ctl.tblbase.decode(), ctl.nu[3], "O3", 2))
    _same_as_files(gpu, ctl, torch)


def test_device_tables_decreasing_column(gpu, tmp_path):
    torch = pytest.importorskip("torch")
    ctl = _scene_copy(tmp_path, scenes.limb_3gas_ctl(), "nonmono")
    fn = scenes.table_name(ctl.tblbase.decode(), ctl.nu[1], "H2O", 2)
    pg, ts, cols = scenes.read_table_bin(fn)
    tg = np.array(ts[0])
    u = cols[0][0][0]
    eps = np.array([[cols[ip][it][1] for it in range(len(ts[ip]))] for ip in range(len(pg))],
                   dtype=np.float32)
    eps[5, 3, 40] = eps[5, 3, 38]  # a dip: column (5, 3) decreases at knot 40
    eps[5, 3, 39] = eps[5, 3, 41]
    scenes.write_table_bin(fn, pg, tg, u, eps)
    _same_as_files(gpu, ctl, torch)


def test_device_tables_adjoint(gpu):
    torch = pytest.importorskip("torch")
    ctl = scenes.nadir_jac_ctl(nd=8)
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    ctl.retq_zmin[1], ctl.retq_zmax[1] = -100, 100
    nu = [ctl.nu[i] for i in range(ctl.nd)]
    tabs = _tbl_arrays(ctl, torch)
    a_d, prof_d, geom_d, nray = _batch(ctl, 2, lambda: scenes.nadir_geometry(705.0, [0.0, 25.0]), torch)
    atm0 = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=0)
    xm, n = gpu.state_map(ctl, atm0)
    pmap = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    out = []
    for make in (lambda: gpu.Context(ctl), lambda: gpu.Context.from_device_tables(ctl, nu, tabs)):
        ctx = make()
        try:
            rad = torch.zeros(nray, ctl.nd, dtype=torch.float64, device="cuda:0")
            tau = torch.zeros_like(rad)
            kd = torch.zeros(nray, ctl.nd, n, dtype=torch.float64, device="cuda:0")
            ctx.jacobian_device(2, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(),
                                tau.data_ptr(), xm[None, :], pmap.data_ptr(), n, kd.data_ptr(),
                                precision=abi.JUR_FP32)
            torch.cuda.synchronize()
            out.append((rad.cpu().numpy(), kd.cpu().numpy()))
        finally:
            ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    # T and p columns are summed with atomics (order-dependent last bit)
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-12, atol=1e-12 * np.abs(out[0][1]).max())
    assert np.abs(out[0][1]).max() > 0


def test_device_tables_capacity_errors(gpu):
    torch = pytest.importorskip("torch")
    ctl = scenes.limb_3gas_ctl()
    tabs = _tbl_arrays(ctl, torch)
    tabs["np"][0, 0] = TP + 1
    with pytest.raises(gpu.JurassicError) as e:
        gpu.Context.from_device_tables(ctl, [ctl.nu[i] for i in range(ctl.nd)], tabs)
    assert "pressure levels" in str(e.value)
