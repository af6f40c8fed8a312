"""The request server: every result is the same alone and in a mix, and
equals the direct entry points (formod, kernel, optimal_estimation).

Each lane of a device pass computes one ray from its own inputs, and a
retrieval's dense algebra runs in its own workgroup, so batching changes no
arithmetic: forward models and Jacobian rows agree to the last bit except
where the adjoint adds the T and p entries of one K row from several
emitter chains in a different order (relative 1e-12 bound), which a
retrieval then carries into its final state (1e-9).
"""
import numpy as np
import pytest

from jurassic_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from jurassic_amd import gpu as G
    G.lib()
    return G


def _ret(nd):
    ret = abi.ret_t()
    ret.kernel_recomp = 3
    ret.conv_itmax = 6
    ret.conv_dmin = 0.1
    for i in range(nd):
        ret.err_noise[i] = np.sqrt(0.1)
    ret.err_temp, ret.err_temp_cz, ret.err_temp_ch = 2.0, 5.0, 500.0
    for ig in range(3):
        ret.err_q[ig], ret.err_q_cz[ig], ret.err_q_ch[ig] = 20.0, 5.0, 500.0
    return ret


def _mix(oracle):
    """A server ctl (limb channels, brightness temperatures) and a mixed set
    of requests: forward, Jacobian (T + H2O), retrieval (T), limb and nadir."""
    ctl = scenes.limb_3gas_ctl()
    ctl.write_bbt = 1
    reqs = []
    limb = scenes.make_obs(scenes.limb_geometry(800.0, scenes.limb_tangent_heights()[:5]))
    nadir = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0, 20.0, 40.0]))
    for i, geo in enumerate([limb, nadir, limb]):
        atm = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=10 + i)
        reqs.append(dict(kind=abi.JUR_REQ_FORWARD, atm=atm, obs=abi.copy_struct(geo)))
        reqs.append(dict(kind=abi.JUR_REQ_JACOBIAN, atm=atm, obs=abi.copy_struct(geo), rett=(-100, 100),
                         retq={1: (-100, 100)}))
        meas = abi.copy_struct(geo)
        oracle.formod(ctl, abi.copy_struct(atm), meas)
        apr = scenes.make_profile_atm(["CO2", "H2O", "O3"], perturb=False)
        reqs.append(dict(kind=abi.JUR_REQ_RETRIEVAL, atm=apr, obs=meas, rett=(-100, 100)))
    return ctl, reqs


def _run(gpu, ctl, ret, reqs, precision, lanes=0, forward_lane=0, path_split=0):
    srv = gpu.Server(ctl, ret, precision=precision, lanes=lanes, forward_lane=forward_lane,
                     path_split=path_split)
    try:
        keep = []
        ids = []
        for r in reqs:
            q = gpu.Server.request(r["kind"], r["atm"], r["obs"], rett=r.get("rett", (-999.0, -999.0)),
                                   retq=r.get("retq"))
            keep.append(q)
            ids.append(srv.submit(q))
        out = []
        for i in ids:
            res = srv.wait(i, 120.0)
            assert res.status == 0, res.message
            out.append((res, srv.fetch(i, res)))
        # worker seconds by phase, then the process's allocation seconds,
        # count, and the seconds of those that freed a smaller buffer
        prof = srv.profile()
        assert len(prof) == 7 and min(prof) >= 0
        assert prof[5] >= 1 and prof[6] <= prof[4]
        return out, srv.stats()
    finally:
        srv.close()


@pytest.mark.parametrize("precision", [abi.JUR_FP32, abi.JUR_FP64])
def test_server_alone_and_in_mix(gpu, oracle, precision):
    ctl, reqs = _mix(oracle)
    ret = _ret(ctl.nd)
    mixed, (rounds, max_active) = _run(gpu, ctl, ret, reqs, precision)
    assert max_active > 1  # the requests shared device passes
    for r, (res, (obs, atm, k, hist)) in zip(reqs, mixed):
        (res1, (obs1, atm1, k1, hist1)), = _run(gpu, ctl, ret, [r], precision)[0]
        r0, t0 = abi.obs_arrays(obs1, ctl)
        r1, t1 = abi.obs_arrays(obs, ctl)
        if r["kind"] == abi.JUR_REQ_RETRIEVAL:
            # the final state went through Jacobians whose T and p entries
            # were summed in a different order
            np.testing.assert_allclose(r1, r0, rtol=1e-9)
        else:
            assert np.array_equal(r0, r1) and np.array_equal(t0, t1), r["kind"]
        if r["kind"] == abi.JUR_REQ_JACOBIAN:
            np.testing.assert_allclose(k, k1, rtol=1e-12, atol=1e-12 * np.abs(k1).max())
        if r["kind"] == abi.JUR_REQ_RETRIEVAL:
            assert res.niter == res1.niter
            np.testing.assert_allclose(hist, hist1, rtol=1e-9)
            np.testing.assert_allclose(np.ctypeslib.as_array(atm.t)[:atm.np],
                                       np.ctypeslib.as_array(atm1.t)[:atm1.np], rtol=1e-9)


def test_server_matches_direct_calls(gpu, oracle):
    ctl, reqs = _mix(oracle)
    ret = _ret(ctl.nd)
    prec = abi.JUR_FP64
    mixed, _ = _run(gpu, ctl, ret, reqs, prec)
    for r, (res, (obs, atm, k, hist)) in zip(reqs, mixed):
        if r["kind"] == abi.JUR_REQ_FORWARD:
            # the server evaluates forward models in its Jacobian-mode pass,
            # whose forward sweep sums the terms in its own order
            o = abi.copy_struct(r["obs"])
            gpu.formod(ctl, abi.copy_struct(r["atm"]), o, prec)
            np.testing.assert_allclose(abi.obs_arrays(obs, ctl)[0], abi.obs_arrays(o, ctl)[0], rtol=1e-12)
        elif r["kind"] == abi.JUR_REQ_JACOBIAN:
            c2 = abi.copy_struct(ctl)
            c2.rett_zmin, c2.rett_zmax = -100, 100
            c2.retq_zmin[1], c2.retq_zmax[1] = -100, 100
            m, n = oracle.sizes(c2, r["atm"], r["obs"])
            kd = gpu.kernel(c2, abi.copy_struct(r["atm"]), abi.copy_struct(r["obs"]),
                            method=abi.JUR_JAC_ADJOINT, precision=prec, m=m, n=n)
            np.testing.assert_allclose(k, kd, rtol=1e-12, atol=1e-12 * np.abs(kd).max())
        else:
            c2 = abi.copy_struct(ctl)
            c2.rett_zmin, c2.rett_zmax = -100, 100
            a1, _, h1, n1 = gpu.optimal_estimation(ret, c2, abi.copy_struct(r["obs"]),
                                                   abi.copy_struct(r["atm"]), precision=prec)
            assert res.niter == n1
            np.testing.assert_allclose(hist, h1, rtol=1e-9)


def test_server_forward_lane(gpu, oracle):
    """With a forward lane, forward requests take the forward pass (the
    kernels of formod: equal to the last bit), and Jacobians and retrievals
    are those of the one-lane server; so are they with two general lanes,
    and with the limb requests on one or two lanes and the others on a lane
    of their own."""
    ctl, reqs = _mix(oracle)
    ret = _ret(ctl.nd)
    prec = abi.JUR_FP64
    base, _ = _run(gpu, ctl, ret, reqs, prec)
    for lanes, fl, ps in ((1, 1, 0), (2, 0, 0), (1, 0, 1), (2, 0, 1)):
        got, (rounds, _) = _run(gpu, ctl, ret, reqs, prec, lanes=lanes, forward_lane=fl, path_split=ps)
        assert rounds >= 1
        for r, (res, (obs, atm, k, hist)), (res0, (obs0, atm0, k0, hist0)) in zip(reqs, got, base):
            if r["kind"] == abi.JUR_REQ_FORWARD and fl:
                o = abi.copy_struct(r["obs"])
                gpu.formod(ctl, abi.copy_struct(r["atm"]), o, prec)
                assert np.array_equal(abi.obs_arrays(obs, ctl)[0], abi.obs_arrays(o, ctl)[0])
            else:
                np.testing.assert_allclose(abi.obs_arrays(obs, ctl)[0], abi.obs_arrays(obs0, ctl)[0],
                                           rtol=1e-12)
            if r["kind"] == abi.JUR_REQ_JACOBIAN:
                np.testing.assert_allclose(k, k0, rtol=1e-12, atol=1e-12 * np.abs(k0).max())
            if r["kind"] == abi.JUR_REQ_RETRIEVAL:
                assert res.niter == res0.niter
                np.testing.assert_allclose(hist, hist0, rtol=1e-9)


def test_server_reports_request_errors(gpu, oracle):
    ctl, reqs = _mix(oracle)
    srv = gpu.Server(ctl, _ret(ctl.nd))
    try:
        bad = abi.copy_struct(reqs[0]["obs"])
        bad.obsz[0] = -10.0  # observer below the surface
        q = gpu.Server.request(abi.JUR_REQ_FORWARD, reqs[0]["atm"], bad)
        good = gpu.Server.request(abi.JUR_REQ_FORWARD, reqs[0]["atm"], reqs[0]["obs"])
        i_bad, i_good = srv.submit(q), srv.submit(good)
        res = srv.wait(i_bad, 60.0)
        assert res.status == 4
        res2 = srv.wait(i_good, 60.0)
        assert res2.status == 0  # a failing request does not fail the others
        with pytest.raises(gpu.JurassicError):
            srv.fetch(i_bad, res)
    finally:
        srv.close()
