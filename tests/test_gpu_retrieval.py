"""GPU optimal estimation against the oracle's restatement of the reference
LM retrieval (optimal_estimation, src/retrieval.c:448-666).

With kernel()'s finite-difference Jacobian (JUR_JAC_FD) the GPU retrieval
runs the same iteration as the oracle on the same FP64 forward model, so the
cost history and the retrieved state agree to rounding amplified over the
iterations.  With the adjoint Jacobian (path held fixed, exact derivative
instead of kernel()'s secant) the iteration differs, and the check is that
it reaches the same minimum.  A batch of retrievals of different state
sizes gives the results of the single runs.
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


def _problem(iprof=3, nrays=6, nlev=81, gases_ret=(1,)):
    """Limb scene (3 emitters, 12 channels), brightness temperatures of a
    perturbed atmosphere as the measurement, the unperturbed profile as a
    priori; state: T and the volume mixing ratios of gases_ret."""
    ctl, _, obs = scenes.limb_3gas(iprof, nrays=nrays)
    ctl.write_bbt = 1
    ctl.rett_zmin, ctl.rett_zmax = -100, 100
    for ig in gases_ret:
        ctl.retq_zmin[ig], ctl.retq_zmax[ig] = -100, 100
    z = np.arange(nlev, dtype=float)
    atm_true = scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=iprof, z=z)
    atm_apr = scenes.make_profile_atm(["CO2", "H2O", "O3"], z=z, perturb=False)
    ret = abi.ret_t()
    ret.kernel_recomp = 3
    ret.conv_itmax = 10
    ret.conv_dmin = 0.1
    for i in range(ctl.nd):
        ret.err_noise[i] = np.sqrt(0.1)
    ret.err_temp, ret.err_temp_cz, ret.err_temp_ch = 2.0, 5.0, 500.0
    for ig in range(3):
        ret.err_q[ig], ret.err_q_cz[ig], ret.err_q_ch[ig] = 20.0, 5.0, 500.0
    return ctl, atm_true, atm_apr, obs, ret


def _measure(oracle, ctl, atm_true, obs):
    o = abi.copy_struct(obs)
    oracle.formod(ctl, abi.copy_struct(atm_true), o)
    return o


def _state(ctl, atm, gases_ret=(1,)):
    x = [np.ctypeslib.as_array(atm.t)[:atm.np]]
    for ig in gases_ret:
        x.append(np.ctypeslib.as_array(atm.q[ig])[:atm.np])
    return x


def test_oe_fd_matches_oracle(gpu, oracle):
    ctl, atm_true, atm_apr, obs, ret = _problem()
    meas = _measure(oracle, ctl, atm_true, obs)
    a0, _, h0, n0 = oracle.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr))
    a1, _, h1, n1 = gpu.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr),
                                           method=abi.JUR_JAC_FD, precision=abi.JUR_FP64)
    assert n1 == n0 and len(h1) == len(h0)
    np.testing.assert_allclose(h1, h0, rtol=1e-6)
    for v1, v0 in zip(_state(ctl, a1), _state(ctl, a0)):
        np.testing.assert_allclose(v1, v0, rtol=1e-6, atol=1e-12)
    assert h1[-1] < 0.05 * h1[0]


@pytest.mark.parametrize("precision", [abi.JUR_FP64, abi.JUR_FP32])
def test_oe_adjoint_reaches_the_minimum(gpu, oracle, precision):
    ctl, atm_true, atm_apr, obs, ret = _problem()
    meas = _measure(oracle, ctl, atm_true, obs)
    a0, _, h0, _ = oracle.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr))
    a1, _, h1, _ = gpu.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr),
                                           method=abi.JUR_JAC_ADJOINT, precision=precision)
    # same a priori cost, a final cost as low as the reference iteration's
    assert h1[0] == pytest.approx(h0[0], rel=1e-6 if precision == abi.JUR_FP64 else 1e-3)
    assert h1[-1] <= 1.1 * h0[-1] + 1e-3
    t0, t1 = _state(ctl, a0)[0], _state(ctl, a1)[0]
    assert np.abs(t1 - t0).max() < 0.3


def test_oe_batch_matches_single(gpu, oracle):
    probs = [_problem(iprof=3), _problem(iprof=4, nlev=61), _problem(iprof=5, nrays=4, gases_ret=(0, 1))]
    for method in (abi.JUR_JAC_FD, abi.JUR_JAC_ADJOINT):
        # one ctl: the state windows of the third problem (T, CO2, H2O) and
        # the 61-level z grid of the second give three state sizes
        ctl = probs[2][0]
        meas, apr = [], []
        for (_, at, aa, ob, _) in probs:
            meas.append(_measure(oracle, ctl, at, ob))
            apr.append(aa)
        ret = probs[0][4]
        single = [gpu.optimal_estimation(ret, ctl, abi.copy_struct(m), abi.copy_struct(a), method=method)
                  for m, a in zip(meas, apr)]
        batch = gpu.optimal_estimation(ret, ctl, [abi.copy_struct(m) for m in meas],
                                       [abi.copy_struct(a) for a in apr], method=method)
        for i, s in enumerate(single):
            assert batch[3][i] == s[3]
            np.testing.assert_allclose(batch[2][i], s[2], rtol=1e-9)
            for v1, v0 in zip(_state(ctl, batch[0][i], (0, 1)), _state(ctl, s[0], (0, 1))):
                np.testing.assert_allclose(v1, v0, rtol=1e-9, atol=1e-15)


def test_oe_adjoint_hydrostatic_unsupported(gpu, oracle):
    ctl, atm_true, atm_apr, obs, ret = _problem(nrays=2)
    ctl.hydz = 0.0
    meas = _measure(oracle, ctl, atm_true, obs)
    with pytest.raises(gpu.JurassicError) as e:
        gpu.optimal_estimation(ret, ctl, meas, atm_apr, method=abi.JUR_JAC_ADJOINT)
    assert e.value.code == 6


def test_oe_error_analysis_matches_oracle(gpu, oracle):
    """Error analysis (src/retrieval.c:668-726) of the GPU retrieval against
    the oracle's on the same problem (finite-difference Jacobians, FP64: the
    same iteration, so the same final Jacobian to the retrieval's own
    agreement), and the atmosphere outputs against the state vectors."""
    ctl, atm_true, atm_apr, obs, ret = _problem()
    for i in range(ctl.nd):
        ret.err_formod[i] = 1.0 if i % 2 else 0.0
    meas = _measure(oracle, ctl, atm_true, obs)
    *_, a0 = oracle.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr), analysis=True)
    atm_i, _, _, _, a1 = gpu.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr),
                                                method=abi.JUR_JAC_FD, precision=abi.JUR_FP64, analysis=True)
    for k in ("cov", "corr", "gain", "avk", "err_total", "err_noise", "err_formod", "cont", "res"):
        scale = np.abs(a0[k]).max()
        np.testing.assert_allclose(a1[k], a0[k], rtol=1e-5, atol=1e-7 * scale, err_msg=k)
    # atm_err_*, atm_cont, atm_res: the state vectors at the state elements
    # (T and H2O here), zero elsewhere (copy_atm with init = 1)
    xm, n = gpu.state_map(ctl, atm_apr)
    for name, vec in (("atm_err_total", "err_total"), ("atm_err_noise", "err_noise"),
                      ("atm_err_formod", "err_formod"), ("atm_cont", "cont"), ("atm_res", "res")):
        a = a1[name]
        assert a.np == atm_i.np
        for ip in range(a.np):
            assert a.t[ip] == a1[vec][xm[abi.xmap_lev(1, ip)]]
            assert a.q[1][ip] == a1[vec][xm[abi.xmap_lev(3, ip)]]
            assert a.p[ip] == 0 and a.q[0][ip] == 0 and a.q[2][ip] == 0
            assert a.z[ip] == atm_i.z[ip]


@pytest.mark.parametrize("precision", [abi.JUR_FP64, abi.JUR_FP32])
def test_oe_error_analysis_adjoint(gpu, oracle, precision):
    """With adjoint Jacobians: cov is symmetric, the total error squared is
    its diagonal, the noise error is below the total error (G S_noise G^T <=
    G S_eps G^T = cov - cov Sa^-1 cov), and A = I - cov Sa^-1 has its
    eigenvalues in [0, 1), so 0 < trace(A) < n."""
    ctl, atm_true, atm_apr, obs, ret = _problem()
    meas = _measure(oracle, ctl, atm_true, obs)
    *_, a1 = gpu.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr),
                                    method=abi.JUR_JAC_ADJOINT, precision=precision, analysis=True)
    cov, avk = a1["cov"], a1["avk"]
    assert np.allclose(cov, cov.T, rtol=1e-10, atol=1e-14 * np.abs(cov).max())
    np.testing.assert_allclose(a1["err_total"] ** 2, np.diag(cov), rtol=1e-12)
    assert np.all(a1["err_total"] > 0)
    assert np.all(a1["err_noise"] <= a1["err_total"] * (1 + 1e-9))
    assert 0 < np.trace(avk) < avk.shape[0]


@pytest.mark.parametrize("gases_ret", [(0,), (0, 1)], ids=["n512", "n768"])
def test_oe_fd_large_state_matches_oracle(gpu, oracle, gases_ret):
    """State sizes at and above the staged dense-algebra limit (CHOL_NMAX =
    512): T and one or two vmr profiles on 256 levels (n = 512 and 768), two
    nadir rays; the FD iteration against the oracle's as above."""
    gases, nus = ["CO2", "H2O"], [667.8, 700.0]
    tb = scenes.write_scene_tables(scenes.scene_dir("oe_large"), gases, nus)
    ctl = abi.make_ctl(gases, nus, tblbase=tb, write_bbt=1, rett_zmin=-100, rett_zmax=100)
    for ig in gases_ret:
        ctl.retq_zmin[ig], ctl.retq_zmax[ig] = -100, 100
    z = np.linspace(0.0, 80.0, 256)
    atm_true = scenes.make_profile_atm(gases, iprof=5, z=z)
    atm_apr = scenes.make_profile_atm(gases, z=z, perturb=False)
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, [0.0, 30.0]))
    ret = abi.ret_t()
    ret.kernel_recomp = 3
    ret.conv_itmax = 4
    ret.conv_dmin = 1e-4
    for i in range(ctl.nd):
        ret.err_noise[i] = np.sqrt(0.1)
    ret.err_temp, ret.err_temp_cz, ret.err_temp_ch = 2.0, 5.0, 500.0
    for ig in range(2):
        ret.err_q[ig], ret.err_q_cz[ig], ret.err_q_ch[ig] = 20.0, 5.0, 500.0
    meas = _measure(oracle, ctl, atm_true, obs)
    m, n = oracle.sizes(ctl, atm_apr, meas)
    assert n == 256 * (1 + len(gases_ret))
    a0, _, h0, n0 = oracle.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr))
    a1, _, h1, n1 = gpu.optimal_estimation(ret, ctl, abi.copy_struct(meas), abi.copy_struct(atm_apr),
                                           method=abi.JUR_JAC_FD, precision=abi.JUR_FP64)
    assert n0 >= 3
    assert n1 == n0 and len(h1) == len(h0)
    np.testing.assert_allclose(h1, h0, rtol=1e-6)
    for v1, v0 in zip(_state(ctl, a1, gases_ret), _state(ctl, a0, gases_ret)):
        np.testing.assert_allclose(v1, v0, rtol=1e-6, atol=1e-12)
    assert h1[-1] < h1[0]
