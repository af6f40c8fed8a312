"""jur_ctx_formod_device_noalloc, the entry point that allocates no device
memory (its work buffers are slices of the library's fixed scratch array),
against jur_ctx_formod_device on the same batches: bit-identical radiances,
transmittances and tangent points, FP64 and FP32, on two calls in a row
(the second smaller), and a clean JUR_ERR_MEM for a batch whose per-ray
buffers exceed the array, after which calls still work.  Through the C-ABI.
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


def _batch(torch, ctl, nprof, heights):
    geom = scenes.limb_geometry(800.0, heights)
    nr = len(heights)
    nray = nprof * nr
    atms = (abi.atm_t * nprof)(*[scenes.make_profile_atm(["CO2", "H2O", "O3"], iprof=k) for k in range(nprof)])
    prof = np.array([i % nprof for i in range(nray)], dtype=np.int32)
    g = np.zeros((nray, 7))
    for i in range(nray):
        j = i // nprof
        g[i] = [0.0, geom["obsz"][j], geom["obslon"][j], geom["obslat"][j], geom["vpz"][j], geom["vplon"][j],
                geom["vplat"][j]]
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(np.frombuffer(bytes(atms), dtype=np.uint8).copy()).to(dev)
    return nprof, a_d, torch.from_numpy(prof).to(dev), torch.from_numpy(g).to(dev), nray


def _run(torch, ctx, fn, b, prec):
    nprof, a_d, prof_d, geom_d, nray = b
    rad = torch.full((nray, ctx.nd), -1.0, dtype=torch.float64, device="cuda:0")
    tau = torch.full_like(rad, -1.0)
    tp = torch.full((nray, 3), -1.0, dtype=torch.float64, device="cuda:0")
    fn(nprof, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(), rad.data_ptr(), tau.data_ptr(),
       tp.data_ptr(), precision=prec)
    torch.cuda.synchronize()
    return rad.cpu().numpy(), tau.cpu().numpy(), tp.cpu().numpy()


def test_noalloc_equals_formod_device(gpu):
    torch = pytest.importorskip("torch")
    ctl = scenes.limb_3gas_ctl()
    ctx = gpu.Context(ctl)
    th = scenes.limb_tangent_heights()
    try:
        # 24 x 32 rays = 12 ray blocks: more than one chunk of the scratch
        # array (about 6 blocks of limb LOS), then a smaller second call
        for prec in (abi.JUR_FP64, abi.JUR_FP32):
            for nprof, heights in ((24, th), (2, th[::4])):
                b = _batch(torch, ctl, nprof, heights)
                ref = _run(torch, ctx, ctx.formod_device, b, prec)
                got = _run(torch, ctx, ctx.formod_device_noalloc, b, prec)
                for r, g, what in zip(ref, got, ("rad", "tau", "tp")):
                    assert np.array_equal(r, g), (prec, nprof, what)
                assert np.isfinite(ref[0]).all() and (ref[0] > 0).all()
    finally:
        ctx.close()


def test_noalloc_rejects_batch_beyond_scratch(gpu):
    """About 2.1 million rays: the per-ray state alone exceeds the 256 MiB
    array, so the call returns JUR_ERR_MEM before tracing any ray; a normal
    call works afterwards."""
    torch = pytest.importorskip("torch")
    ctl = scenes.limb_3gas_ctl()
    ctx = gpu.Context(ctl)
    try:
        small = _batch(torch, ctl, 2, scenes.limb_tangent_heights()[::8])
        nprof, a_d, _, _, _ = small
        nray = 2 ** 21
        dev = torch.device("cuda:0")
        prof_d = torch.zeros(nray, dtype=torch.int32, device=dev)
        geom_d = torch.zeros(nray, 7, dtype=torch.float64, device=dev)
        rad = torch.zeros(nray, ctx.nd, dtype=torch.float64, device=dev)
        tau = torch.zeros_like(rad)
        with pytest.raises(gpu.JurassicError) as e:
            ctx.formod_device_noalloc(nprof, a_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(),
                                      rad.data_ptr(), tau.data_ptr(), precision=abi.JUR_FP32)
        assert e.value.code == abi.JUR_ERR_MEM and "scratch" in str(e.value)
        del prof_d, geom_d, rad, tau
        ref = _run(torch, ctx, ctx.formod_device, small, abi.JUR_FP32)
        got = _run(torch, ctx, ctx.formod_device_noalloc, small, abi.JUR_FP32)
        assert all(np.array_equal(r, g) for r, g in zip(ref, got))
    finally:
        ctx.close()
