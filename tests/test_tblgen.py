"""tblgen (src/tblgen.c:44-140): emissivity rows from a monochromatic
optical-depth spectrum.

The oracle's restatement is pinned by the mathematics, since the reference
holds no tblgen fixtures:
  * a constant spectrum gives eps(u) = 1 - exp(-a u) whatever the filter,
    so the rows, their column densities and the termination are known in
    closed form;
  * a spectrum of two interleaved levels under a linear filter gives
    eps(u) = 1 - (W1 exp(-a1 u) + W2 exp(-a2 u)) / (W1 + W2), with the
    filter sums W1, W2 evaluated independently here (numpy).
The GPU entry point (jur_tblgen, through the C-ABI) is checked against the
oracle on line spectra, on a grid whose points fall on filter nodes, and
through the command-line mirror of the tool.
"""
import io
import contextlib

import numpy as np
import pytest

from jurassic_amd import tblgen as T

KB = 1.380658e-23


def uref(press, temp):
    return 1e-6 * press * 100 / (KB * temp) * 1000 / 1e4


def u_grid():
    u, out = 1.0, []
    while u <= 1e30:
        out.append(u)
        u *= 1.122
    return np.array(out)


def expected_rows(eps_of_u):
    """Rows the tool keeps for an emissivity curve, by its rules."""
    us, es, old = [], [], 0.0
    for u in u_grid():
        e = eps_of_u(u)
        if 1e-6 <= e <= 0.999999 and e > old:
            us.append(u)
            es.append(e)
        old = e
        if e > 0.999999:
            break
    return np.array(us), np.array(es)


def line_spectrum(n, nu0, nu1, seed=0, nlines=40):
    rng = np.random.default_rng(seed)
    nu = np.linspace(nu0, nu1, n)
    tau = np.full(n, 1e-4)
    for c, s, w in zip(rng.uniform(nu0, nu1, nlines), 10 ** rng.uniform(-2, 1, nlines),
                       rng.uniform(0.01, 0.1, nlines)):
        tau += s * w / np.pi / ((nu - c) ** 2 + w * w)
    return tau


def test_oracle_constant_spectrum_closed_form(oracle):
    press, temp, tau0 = 250.0, 230.0, 3e-3
    nuf = np.array([700.0, 700.5, 701.0, 701.5])
    filt = np.array([0.2, 1.0, 0.7, 0.1])
    tau = np.full(4001, tau0)
    u, e = oracle.tblgen(press, temp, tau, 699.0, 703.0, nuf, filt)
    a = tau0 / uref(press, temp)
    ue, ee = expected_rows(lambda x: 1 - np.exp(-a * x))
    assert len(u) == len(ue) > 50
    np.testing.assert_array_equal(u, ue)
    # eps = 1 - S / fsum: the rounding of the ~1000-term sums, ~1e-16 each,
    # is absolute in eps
    np.testing.assert_allclose(e, ee, rtol=1e-12, atol=2e-14)


def test_oracle_two_level_spectrum(oracle):
    # levels a1 (even points) and a2 (odd points) under a linear filter
    press, temp = 100.0, 210.0
    n, nu0, nu1 = 2001, 800.0, 802.0
    nu = nu0 + (nu1 - nu0) / (n - 1) * np.arange(n)
    tau = np.where(np.arange(n) % 2 == 0, 5e-4, 4e-2)
    nuf = np.array([800.3, 801.7])
    filt = np.array([1.0, 3.0])
    u, e = oracle.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
    sel = (nu >= nuf[0]) & (nu <= nuf[1])
    w = np.interp(nu[sel], nuf, filt)
    ev = (np.arange(n)[sel] % 2) == 0
    a1, a2 = 5e-4 / uref(press, temp), 4e-2 / uref(press, temp)
    W1, W2 = w[ev].sum(), w[~ev].sum()
    ue, ee = expected_rows(lambda x: 1 - (W1 * np.exp(-a1 * x) + W2 * np.exp(-a2 * x)) / (W1 + W2))
    np.testing.assert_array_equal(u, ue)
    np.testing.assert_allclose(e, ee, rtol=1e-11, atol=2e-14)


def test_oracle_empty_filter_range(oracle):
    # no spectrum point inside the filter: fsum = 0, eps = NaN, no rows
    u, e = oracle.tblgen(300.0, 250.0, np.ones(100), 600.0, 601.0, np.array([700.0, 701.0]),
                         np.array([1.0, 1.0]))
    assert len(u) == 0


def test_spectrum_file_roundtrip(tmp_path):
    tau = line_spectrum(1003, 650.0, 652.0)
    p = tmp_path / "spec.txt"
    T.write_spectrum(str(p), tau, 650.0, 652.0)
    t2, a, b = T.read_spectrum(str(p))
    assert (a, b) == (650.0, 652.0)
    np.testing.assert_allclose(t2, tau, rtol=1e-8)
    f = tmp_path / "filt.txt"
    f.write_text("# filter\n650.5 0.1\n651.0 1.0\nbad line\n651.5 0.2\n")
    nuf, filt = T.read_filter(str(f))
    np.testing.assert_array_equal(nuf, [650.5, 651.0, 651.5])
    np.testing.assert_array_equal(filt, [0.1, 1.0, 0.2])


def test_tblgen_argument_checks_on_host():
    # rejected or empty before any device call, so these run without a GPU
    from jurassic_amd import gpu as G
    with pytest.raises(G.JurassicError):
        G.tblgen(300.0, 250.0, np.ones(1), 600.0, 601.0, np.array([600.2, 600.8]), np.ones(2))
    with pytest.raises(G.JurassicError):
        G.tblgen(300.0, 250.0, np.ones(100), 600.0, 601.0, np.array([600.2]), np.ones(1))
    assert G.tblgen_batch(np.zeros(0), np.zeros(0), np.zeros((0, 100)), 600.0, 601.0,
                          np.array([600.2, 600.8]), np.ones(2)) == []


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from jurassic_amd import gpu as G
    G.lib()
    return G


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(20001, 1), (250001, 2)])
def test_gpu_tblgen_matches_oracle(gpu, oracle, n, seed):
    nu0, nu1 = 700.0, 705.0
    tau = line_spectrum(n, nu0, nu1, seed=seed)
    nuf = np.linspace(701.0, 704.0, 61)
    filt = np.exp(-0.5 * ((nuf - 702.5) / 0.6) ** 2)
    for press, temp in ((1013.25, 290.0), (10.0, 220.0)):
        u0, e0 = oracle.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
        u1, e1 = gpu.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
        print("tblgen n=%d p=%g: %d rows" % (n, press, len(u0)), flush=True)
        assert len(u1) == len(u0) > 20
        np.testing.assert_array_equal(u1, u0)
        # chunked sums against the oracle's sequential ones: the rounding of
        # sums of n terms, ~sqrt(n) 1e-16 of the filter sum, is absolute in eps
        np.testing.assert_allclose(e1, e0, rtol=1e-10, atol=1e-15 * np.sqrt(n))


@pytest.mark.gpu
def test_gpu_tblgen_constant_closed_form(gpu):
    press, temp, tau0 = 500.0, 260.0, 1e-2
    nuf = np.array([900.0, 901.0])
    filt = np.array([1.0, 2.0])
    u, e = gpu.tblgen(press, temp, np.full(5001, tau0), 899.0, 902.0, nuf, filt)
    a = tau0 / uref(press, temp)
    ue, ee = expected_rows(lambda x: 1 - np.exp(-a * x))
    np.testing.assert_array_equal(u, ue)
    np.testing.assert_allclose(e, ee, rtol=1e-12, atol=2e-14)


@pytest.mark.gpu
def test_gpu_tblgen_points_on_filter_nodes(gpu, oracle):
    # grid step 0.01 and filter nodes on grid points: the carried filter
    # bracket decides the bracket of a point on a node
    n, nu0, nu1 = 501, 600.0, 605.0
    tau = line_spectrum(n, nu0, nu1, seed=5, nlines=10)
    nuf = nu0 + (nu1 - nu0) / (n - 1) * np.array([50, 120, 230, 300, 410])
    filt = np.array([0.3, 1.0, 0.8, 0.9, 0.05])
    u0, e0 = oracle.tblgen(700.0, 270.0, tau, nu0, nu1, nuf, filt)
    u1, e1 = gpu.tblgen(700.0, 270.0, tau, nu0, nu1, nuf, filt)
    np.testing.assert_array_equal(u1, u0)
    np.testing.assert_allclose(e1, e0, rtol=1e-12, atol=1e-14)


@pytest.mark.gpu
def test_gpu_tblgen_cli(gpu, oracle, tmp_path):
    tau = line_spectrum(30001, 660.0, 670.0, seed=3)
    spec, filt_p = tmp_path / "spec.txt", tmp_path / "filt.txt"
    T.write_spectrum(str(spec), tau, 660.0, 670.0)
    filt_p.write_text("".join("%.4f %.6f\n" % (x, np.exp(-((x - 665.0) / 2) ** 2))
                              for x in np.linspace(662.0, 668.0, 121)))
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        assert T.main(["tblgen", "300", "250", str(spec), str(filt_p)]) == 0
    t2, a, b = T.read_spectrum(str(spec))
    nuf, filt = T.read_filter(str(filt_p))
    u0, e0 = oracle.tblgen(300.0, 250.0, t2, a, b, nuf, filt)
    lines = out.getvalue().splitlines()
    assert len(lines) == len(u0) > 20
    got = np.array([[float(x) for x in ln.split()] for ln in lines])
    np.testing.assert_allclose(got[:, 2], u0, rtol=1e-5)
    np.testing.assert_allclose(got[:, 3], e0, rtol=1e-5)
    assert lines[0].split()[:2] == ["300", "250"]


@pytest.mark.gpu
def test_gpu_tblgen_batch_matches_oracle(gpu, oracle):
    # one table's (p, T) grid on one spectral grid, each spectrum its own
    n, nu0, nu1 = 8001, 700.0, 704.0
    press = np.repeat([1000.0, 300.0, 50.0, 2.0], 3)
    temp = np.tile([200.0, 250.0, 300.0], 4)
    tau = np.stack([line_spectrum(n, nu0, nu1, seed=10 + s) * (press[s] / 300.0) for s in range(len(press))])
    nuf = np.linspace(700.5, 703.5, 31)
    filt = 1.0 - np.abs(nuf - 702.0) / 1.6
    rows = gpu.tblgen_batch(press, temp, tau, nu0, nu1, nuf, filt)
    assert len(rows) == len(press)
    for s, (u1, e1) in enumerate(rows):
        u0, e0 = oracle.tblgen(press[s], temp[s], tau[s], nu0, nu1, nuf, filt)
        assert len(u1) == len(u0) > 10, s
        np.testing.assert_array_equal(u1, u0)
        np.testing.assert_allclose(e1, e0, rtol=1e-10, atol=1e-15 * np.sqrt(n))


@pytest.mark.gpu
def test_gpu_tblgen_row_capacity(gpu):
    # fewer output slots than rows: the first maxrow rows, and the count of all
    import ctypes as C
    tau = line_spectrum(10001, 700.0, 702.0, seed=4)
    nuf = np.array([700.5, 701.0, 701.5])
    filt = np.array([0.5, 1.0, 0.5])
    u_all, e_all = gpu.tblgen(400.0, 260.0, tau, 700.0, 702.0, nuf, filt)
    dp = C.POINTER(C.c_double)
    u = np.zeros(5)
    e = np.zeros(5)
    n = C.c_int(0)
    rc = gpu.lib().jur_tblgen(400.0, 260.0, tau.ctypes.data_as(dp), len(tau), 700.0, 702.0, nuf.ctypes.data_as(dp),
                              filt.ctypes.data_as(dp), 3, u.ctypes.data_as(dp), e.ctypes.data_as(dp), 5,
                              C.byref(n))
    assert rc == 0 and n.value == len(u_all) > 5
    np.testing.assert_array_equal(u, u_all[:5])
    np.testing.assert_array_equal(e, e_all[:5])


@pytest.mark.gpu
def test_gpu_tblgen_bad_args(gpu):
    with pytest.raises(gpu.JurassicError):
        gpu.tblgen(300.0, 250.0, np.ones(1), 600.0, 601.0, np.array([600.2, 600.8]), np.ones(2))
