"""The DFT oracles, pinned by closed forms and invariants that a dropped
term, a wrong sign or a transposed operand would break, and the
restatement of Obit's FP32 CPU DFT checked against the FP64 oracle within
the CPU's own phase-rounding bound."""
import math

import numpy as np
import pytest

from obit_amd import synth

C_LIGHT = 299792458.0


def _tiny_uv(nrow=5, nchan=4, nif=2, nstok=2, seed=0):
    rng = np.random.default_rng(seed)
    hdr = np.zeros((nrow, 5))
    hdr[:, :3] = rng.uniform(-5e4, 5e4, (nrow, 3))
    freq = 1.4e9
    fa = synth.freq_table(freq, nif, nchan, 2e6)
    return synth.make_uv(hdr, nif, nchan, nstok, freq, fa, rng)


def _sky(rows, modType=0, **kw):
    return synth.SkyModel(comps=np.asarray(rows, np.float32), modType=modType, **kw)


def test_origin_component_is_constant(oracle):
    uv = _tiny_uv()
    m = oracle.dft_model_fp64(uv, _sky([[2.5, 0, 0, 0]]))
    assert np.allclose(m, 2.5, rtol=0, atol=1e-15)


def test_single_point_closed_form(oracle):
    """V = a exp(2 pi i (u l + v m + w (n-1))) nu/nu0, with (l, m) chosen so
    the phase is large (|phi| ~ 1e3 rad)."""
    uv = _tiny_uv()
    l, m = 0.004, -0.003
    row = synth.point_row(np.array([1.3]), np.array([l]), np.array([m]))
    tab = row.astype(np.float32)
    v = oracle.dft_model_fp64(uv, _sky(tab))
    x, y, z = (float(t) for t in tab[0, 1:4])
    uvw = uv.buffer[:, :3].astype(np.float64)
    fr = uv.freqArr / uv.freq
    ph = (uvw @ np.array([x, y, z]))[:, None] * fr[None, :]
    ref = (float(tab[0, 0]) * np.exp(1j * ph)).reshape(uv.nrow, uv.nif, uv.nchan)
    assert np.abs(v - ref).max() < 1e-12 * 1.3
    assert np.abs(ph).max() > 100


def test_conjugate_symmetry_and_linearity(oracle):
    uv = _tiny_uv(nrow=6)
    rng = np.random.default_rng(2)
    a = synth.point_row(rng.uniform(0.1, 1, 7), rng.uniform(-1e-3, 1e-3, 7), rng.uniform(-1e-3, 1e-3, 7))
    b = synth.point_row(rng.uniform(0.1, 1, 5), rng.uniform(-1e-3, 1e-3, 5), rng.uniform(-1e-3, 1e-3, 5))
    va = oracle.dft_model_fp64(uv, _sky(a))
    vb = oracle.dft_model_fp64(uv, _sky(b))
    vab = oracle.dft_model_fp64(uv, _sky(np.concatenate([a, b])))
    assert np.abs(vab - va - vb).max() < 1e-12
    uvn = synth.UVData(**{**uv.__dict__, "buffer": uv.buffer.copy()})
    uvn.buffer[:, :3] *= -1
    assert np.abs(oracle.dft_model_fp64(uvn, _sky(a)) - np.conj(va)).max() < 1e-12


def test_gaussian_textbook_form(oracle):
    """A circular Gaussian of FWHM theta at the origin has visibility
    exp(-(pi theta)^2 (u^2+v^2) / (4 ln 2)); Obit's table builds it from
    konst2 (ObitSkyModel.c:1722, 1939-1947)."""
    uv = _tiny_uv()
    fwhm_deg = 20.0 / 3600.0
    gp = synth.gauss_coeffs(np.array([fwhm_deg]), np.array([fwhm_deg]), np.array([30.0]))
    tab = [[1.0, 0, 0, 0, gp[0][0], gp[1][0], gp[2][0]]]
    v = oracle.dft_model_fp64(uv, _sky(tab, modType=1))
    th = fwhm_deg * math.pi / 180
    fr = uv.freqArr / uv.freq
    uvw = uv.buffer[:, :3].astype(np.float64)
    q = (uvw[:, 0] ** 2 + uvw[:, 1] ** 2)[:, None] * fr[None, :] ** 2
    ref = np.exp(-(math.pi * th) ** 2 * q / (4 * math.log(2))).reshape(v.shape)
    assert np.abs(v - ref).max() < 2e-6  # FP32 table coefficients
    # elliptical with PA: major axis along PA rotates the response
    gp2 = synth.gauss_coeffs(np.array([2 * fwhm_deg]), np.array([fwhm_deg]), np.array([90.0]))
    v2 = oracle.dft_model_fp64(uv, _sky([[1.0, 0, 0, 0, gp2[0][0], gp2[1][0], gp2[2][0]]], modType=1))
    th1, th2 = 2 * th, th
    # PA 90 deg: major axis along x (east) -> stronger decay along u
    qq = ((th1 * uvw[:, 0]) ** 2 + (th2 * uvw[:, 1]) ** 2)[:, None] * fr[None, :] ** 2
    ref2 = np.exp(-(math.pi ** 2) * qq / (4 * math.log(2))).reshape(v.shape)
    assert np.abs(v2 - ref2).max() < 2e-6


def test_gaussian_zero_width_is_point(oracle):
    uv = _tiny_uv()
    rng = np.random.default_rng(4)
    pts = synth.point_row(rng.uniform(0.1, 1, 5), rng.uniform(-1e-3, 1e-3, 5), rng.uniform(-1e-3, 1e-3, 5))
    g = np.concatenate([pts, np.zeros((5, 3))], axis=1)
    assert np.abs(oracle.dft_model_fp64(uv, _sky(g, modType=1)) - oracle.dft_model_fp64(uv, _sky(pts))).max() < 1e-13


def test_spectral_power_law(oracle):
    """PointModSpec with one term: S(nu) = S0 (nu/nu0)^alpha (CPU sign)."""
    uv = _tiny_uv()
    alpha = -0.8
    nu0 = 1.35e9
    v = oracle.dft_model_fp64(uv, _sky([[2.0, 0, 0, 0, alpha]], modType=10, nSpecTerm=1, imgRefFreq=nu0))
    ref = 2.0 * (uv.freqArr / nu0) ** alpha
    assert np.abs(v - ref.reshape(1, uv.nif, uv.nchan)).max() < 1e-6 * 2  # FP32 alpha


def test_tabulated_spectrum(oracle):
    """PointModTSpec: S_bin (nu/nu_bin)^alpha with the bin from specIndex."""
    uv = _tiny_uv()
    nspec = 2
    sf = np.array([uv.freqArr[1], uv.freqArr[uv.nchan + 2]])
    si = np.repeat(np.arange(nspec), uv.nchan).astype(np.int32)
    tab = [[1.0, 0, 0, 0, 0.5, 3.0, 4.0]]
    v = oracle.dft_model_fp64(uv, _sky(tab, modType=20, specFreq=sf, specIndex=si))
    S = np.array([3.0, 4.0])[si]
    ref = S * (uv.freqArr / sf[si]) ** 0.5
    assert np.abs(v - ref.reshape(1, uv.nif, uv.nchan)).max() < 1e-12


def test_flip_multiplies_by_i(oracle):
    uv = _tiny_uv()
    tab = synth.point_row(np.array([1.0]), np.array([1e-3]), np.array([2e-3]))
    a = oracle.dft_model_fp64(uv, _sky(tab))
    b = oracle.dft_model_fp64(uv, _sky(tab), doFlip=True)
    assert np.abs(b - 1j * a).max() < 1e-15


@pytest.mark.parametrize("cfg", [1, 3])
def test_cpu_restatement_within_its_phase_bound(oracle, cfg):
    """Obit's CPU DFT rounds each phase to FP32 and uses Cephes sincos:
    |dV| <= sum_k |a_k| (|phi_k| 2^-24 + 2e-7) (plus Gaussian exp terms)."""
    uv, sky, _ = synth.workload(cfg, rows=40)
    if cfg == 3:
        sky = synth.SkyModel(comps=sky.comps[:3000], modType=1)
    cpu = oracle.dft_obit(uv, sky, op="repl")
    d = cpu[:, 5:].reshape(uv.nrow, uv.nif, uv.nchan, uv.nstok, 3)
    got = d[..., 0, 0] + 1j * d[..., 0, 1]
    ref = oracle.dft_model_fp64(uv, sky)
    uvw = np.abs(uv.buffer[:, :3]).astype(np.float64) * (uv.freqArr.max() / uv.freq)
    phimax = uvw @ np.abs(sky.comps[:, 1:4]).astype(np.float64).max(axis=0)
    bound = np.abs(sky.comps[:, 0]).sum() * (phimax[:, None, None] * 2 ** -24 * 2 + 4e-7)
    assert (np.abs(got - ref) <= bound).all()
    # the bound is tight enough to matter: the restatement is not FP64
    assert np.abs(got - ref).max() > 1e-7 * np.abs(ref).max()


def test_cpu_ops(oracle):
    """Subtract / replace / divide of the restated CPU DFT against the FP64
    model, on unflagged data (flags are covered on the GPU side)."""
    uv, sky, _ = synth.workload(1, rows=20, ncomp=50)
    m = oracle.dft_model_fp64(uv, sky)[..., None]
    dv = uv.data_view()
    data = dv[..., 0] + 1j * dv[..., 1]
    tol = 3e-4 * np.abs(m).max()
    sub = oracle.dft_obit(uv, sky, op="sub")[:, 5:].reshape(dv.shape)
    assert np.abs((sub[..., 0] + 1j * sub[..., 1]) - (data - m)).max() < tol
    rep = oracle.dft_obit(uv, sky, op="repl")[:, 5:].reshape(dv.shape)
    assert np.abs((rep[..., 0] + 1j * rep[..., 1]) - m).max() < tol
    div = oracle.dft_obit(uv, sky, op="div")[:, 5:].reshape(dv.shape)
    q = div[..., 0] + 1j * div[..., 1]
    assert np.median(np.abs(q * m - data)) < 1e-3
    assert np.abs(div[..., 2] - dv[..., 2] * np.abs(m)).max() < tol  # weight x |model|


def test_threaded_decomposition_is_exact(oracle):
    """ObitSkyModelFTDFT's row split changes nothing (rows are independent)."""
    uv, sky, _ = synth.workload(1, rows=1200, ncomp=40)
    a = oracle.dft_obit(uv, sky, nthreads=1)
    b = oracle.dft_obit(uv, sky, nthreads=7)
    assert np.array_equal(a, b)
