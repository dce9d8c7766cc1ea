"""Obit's own UVSub test case as a golden vector for the DFT divide.

The fixture (tests/golden/uvsub_fixture.npz, written by
tests/golden/make_uvsub_fixture.py) holds 2000 seeded rows of
testData/UVSubTestIn.uvtab and of testData/UVSubTestMaster.uvtab (the input
divided by the DFT of CC table 2 of testData/UVSubTestModIn.fits.gz: the
reference's testScripts/testUVSub.py, Mode DFT, Type CC), and the 13
components placed by textbook geometry (no fitted parameter).

Pinned here: the divide operation as the master holds it (data / M, weight
x |M|, flagged and cross-hand correlations untouched) and, end to end, the
oracle's and the HIP kernel's DFT divide against the master under the
reference's own comparison metric (ObitUVUtilVisCompare,
src/ObitUVUtil.c:1405-1480).  The master's model was formed with FP32
phases (|phase| up to ~4e4 rad here) from Obit's own float component
table, so agreement is at the 1e-3 level, not bit level."""
import os

import numpy as np
import pytest

from obit_amd import synth

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uvsub_fixture.npz")
METRIC_TOL = 3e-3   # VisCompare fractional RMS vs the master
WEIGHT_TOL = 3e-3   # median relative weight difference vs the master


@pytest.fixture(scope="module")
def fx():
    d = np.load(FIX, allow_pickle=False)
    return {k: d[k] for k in d.files}


def _uv(fx, rows):
    return synth.UVData(buffer=np.ascontiguousarray(rows, np.float32), nchan=int(fx["nchan"]),
                        nif=int(fx["nif"]), nstok=int(fx["nstok"]), freq=float(fx["f0"]),
                        freqArr=np.asarray(fx["freqs"], np.float64), crval_stokes=-1.0)


def _sky(fx):
    return synth.SkyModel(comps=np.ascontiguousarray(fx["comps_geom"], np.float32), modType=0)


def vis_compare(a, b, nrparm=5):
    """ObitUVUtilVisCompare: sqrt(sum((dRe^2 + dIm^2) / |V_b|) / count) over
    correlations with positive weight in both and |V_b| > 0."""
    x = a[:, nrparm:].reshape(a.shape[0], -1, 3).astype(np.float64)
    y = b[:, nrparm:].reshape(b.shape[0], -1, 3).astype(np.float64)
    amp = np.sqrt(y[..., 0] ** 2 + y[..., 1] ** 2)
    ok = (x[..., 2] > 0) & (y[..., 2] > 0) & (amp > 0)
    s = ((x[..., 0] - y[..., 0]) ** 2 + (x[..., 1] - y[..., 1]) ** 2)[ok] / amp[ok]
    return float(np.sqrt(s.sum() / (2 * ok.sum()))), int(2 * ok.sum())


def _check_against_master(fx, out):
    rin, rma = fx["rows_in"], fx["rows_master"]
    metric, n = vis_compare(out, rma)
    assert n > 7000
    assert metric < METRIC_TOL, metric
    nif, nst = int(fx["nif"]), int(fx["nstok"])
    o = out[:, 5:].reshape(-1, nif, nst, 3)
    m = rma[:, 5:].reshape(-1, nif, nst, 3)
    i = rin[:, 5:].reshape(-1, nif, nst, 3)
    good = m[..., :2, 2] > 0
    rel = np.abs(o[..., :2, 2][good] / m[..., :2, 2][good] - 1.0)
    assert np.median(rel) < WEIGHT_TOL, np.median(rel)
    # flagged samples and the (flagged) cross hands come through unchanged
    assert np.array_equal(o[..., 2:, :], i[..., 2:, :])
    bad = ~good
    assert np.array_equal(o[..., :2, :][bad], i[..., :2, :][bad])
    return metric


def test_master_is_a_divide(fx):
    """What the reference's master holds: V_out = V_in / M with the weight
    scaled by |M| (ObitSkyModel.c:2794-2844), flagged data untouched."""
    rin, rma = fx["rows_in"], fx["rows_master"]
    assert np.array_equal(rin[:, :5], rma[:, :5])
    nif, nst = int(fx["nif"]), int(fx["nstok"])
    i = rin[:, 5:].reshape(-1, nif, nst, 3).astype(np.float64)
    m = rma[:, 5:].reshape(-1, nif, nst, 3).astype(np.float64)
    good = (i[..., :2, 2] > 0)
    assert np.array_equal(good, m[..., :2, 2] > 0)
    M = (i[..., :2, 0] + 1j * i[..., :2, 1])[good] / (m[..., :2, 0] + 1j * m[..., :2, 1])[good]
    ratio = m[..., :2, 2][good] / i[..., :2, 2][good]
    assert np.max(np.abs(np.abs(M) / ratio - 1.0)) < 1e-6
    assert np.array_equal(i[..., 2:, :], m[..., 2:, :])


def test_geometry_model_matches_master(fx):
    """The FP64 DFT of the geometry-placed components against the model the
    master implies (M = V_in / V_out): a position, sign or frequency-scaling
    slip would decorrelate them."""
    rin, rma = fx["rows_in"], fx["rows_master"]
    nif, nst = int(fx["nif"]), int(fx["nstok"])
    i = rin[:, 5:].reshape(-1, nif, nst, 3).astype(np.float64)
    m = rma[:, 5:].reshape(-1, nif, nst, 3).astype(np.float64)
    good = i[..., :2, 2] > 0
    Mobs = np.where(good, (i[..., :2, 0] + 1j * i[..., :2, 1]) /
                    np.where(good, m[..., :2, 0] + 1j * m[..., :2, 1], 1), 0)
    t = fx["comps_geom"].astype(np.float64)
    uvw = rin[:, :3].astype(np.float64)
    fsc = np.asarray(fx["freqs"]) / float(fx["f0"])
    ph = uvw @ t[:, 1:4].T
    mod = np.stack([(t[:, 0] * np.exp(1j * ph * f)).sum(axis=1) for f in fsc], axis=1)[..., None]
    r = (mod - Mobs) * good
    rms = np.sqrt((np.abs(r) ** 2).sum() / good.sum())
    assert rms < 2e-3 * np.abs(Mobs).max(), rms


def test_oracle_divide_reproduces_master(fx, oracle):
    uv = _uv(fx, fx["rows_in"])
    out = oracle.dft_obit(uv, _sky(fx), op="div")
    metric = _check_against_master(fx, out)
    # and a mistake the metric must catch: the model's phase sign flipped
    bad = _sky(fx)
    bad.comps = bad.comps.copy()
    bad.comps[:, 1:4] *= -1
    worse, _ = vis_compare(oracle.dft_obit(uv, bad, op="div"), fx["rows_master"])
    assert worse > 100 * metric


@pytest.mark.gpu
def test_hip_divide_reproduces_master(fx, oracle):
    from obit_amd.skymodel import SkyModelDFT
    uv = _uv(fx, fx["rows_in"])
    with SkyModelDFT(uv, _sky(fx), op="div") as sm:
        sm.set_vis(uv.buffer)
        sm.calc_resident()
        out = sm.get_vis()
    _check_against_master(fx, out)
    # the per-buffer path (Obit's DFTCalc) gives the same result
    with SkyModelDFT(uv, _sky(fx), op="div", nVisPIO=700) as sm:
        buf = sm.calc(uv.buffer.copy())
    assert np.array_equal(buf, out)
    # and the HIP result sits within the CPU restatement's own tolerance of it
    ref = oracle.dft_obit(uv, _sky(fx), op="div")
    a = out[:, 5:].reshape(-1, 3)
    b = ref[:, 5:].reshape(-1, 3)
    ok = b[:, 2] > 0
    d = np.abs((a[ok, 0] - b[ok, 0]) + 1j * (a[ok, 1] - b[ok, 1]))
    assert np.sqrt(np.mean(d ** 2)) < 2e-3


def test_load_comps_builds_the_component_table(fx, oracle):
    """ObitHostLoadCompsPoint (ObitSkyModelLoadComps for a 3-D field, in
    the reference's float arithmetic) against the FP64 textbook geometry:
    agreement to float rounding; with the image on the uv phase centre and
    no rotation it reduces to konst x (deltaX, deltaY, 0); and the table it
    builds for the test's two fields x factor 0.5 reproduces the master
    through the oracle's divide as the textbook table does."""
    from obit_amd.skymodel import load_comps_point
    one = load_comps_point(fx["cc"], fx["img_crval"], fx["img_crota"], fx["uv_centre"],
                           fx["uv_crota"], 1.0)
    geo = fx["comps_geom"]
    assert np.array_equal(one[:, 0], geo[:, 0])
    rel = np.abs(one[:, 1:] - geo[:, 1:]) / np.abs(geo[:, 1:]).max(axis=0)
    assert rel.max() < 1e-5, rel.max()
    # a field centred on the phase centre, no rotation: no 3-D matrix
    same = load_comps_point(fx["cc"], fx["uv_centre"], 0.0, fx["uv_centre"], 0.0, 1.0)
    konst = np.float32(np.pi / 180.0 * 2.0 * np.pi)
    assert np.array_equal(same[:, 1], fx["cc"][:, 1] * konst)
    assert np.array_equal(same[:, 2], fx["cc"][:, 2] * konst)
    assert np.all(same[:, 3] == 0)
    # the UVSub run: two fields of the same image, Factor 0.5
    comps = np.concatenate([load_comps_point(fx["cc"], fx["img_crval"], fx["img_crota"],
                                             fx["uv_centre"], fx["uv_crota"], float(fx["factor"]))]
                           * int(fx["nfield"]))
    out = oracle.dft_obit(_uv(fx, fx["rows_in"]), synth.SkyModel(comps=comps, modType=0), op="div")
    metric = _check_against_master(fx, out)
    print(f"builder table: VisCompare against the master {metric:.2e}")
