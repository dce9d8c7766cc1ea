"""Jones-corrected DFT on the GPU (ObitHIPBeamDFT*) against the FP64
oracle (oracle.beam_model_fp64 + beam_apply).

Cases cover 2 and 4 correlations, point and Gaussian tables, the four
Stokes models, subtract / replace / divide, beam-channel maps that keep an
8-channel chunk in one beam channel (LDS path) and maps that do not
(per-channel products), more antenna types than the LDS path stages, rows
whose antennas fall outside the type table, and flagged samples.
Tolerance: the model is within 1e-5 of its peak amplitude (the split DFT's
bound); entries the application rules leave alone are bit-identical."""
import numpy as np
import pytest

from obit_amd import synth
from obit_amd.beam import BeamDFT
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

NANT = 27


def _data(ncorr, rows, nif=2, nchan=40, seed=0):
    """Workload 1's uvw and baselines, ncorr correlations per channel of
    N(0,1) data with unit weights, 8 % of samples flagged, 2 % of rows with
    an antenna outside the 27-entry type table."""
    rng = np.random.default_rng(seed)
    base, _, _ = synth.workload(1, rows=rows)
    nrparm = base.nrparm
    lrec = nrparm + nif * nchan * 3 * ncorr
    buf = np.zeros((rows, lrec), np.float32)
    buf[:, :nrparm] = base.buffer[:, :nrparm]
    vis = buf[:, nrparm:].reshape(rows, nif, nchan, ncorr, 3)
    vis[..., 0] = rng.normal(size=vis.shape[:-1])
    vis[..., 1] = rng.normal(size=vis.shape[:-1])
    vis[..., 2] = np.where(rng.uniform(size=vis.shape[:-1]) < 0.08, -1.0, 1.0)
    vis[rng.uniform(size=(rows, nif, nchan)) < 0.02, :, 2] = 0.0     # whole channels flagged
    bad = rng.uniform(size=rows) < 0.02
    cb = buf[:, base.ilocb].astype(np.int64)
    buf[bad, base.ilocb] = ((cb[bad] // 256) * 256 + 31).astype(np.float32)   # antenna 31
    f0 = base.freq
    freqArr = (f0 * (1.0 + 0.02 * np.arange(nif))[:, None] + 250e3 * np.arange(nchan)[None, :]).ravel()
    uv = synth.UVData(buffer=buf, nchan=nchan, nif=nif, nstok=ncorr, freq=f0, freqArr=freqArr,
                      crval_stokes=-5.0, nrparm=nrparm)
    return uv


def _comps(gauss, ncomp=48, seed=1):
    rng = np.random.default_rng(seed)
    _, sky, _ = synth.workload(1, rows=8)
    c = sky.comps[:ncomp].astype(np.float64).copy()
    c[0, 0] = 1.0                                  # a dominant component (division)
    if gauss:
        g = np.zeros((ncomp, 3))
        isg = rng.uniform(size=ncomp) < 0.4
        maj = rng.uniform(2.0, 30.0, ncomp) / 3600.0
        gp = np.stack(synth.gauss_coeffs(maj, maj * rng.uniform(0.3, 1.0, ncomp),
                                         rng.uniform(0, 180, ncomp)), axis=-1)
        g[isg] = gp[isg]
        c = np.concatenate([c, g], axis=1)
    return c.astype(np.float32)


def _jones(ntype, ncomp, njchan, seed=2):
    rng = np.random.default_rng(seed)
    j = np.zeros((ntype, ncomp, njchan, 2, 2), np.complex128)
    j[..., 0, 0] = 1.0
    j[..., 1, 1] = 1.0
    j += 0.15 * (rng.normal(size=j.shape) + 1j * rng.normal(size=j.shape))
    return j.astype(np.complex64).astype(np.complex128)      # the device holds FP32


CASES = [
    # ncorr, gauss, stokes, op, uniform beam channels, antenna types
    (2, False, 1, "sub", True, 2),
    (2, True, 2, "repl", False, 2),
    (4, False, 3, "sub", True, 3),
    (4, True, 4, "repl", False, 2),
    (2, False, 1, "div", True, 1),
    (4, True, 2, "div", False, 2),
    (4, False, 4, "sub", True, 5),      # 25 type pairs: per-channel products
    (2, True, 3, "sub", False, 1),
]


@pytest.mark.parametrize("ncorr,gauss,stok,op,uniform,ntype", CASES)
def test_beam_dft_matches_oracle(ncorr, gauss, stok, op, uniform, ntype):
    uv = _data(ncorr, rows=1500, seed=ncorr + 10 * stok)
    comps = _comps(gauss)
    rng = np.random.default_rng(3)
    ant_type = rng.integers(0, ntype, NANT)
    nf = uv.nif * uv.nchan
    ch = np.arange(nf) % uv.nchan
    jmap = (np.arange(nf) // uv.nchan) * 5 + (ch // 8 if uniform else ch // 6)
    njchan = int(jmap.max()) + 1
    jones = _jones(ntype, len(comps), njchan)
    chi = 0.41
    sf = (1.0, 0.9, 0.8, 1.1)
    with BeamDFT(uv, comps, ant_type, jmap, njchan, ntype=ntype, stok_type=stok, op=op,
                 stok_fact=sf) as bm:
        bm.set_jones(jones)
        got = bm.calc(uv.buffer, par_ang=chi)
    model, live = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, ant_type, jmap, jones, stok,
                                      chi, ilocb=uv.ilocb, stok_fact=sf)
    want = orc.beam_apply(uv.buffer, model, live, uv.nrparm, uv.nif, uv.nchan, uv.incf, uv.incif, ncorr, op)
    assert (~live).sum() > 0 and live.sum() > 0
    # the entries the rules select: a NaN model marks them
    moved = np.isnan(orc.beam_apply(uv.buffer, np.full_like(model, complex(np.nan, np.nan)), live, uv.nrparm, uv.nif,
                                    uv.nchan, uv.incf, uv.incif, ncorr, op))
    assert np.array_equal(got[~moved], uv.buffer[~moved]), "entries the model must not touch changed"
    peak = np.abs(model[live]).max()
    err = np.abs(got[moved].astype(np.float64) - want[moved]).max()
    if op == "div":
        scale = np.abs(want[moved]).max()
        assert err < 2e-5 * scale, (err, scale)
    else:
        assert err < 1e-5 * peak, (err, peak)


def test_beam_unit_jones_matches_split_dft():
    """Unit Jones, Stokes I, 2 correlations: the subtracted model is the
    plain DFT's (the oracle's FP64 model of the same table)."""
    uv = _data(2, rows=1200, nif=1, nchan=64, seed=5)
    comps = _comps(False)
    with BeamDFT(uv, comps, np.zeros(NANT, int), np.zeros(uv.nchan, int), 1, stok_type=1, op="sub") as bm:
        got = bm.calc(uv.buffer)
    sky = synth.SkyModel(comps=comps, modType=0)
    plain = orc.dft_model_fp64(uv, sky)                          # [nrow, nif, nchan]
    cb = uv.buffer[:, uv.ilocb].astype(np.int64)
    live = (cb // 256 <= NANT) & (cb % 256 <= NANT)
    d0 = uv.data_view()
    d1 = got[:, uv.nrparm:].reshape(d0.shape)
    for h in range(2):
        wt = d0[..., h, 2] > 0
        sel = wt & live[:, None, None]
        mod = (d0[..., h, 0] - d1[..., h, 0]) + 1j * (d0[..., h, 1] - d1[..., h, 1])
        assert np.abs(mod[sel] - plain[sel]).max() < 1e-5 * np.abs(plain).max()
        assert np.all(mod[~sel] == 0)


def test_beam_refuses_bad_tables():
    uv = _data(2, rows=64, nif=1, nchan=8)
    comps = _comps(False, ncomp=4)
    with pytest.raises(Exception, match="type"):
        BeamDFT(uv, comps, np.full(NANT, 3), np.zeros(8, int), 1, ntype=2)
    with pytest.raises(Exception, match="beam channel"):
        BeamDFT(uv, comps, np.zeros(NANT, int), np.full(8, 4), 2)


FBLANK = -1.0e30


def _beams(ntype, nch, proj, nx=48, ny=40):
    """Smooth beam images with a blanked patch: P1, P2 real parts a
    Gaussian bump (<= 0.95), the other six small ripples."""
    y, x = np.mgrid[0:ny, 0:nx] / 10.0
    images = []
    for t in range(ntype):
        imgs = np.zeros((8, nch, ny, nx), np.float32)
        for c in range(nch):
            bump = np.exp(-((x - nx / 20.0) ** 2 + (y - ny / 20.0) ** 2) / (6.0 + c + 0.5 * t))
            imgs[0, c] = 0.95 * bump
            imgs[2, c] = 0.9 * bump
            for i in (1, 3, 4, 5, 6, 7):
                imgs[i, c] = 0.05 * np.sin(0.3 * (i + 1) * x + 0.2 * t * y + c)
        imgs[:, :, 3:6, 7:10] = FBLANK
        images.append(imgs)
    descs = [dict(nx=nx, ny=ny, proj=proj, hwidth=2, crval=(0.0, 0.0), crpix=(nx / 2 + 0.5, ny / 2 + 0.5),
                  cdelt=(-0.1, 0.1)) for _ in range(ntype)]
    return descs, images


def _positions(n, seed=9):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.8, 1.8, (n, 2)).astype(np.float32)
    pos[0] = (9.0, 9.0)          # off the images: unit matrices
    pos[1] = (2.2, -1.4)         # on the blanked patch
    return pos


@pytest.mark.parametrize("proj", [0, 1])
def test_jones_from_beams_matches_oracle(proj):
    uv = _data(2, rows=64, nif=1, nchan=24)
    comps = _comps(True)
    nch = 3
    descs, images = _beams(2, nch, proj)
    bfreq = np.array([1.3e9, 1.4e9, 1.5e9])
    fact = 2.0e-10
    pos = _positions(len(comps))
    ant_type = np.arange(NANT) % 2
    jmap = np.arange(uv.nchan) // 8
    with BeamDFT(uv, comps, ant_type, jmap, nch, ntype=2) as bm:
        bm.set_beams(descs, images, bfreq, fact, FBLANK)
        bm.jones_from_beams(pos, par_ang=0.3)
        got = bm.get_jones().astype(np.complex128)
    want = orc.beam_jones(descs, images, bfreq, fact, FBLANK, pos, 0.3)
    assert np.array_equal(got[:, 0], want[:, 0])                        # unit matrices
    assert np.abs(got - want).max() < 2e-5, np.abs(got - want).max()
    assert np.abs(want - np.eye(2)).max() > 0.1                          # the beams do act


def test_beam_chain_from_images_matches_oracle():
    """SetBeams + JonesFromBeams + Calc against the oracle's Jones table,
    model and application."""
    uv = _data(4, rows=800, nif=1, nchan=24, seed=21)
    comps = _comps(False)
    nch = 3
    descs, images = _beams(2, nch, 1)
    bfreq = np.array([1.3e9, 1.4e9, 1.5e9])
    pos = _positions(len(comps))
    ant_type = np.arange(NANT) % 2
    jmap = np.arange(uv.nchan) // 8
    chi = -0.6
    with BeamDFT(uv, comps, ant_type, jmap, nch, ntype=2, stok_type=2, op="sub") as bm:
        bm.set_beams(descs, images, bfreq, 2.0e-10, FBLANK)
        bm.jones_from_beams(pos, par_ang=chi)
        got = bm.calc(uv.buffer, par_ang=chi)
    J = orc.beam_jones(descs, images, bfreq, 2.0e-10, FBLANK, pos, chi)
    model, live = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, ant_type, jmap, J, 2, chi,
                                      ilocb=uv.ilocb)
    want = orc.beam_apply(uv.buffer, model, live, uv.nrparm, uv.nif, uv.nchan, uv.incf, uv.incif, 4, "sub")
    moved = np.isnan(orc.beam_apply(uv.buffer, np.full_like(model, complex(np.nan, np.nan)), live, uv.nrparm,
                                    uv.nif, uv.nchan, uv.incf, uv.incif, 4, "sub"))
    assert np.array_equal(got[~moved], uv.buffer[~moved])
    peak = np.abs(model[live]).max()
    assert np.abs(got[moved].astype(np.float64) - want[moved]).max() < 1e-5 * peak


@pytest.mark.parametrize("ncorr,gauss,uniform,op", [(2, False, True, "sub"), (4, True, False, "repl"),
                                                    (2, True, True, "div")])
def test_beam_spectral_table_matches_oracle(ncorr, gauss, uniform, op):
    """Tabulated spectra (flux and spectral index per subband, the
    reference's TSpec form) through ObitHIPBeamDFTSetAmps."""
    from obit_amd.beam import tspec_amps
    uv = _data(ncorr, rows=1000, seed=31 + ncorr)
    comps = _comps(gauss)
    rng = np.random.default_rng(4)
    nf = uv.nif * uv.nchan
    nspec = 4
    spec_index = (np.arange(nf) * nspec) // nf
    fs = uv.freqArr / uv.freq
    freq_rat = 1.0 / np.array([fs[spec_index == s].mean() for s in range(nspec)])
    flux = np.abs(comps[:, :1]) * rng.uniform(0.5, 1.5, (len(comps), nspec))
    flux[0] = 1.0
    si = rng.normal(-0.7, 0.3, (len(comps), nspec))
    amps = tspec_amps(flux, si, spec_index, fs, freq_rat)
    ant_type = rng.integers(0, 2, NANT)
    ch = np.arange(nf) % uv.nchan
    jmap = (np.arange(nf) // uv.nchan) * 5 + (ch // 8 if uniform else ch // 6)
    njchan = int(jmap.max()) + 1
    jones = _jones(2, len(comps), njchan, seed=6)
    with BeamDFT(uv, comps, ant_type, jmap, njchan, ntype=2, stok_type=1, op=op) as bm:
        bm.set_jones(jones)
        bm.set_amps(amps)
        got = bm.calc(uv.buffer, par_ang=0.2)
    model, live = orc.beam_model_fp64(uv.buffer, comps, fs, ant_type, jmap, jones, 1, 0.2, ilocb=uv.ilocb,
                                      amps=amps)
    want = orc.beam_apply(uv.buffer, model, live, uv.nrparm, uv.nif, uv.nchan, uv.incf, uv.incif, ncorr, op)
    moved = np.isnan(orc.beam_apply(uv.buffer, np.full_like(model, complex(np.nan, np.nan)), live, uv.nrparm,
                                    uv.nif, uv.nchan, uv.incf, uv.incif, ncorr, op))
    assert np.array_equal(got[~moved], uv.buffer[~moved])
    err = np.abs(got[moved].astype(np.float64) - want[moved]).max()
    if op == "div":
        assert err < 2e-5 * np.abs(want[moved]).max(), err
    else:
        assert err < 1e-5 * np.abs(model[live]).max(), err


def test_beam_antenna_parameters():
    """Rows with ANTENNA1/ANTENNA2 parameters instead of BASELINE give the
    same result as the BASELINE rows they were made from."""
    uv = _data(2, rows=600, nif=1, nchan=16, seed=41)
    cb = uv.buffer[:, uv.ilocb].astype(np.int64)
    buf2 = np.zeros((uv.nrow, uv.lrec + 2), np.float32)
    buf2[:, :uv.nrparm] = uv.buffer[:, :uv.nrparm]
    buf2[:, uv.nrparm] = (cb // 256).astype(np.float32)          # ANTENNA1
    buf2[:, uv.nrparm + 1] = (cb % 256).astype(np.float32)       # ANTENNA2
    buf2[:, uv.nrparm + 2:] = uv.buffer[:, uv.nrparm:]
    buf2[:, uv.ilocb] = 0.0                                      # no BASELINE
    uv2 = synth.UVData(buffer=buf2, nchan=uv.nchan, nif=uv.nif, nstok=2, freq=uv.freq, freqArr=uv.freqArr,
                       crval_stokes=-5.0, nrparm=uv.nrparm + 2)
    comps = _comps(True)
    rng = np.random.default_rng(2)
    at = rng.integers(0, 2, NANT)
    jones = _jones(2, len(comps), 2)
    jmap = np.arange(uv.nchan) // 8
    with BeamDFT(uv, comps, at, jmap, 2, ntype=2, stok_type=3) as bm:
        bm.set_jones(jones)
        ref = bm.calc(uv.buffer, par_ang=0.5)
    with BeamDFT(uv2, comps, at, jmap, 2, ntype=2, stok_type=3, iloca=(uv.nrparm, uv.nrparm + 1)) as bm:
        bm.set_jones(jones)
        got = bm.calc(buf2, par_ang=0.5)
    assert np.array_equal(got[:, uv.nrparm + 2:], ref[:, uv.nrparm:])


def test_beam_images_refusals():
    uv = _data(2, rows=64, nif=1, nchan=8)
    comps = _comps(False, ncomp=4)
    descs, images = _beams(1, 2, 0, nx=16, ny=16)
    with BeamDFT(uv, comps, np.zeros(NANT, int), np.arange(8) // 4, 2, ntype=1) as bm:
        with pytest.raises(Exception, match="no beam images"):
            bm.jones_from_beams(np.zeros((4, 2), np.float32))
        with pytest.raises(Exception, match="projection"):
            bm.set_beams([dict(descs[0], proj=5)], images, [1e9, 1.1e9], 1e-10, FBLANK)
        with pytest.raises(Exception, match="hwidth"):
            bm.set_beams([dict(descs[0], hwidth=5)], images, [1e9, 1.1e9], 1e-10, FBLANK)
    with BeamDFT(uv, comps, np.zeros(NANT, int), np.arange(8) // 2, 4, ntype=1) as bm:
        with pytest.raises(Exception, match="beam planes"):         # the library's own check
            bm.set_beams(descs, images, [1e9, 1.1e9, 1.2e9, 1.3e9], 1e-10, FBLANK)
        with pytest.raises(ValueError, match="beam frequencies"):
            bm.set_beams(descs, images, [1e9, 1.1e9], 1e-10, FBLANK)


def test_beam_channel_if_selection():
    """Only channels chanb..chane of IFs IFb..IFe change (the reference's
    selection test); the rest is bit-identical."""
    uv = _data(4, rows=700, nif=2, nchan=40, seed=51)
    comps = _comps(True)
    rng = np.random.default_rng(5)
    at = rng.integers(0, 2, NANT)
    nf = uv.nif * uv.nchan
    jmap = (np.arange(nf) // uv.nchan) * 5 + (np.arange(nf) % uv.nchan) // 8
    njchan = int(jmap.max()) + 1
    jones = _jones(2, len(comps), njchan, seed=7)
    sel = dict(chan_sel=(5, 30), if_sel=(1, 1))
    with BeamDFT(uv, comps, at, jmap, njchan, ntype=2, stok_type=4, op="sub", **sel) as bm:
        bm.set_jones(jones)
        got = bm.calc(uv.buffer, par_ang=0.1)
    model, live = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, at, jmap, jones, 4, 0.1,
                                      ilocb=uv.ilocb)
    want = orc.beam_apply(uv.buffer, model, live, uv.nrparm, uv.nif, uv.nchan, uv.incf, uv.incif, 4, "sub", **sel)
    moved = np.isnan(orc.beam_apply(uv.buffer, np.full_like(model, complex(np.nan, np.nan)), live, uv.nrparm,
                                    uv.nif, uv.nchan, uv.incf, uv.incif, 4, "sub", **sel))
    assert moved.any()
    d = moved[:, uv.nrparm:].reshape(uv.nrow, uv.nif, uv.nchan, -1)
    assert not d[:, 0].any() and not d[:, 1, :5].any() and not d[:, 1, 31:].any()
    assert np.array_equal(got[~moved], uv.buffer[~moved])
    assert np.abs(got[moved].astype(np.float64) - want[moved]).max() < 1e-5 * np.abs(model[live]).max()
    with pytest.raises(Exception, match="selection"):
        BeamDFT(uv, comps, at, jmap, njchan, ntype=2, chan_sel=(10, 40))
