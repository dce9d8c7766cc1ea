"""GPU DFT parity: the split-precision kernel against the FP64 oracle
(max |dV| <= 1e-5 x peak |V_model|), the operation semantics of Obit's CPU
DFT (subtract / divide / replace, flags, selections, flip, Stokes factor),
the per-buffer, resident, graph and FP64-comparator paths."""
import numpy as np
import pytest

from obit_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5  # max |dV| / peak |V_model| against the FP64 oracle


def _model_from(uv, buf, stokes=0):
    d = buf[:, uv.nrparm:].reshape(uv.nrow, uv.nif, uv.nchan, uv.nstok, 3)
    return d[..., stokes, 0] + 1j * d[..., stokes, 1]


def _run(uv, sky, path="resident", **kw):
    from obit_amd.skymodel import SkyModelDFT
    variant = kw.pop("variant", None)
    prec = kw.pop("prec", None)
    with SkyModelDFT(uv, sky, **kw) as sm:
        if prec is not None:
            sm.set_precision(prec)
        if variant is not None:
            sm.set_variant(*variant)
        if path == "buffer":
            return sm.calc(uv.buffer.copy())
        sm.set_vis(uv.buffer)
        if path == "resident":
            sm.calc_resident()
        elif path == "graph":
            sm.graph_launch()
            sm.graph_launch()  # replayed: replace is idempotent
        elif path == "fp64":
            sm.calc_fp64()
        return sm.get_vis()


@pytest.fixture(scope="module")
def w1():
    uv, sky, _ = synth.workload(1, rows=1024)
    return uv, sky


@pytest.mark.parametrize("variant", [(16, 32, 1), (16, 64, 2), (32, 64, 2), (32, 128, 4), (64, 64, 2)])
def test_point_model_vs_fp64(oracle, w1, variant):
    uv, sky = w1
    out = _run(uv, sky, op="repl", variant=variant)
    ref = oracle.dft_model_fp64(uv, sky)
    got = _model_from(uv, out)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    rms = np.sqrt(np.mean(np.abs(got - ref) ** 2)) / np.abs(ref).max()
    print(f"variant {variant}: max {err:.3e} rms {rms:.3e}", flush=True)
    assert err <= TOL
    assert np.array_equal(_model_from(uv, out, 1), got)  # LL = RR model


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_precision_policies(oracle, w1, prec):
    uv, sky = w1
    out = _run(uv, sky, op="repl", prec=prec)
    ref = oracle.dft_model_fp64(uv, sky)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"prec {prec}: max {err:.3e}", flush=True)
    if prec != 0:
        assert err <= TOL
    else:
        assert err <= 1e-3  # hardware sin/cos for the step: reported, not the default


def test_subtract_matches_cpu_semantics(oracle, w1):
    uv, sky = w1
    out = _run(uv, sky, op="sub")
    ref = oracle.dft_model_fp64(uv, sky)
    d0 = uv.buffer[:, 5:].reshape(uv.nrow, 1, 64, 2, 3)
    d1 = out[:, 5:].reshape(d0.shape)
    peak = np.abs(ref).max()
    for s in range(2):
        sub = (d0[..., s, 0] + 1j * d0[..., s, 1]) - (d1[..., s, 0] + 1j * d1[..., s, 1])
        assert np.abs(sub - ref).max() <= TOL * peak + 4e-7 * np.abs(d0[..., s, :2]).max()
    assert np.array_equal(d0[..., 2], d1[..., 2])  # weights untouched
    assert np.array_equal(out[:, :5], uv.buffer[:, :5])  # random parameters untouched


def test_divide_and_flags(oracle):
    """Flags follow the CPU DFT: a correlation with weight <= 0 is left
    alone by subtract and divide and overwritten by replace; divide gives
    data/model with weight x |model| (ObitSkyModel.c:2794-2844)."""
    uv, sky, _ = synth.workload(1, rows=300)
    dv = uv.data_view()
    rng = np.random.default_rng(7)
    dv[..., 2][rng.uniform(size=dv.shape[:-1]) < 0.1] = 0.0
    dv[..., 2][rng.uniform(size=dv.shape[:-1]) < 0.05] = -1.0
    fl = dv[..., 2] <= 0
    m = oracle.dft_model_fp64(uv, sky)[..., None]  # [row, IF, chan, 1]
    peak = np.abs(m).max()
    data = dv[..., 0] + 1j * dv[..., 1]
    for op in ("div", "sub", "repl"):
        out = _run(uv, sky, op=op)
        o = out[:, 5:].reshape(dv.shape)
        oc = o[..., 0] + 1j * o[..., 1]
        cpu = oracle.dft_obit(uv, sky, op=op)[:, 5:].reshape(dv.shape)
        if op == "repl":
            assert np.array_equal(o[..., 2], dv[..., 2])
            assert np.abs(oc - m).max() <= 1e-5 * peak
            continue
        assert np.array_equal(o[fl], dv[fl]), op       # flagged: untouched
        assert np.array_equal(cpu[fl], dv[fl]), op     # as in the CPU DFT
        ok = ~fl
        if op == "sub":
            exp = data - m
            assert np.abs(oc - exp)[ok].max() <= 1e-5 * peak + 2e-7 * np.abs(data).max()
        else:
            exp = data / m
            tol = 1e-5 * peak * np.abs(data) / np.abs(m) ** 2 + 1e-6 * np.abs(exp)
            assert (np.abs(oc - exp) <= tol)[ok].all()
            am = np.abs(m)  # broadcasts over Stokes
            wexp = dv[..., 2] * am
            wtol = (2e-5 * peak / am + 1e-6) * np.abs(wexp)
            assert (np.abs(o[..., 2] - wexp) <= wtol)[ok].all()


def test_selection_flip_and_stokes_factor(oracle):
    uv, sky, _ = synth.workload(2, rows=256, ncomp=500)
    kw = dict(op="sub", stokFactor=-1.0, doFlip=True)
    out = _run(uv, sky, startChannel=5, numberChannel=40, startIF=2, numberIF=2, **kw)
    ref = oracle.dft_model_fp64(uv, sky, doFlip=True, startChan=4, numChan=40, startIF=1, numIF=2)
    d0 = uv.buffer[:, 5:].reshape(uv.nrow, 4, 64, 2, 3)
    d1 = out[:, 5:].reshape(d0.shape)
    peak = np.abs(ref).max()
    sel = np.zeros((4, 64), bool)
    sel[1:3, 4:44] = True
    for s, sign in ((0, 1.0), (1, -1.0)):
        sub = (d0[..., s, 0] + 1j * d0[..., s, 1]) - (d1[..., s, 0] + 1j * d1[..., s, 1])
        assert np.abs(sub[:, sel] - sign * ref[:, sel]).max() <= TOL * peak + 1e-6
        assert np.array_equal(d1[:, ~sel, s], d0[:, ~sel, s])  # outside the selection: untouched


def test_paths_agree(oracle, w1):
    """The per-buffer path (ObitCUDASkyModelDFTCalc: row segments on the
    caller's streams, channels per thread chosen to fill the GPU) and the
    resident path agree within the split tolerance; graph replay equals the
    resident launch exactly."""
    uv, sky = w1
    a = _run(uv, sky, op="repl", path="buffer")
    b = _run(uv, sky, op="repl", path="resident")
    ref = oracle.dft_model_fp64(uv, sky)
    peak = np.abs(ref).max()
    assert np.abs(_model_from(uv, a) - ref).max() <= TOL * peak
    assert np.abs(_model_from(uv, a) - _model_from(uv, b)).max() <= TOL * peak
    assert np.array_equal(a[:, :5], uv.buffer[:, :5])
    g = _run(uv, sky, op="repl", path="graph")
    assert np.array_equal(g, b)


def test_fp64_comparator(oracle, w1):
    uv, sky = w1
    out = _run(uv, sky, op="repl", path="fp64")
    ref = oracle.dft_model_fp64(uv, sky)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"fp64 comparator: max {err:.3e}", flush=True)
    assert err <= 2e-7  # FP32 output rounding only


def test_gaussian_mixed_table(oracle):
    uv, sky, _ = synth.workload(3, rows=256)
    out = _run(uv, sky, op="repl")
    ref = oracle.dft_model_fp64(uv, sky)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"gauss (cfg3, {sky.comps.shape[0]} comps): max {err:.3e}", flush=True)
    assert err <= TOL


def test_parametric_spectrum(oracle):
    uv, sky, _ = synth.workload(5, rows=48)
    out = _run(uv, sky, op="repl")
    ref = oracle.dft_model_fp64(uv, sky, nthreads=8)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"spectral (cfg5, {sky.comps.shape[0]} comps): max {err:.3e}", flush=True)
    assert err <= TOL


def test_tabulated_spectrum(oracle):
    """PointTSpec: per-component spectral index and 4 coarse-bin fluxes."""
    uv, sky0, _ = synth.workload(2, rows=256, ncomp=300)
    rng = np.random.default_rng(3)
    nspec = 4
    k = sky0.comps.shape[0]
    alpha = rng.normal(-0.7, 0.3, k)
    S = np.abs(sky0.comps[:, :1]) * rng.uniform(0.5, 1.5, (k, nspec))
    comps = np.concatenate([sky0.comps[:, :4], alpha[:, None], S], axis=1).astype(np.float32)
    specFreq = np.array([uv.freqArr[i * 64 + 32] for i in range(4)])
    specIndex = np.repeat(np.arange(4), 64).astype(np.int32)
    sky = synth.SkyModel(comps=comps, modType=20, specFreq=specFreq, specIndex=specIndex)
    out = _run(uv, sky, op="repl")
    ref = oracle.dft_model_fp64(uv, sky)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"tabulated spectrum: max {err:.3e}", flush=True)
    assert err <= TOL


def test_full_size_workload2_sampled(oracle):
    """Workload 2 at full size (907,200 rows x 256 ch x 10,000 comps):
    one resident launch, then 64 seeded rows checked against the FP64
    oracle and a checksum of the untouched weights."""
    uv, sky, _ = synth.workload(2)
    from obit_amd.skymodel import SkyModelDFT
    rows = np.sort(np.random.default_rng(11).choice(uv.nrow, 64, replace=False))
    with SkyModelDFT(uv, sky, op="sub") as sm:
        sm.set_vis(uv.buffer)
        sm.calc_resident()
        out = sm.get_vis()
    sub = uv.subset(rows)
    ref = oracle.dft_model_fp64(sub, sky, nthreads=8)
    d0 = sub.buffer[:, 5:].reshape(len(rows), 4, 64, 2, 3)
    d1 = out[rows, 5:].reshape(d0.shape)
    got = (d0[..., 0, 0] + 1j * d0[..., 0, 1]) - (d1[..., 0, 0] + 1j * d1[..., 0, 1])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"full workload 2 (sampled rows): max {err:.3e}", flush=True)
    assert err <= TOL + 4e-7 * np.abs(d0[..., :2]).max() / np.abs(ref).max()
    w0 = uv.buffer[:, 5:].reshape(uv.nrow, -1, 3)[:, :, 2]
    w1_ = out[:, 5:].reshape(uv.nrow, -1, 3)[:, :, 2]
    assert np.array_equal(w0, w1_)


def _gauss_cols(k, rng, frac=0.5):
    """gp1..gp3 columns (ObitSkyModel.c:1939-1947) for k rows, a fraction
    frac of them Gaussian (FWHM 2-30 arcsec) and the rest points (zeros)."""
    maj = rng.uniform(2.0, 30.0, k) / 3600.0
    mnr = maj * rng.uniform(0.3, 1.0, k)
    pa = rng.uniform(0.0, 180.0, k)
    gp = np.stack(synth.gauss_coeffs(maj, mnr, pa), axis=-1)
    gp[rng.uniform(size=k) >= frac] = 0.0
    return gp


def test_gauss_parametric_spectrum(oracle):
    """GaussSpec, the type of ObitCUDASkyModel.cu:492-597: rows [a, x, y, z,
    gp1, gp2, gp3, c1, c2] (coefficients at offset 7), half of them points,
    on workload 2's 4 IF x 64 channel layout, against the FP64 oracle."""
    uv, sky0, _ = synth.workload(2, rows=256, ncomp=400)
    rng = np.random.default_rng(21)
    k = sky0.comps.shape[0]
    coef = np.stack([rng.normal(-0.7, 0.3, k), rng.normal(0.0, 0.1, k)], axis=-1)
    comps = np.concatenate([sky0.comps[:, :4], _gauss_cols(k, rng), coef], axis=1).astype(np.float32)
    sky = synth.SkyModel(comps=comps, modType=11, nSpecTerm=2, imgRefFreq=float(uv.freqArr[128]))
    out = _run(uv, sky, op="repl")
    ref = oracle.dft_model_fp64(uv, sky)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"GaussSpec: max {err:.3e}", flush=True)
    assert err <= TOL
    # the spectrum matters at this tolerance: without it the model differs
    flat = synth.SkyModel(comps=comps[:, :7].copy(), modType=1)
    assert np.abs(oracle.dft_model_fp64(uv, flat) - ref).max() > 100 * TOL * np.abs(ref).max()


def test_gauss_tabulated_spectrum(oracle):
    """GaussTSpec, the type of ObitCUDASkyModel.cu:608-709: rows [a, x, y, z,
    gp1, gp2, gp3, alpha, S_1..S_4] with 4 coarse bins, half of them points,
    against the FP64 oracle."""
    uv, sky0, _ = synth.workload(2, rows=256, ncomp=300)
    rng = np.random.default_rng(23)
    nspec = 4
    k = sky0.comps.shape[0]
    alpha = rng.normal(-0.7, 0.3, k)
    S = np.abs(sky0.comps[:, :1]) * rng.uniform(0.5, 1.5, (k, nspec))
    comps = np.concatenate([sky0.comps[:, :4], _gauss_cols(k, rng), alpha[:, None], S],
                           axis=1).astype(np.float32)
    specFreq = np.array([uv.freqArr[i * 64 + 32] for i in range(4)])
    specIndex = np.repeat(np.arange(4), 64).astype(np.int32)
    sky = synth.SkyModel(comps=comps, modType=21, specFreq=specFreq, specIndex=specIndex)
    for path in ("resident", "buffer"):
        out = _run(uv, sky, op="repl", path=path)
        ref = oracle.dft_model_fp64(uv, sky)
        err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
        print(f"GaussTSpec ({path}): max {err:.3e}", flush=True)
        assert err <= TOL


def _lsq_lines(fs32, nif, nchan):
    """Per-IF least-squares line through an FP32 freqScale table (FP64)."""
    out = np.empty(nif * nchan)
    c = np.arange(nchan, dtype=np.float64)
    for i in range(nif):
        y = fs32[i * nchan:(i + 1) * nchan].astype(np.float64)
        b, a = np.polyfit(c, y, 1)
        out[i * nchan:(i + 1) * nchan] = a + b * c
    return out


def test_fp32_freqscale_caller(oracle):
    """The frequencies an unmodified ObitGPUSkyModel caller hands over: the
    FP32 fscale table only (ObitGPUSkyModel.c:265), no FP64 frequencies,
    through the per-buffer ObitCUDASkyModelDFTCalc on workload 2's layout.
    The library fits each IF's channels with a line; the test pins (1) that
    the line recovers the true frequencies far below FP32 rounding, (2) the
    kernel at those frequencies to the split tolerance, and (3) that the
    result is no less accurate against the true sky than Obit's CPU DFT fed
    the same FP32 table."""
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, _ = synth.workload(2, rows=200, ncomp=2000)
    with SkyModelDFT(uv, sky, op="repl", useFreq64=False, nVisPIO=64) as sm:
        out = sm.calc(uv.buffer.copy())
        assert sm.path() == 0  # the split kernel ran
    got = _model_from(uv, out)
    true = uv.freqArr / uv.freq
    fit = _lsq_lines(uv.fscale, uv.nif, uv.nchan)
    ferr = np.abs(fit / true - 1.0).max()
    f32err = np.abs(uv.fscale.astype(np.float64) / true - 1.0).max()
    print(f"frequency: line {ferr:.2e}, FP32 table {f32err:.2e} (relative)", flush=True)
    assert ferr < 0.5 * f32err
    ref_fit = oracle.dft_model_fp64(uv, sky, freq_ratio=fit)
    peak = np.abs(ref_fit).max()
    e_fit = np.abs(got - ref_fit).max() / peak
    ref = oracle.dft_model_fp64(uv, sky)
    cpu = _model_from(uv, oracle.dft_obit(uv, sky, op="repl", nthreads=8))
    e_gpu = np.abs(got - ref).max() / peak
    e_cpu = np.abs(cpu - ref).max() / peak
    print(f"vs FP64 at the fitted frequencies {e_fit:.2e}; vs the true sky: GPU {e_gpu:.2e}, "
          f"Obit CPU {e_cpu:.2e}", flush=True)
    assert e_fit <= TOL
    assert e_gpu <= max(TOL, e_cpu)


@pytest.mark.parametrize("freq64", [False, True])
def test_nonuniform_channels_report_fp64_kernel(oracle, freq64):
    """A channel 3 kHz off its IF's line: the split kernel cannot walk it,
    the launch runs on the FP64 kernel (reported by ObitHIPSkyModelDFTGetPath)
    at every channel's own frequency."""
    import dataclasses
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, _ = synth.workload(2, rows=128, ncomp=500)
    fa = uv.freqArr.copy()
    fa[70] += 3.0e3
    uv = dataclasses.replace(uv, freqArr=fa)
    with SkyModelDFT(uv, sky, op="repl", useFreq64=freq64, nVisPIO=64) as sm:
        out = sm.calc(uv.buffer.copy())
        assert sm.path() == 1
    fr = uv.freqArr / uv.freq if freq64 else uv.fscale.astype(np.float64)
    ref = oracle.dft_model_fp64(uv, sky, freq_ratio=fr)
    err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
    print(f"non-uniform table (freq64={freq64}): max {err:.3e}", flush=True)
    assert err <= 2e-7  # FP32 output rounding only


def _split_sky(kind, uv, sky0, rng):
    """A model of the given kind on workload 2's layout (4 IF x 64 ch)."""
    k = sky0.comps.shape[0]
    base = sky0.comps[:, :4]
    if kind == "point":
        return sky0
    if kind == "gauss":
        comps = np.concatenate([base, _gauss_cols(k, rng)], axis=1).astype(np.float32)
        return synth.SkyModel(comps=comps, modType=1)
    if kind == "gaussspec":
        coef = np.stack([rng.normal(-0.7, 0.3, k), rng.normal(0.0, 0.1, k)], axis=-1)
        comps = np.concatenate([base, _gauss_cols(k, rng), coef], axis=1).astype(np.float32)
        return synth.SkyModel(comps=comps, modType=11, nSpecTerm=2, imgRefFreq=float(uv.freqArr[128]))
    alpha = rng.normal(-0.7, 0.3, k)
    S = np.abs(base[:, :1]) * rng.uniform(0.5, 1.5, (k, 4))
    comps = np.concatenate([base, _gauss_cols(k, rng), alpha[:, None], S], axis=1).astype(np.float32)
    specFreq = np.array([uv.freqArr[i * 64 + 32] for i in range(4)])
    specIndex = np.repeat(np.arange(4), 64).astype(np.int32)
    return synth.SkyModel(comps=comps, modType=21, specFreq=specFreq, specIndex=specIndex)


@pytest.mark.parametrize("kind", ["point", "gauss", "gaussspec", "gausstspec"])
def test_buffer_component_parts(oracle, kind):
    """Small per-buffer pieces cut the component list into parts whose FP64
    sums k_dft_combine adds and applies: the model (replace) against the
    FP64 oracle, and subtract / divide against the resident launch, for
    point, Gaussian (points and Gaussians in parts on their own kernels)
    and spectral types; the last piece is ragged."""
    from obit_amd.skymodel import SkyModelDFT
    uv, sky0, _ = synth.workload(2, rows=1100, ncomp=2000)
    sky = _split_sky(kind, uv, sky0, np.random.default_rng(31))
    with SkyModelDFT(uv, sky, op="repl", nVisPIO=512) as sm:
        out = sm.calc(uv.buffer.copy())
        cpt, nsplit = sm.buffer_shape()
    print(f"{kind}: {cpt} channels per thread, {nsplit} component parts", flush=True)
    assert nsplit > 1
    if kind in ("point", "gauss"):
        # points on the matrix-unit kernel (a Gaussian table's points too,
        # its Gaussians in their own parts on the VALU kernel)
        assert cpt == 64
    ref = oracle.dft_model_fp64(uv, sky)
    peak = np.abs(ref).max()
    err = np.abs(_model_from(uv, out) - ref).max() / peak
    print(f"{kind} (replace): max {err:.3e}", flush=True)
    assert err <= TOL
    assert np.array_equal(out[:, :uv.nrparm], uv.buffer[:, :uv.nrparm])
    for op in ("sub", "div"):
        a = _run(uv, sky, op=op, path="buffer", nVisPIO=512)
        b = _run(uv, sky, op=op, path="resident")
        da = a[:, uv.nrparm:].reshape(uv.nrow, -1, 3)
        db = b[:, uv.nrparm:].reshape(uv.nrow, -1, 3)
        if op == "sub":
            assert np.abs(da[..., :2] - db[..., :2]).max() <= TOL * peak
            assert np.array_equal(da[..., 2], db[..., 2])
        else:
            # where the model is not small the quotient is well conditioned
            m = np.repeat(np.abs(ref).reshape(uv.nrow, -1), uv.nstok, axis=1) > 0.1 * peak
            rel = np.abs(da[..., :2] - db[..., :2]) / (np.abs(db[..., :2]).max() + 1e-30)
            assert rel[m].max() <= 1e-4
            assert np.allclose(da[..., 2][m], db[..., 2][m], rtol=1e-4)


@pytest.mark.parametrize("op", ["repl", "sub", "div"])
def test_matrix_core_kernel(oracle, op):
    """The matrix-core kernel (variant cpt 64: 64-channel groups as a
    complex product on v_mfma_f32_16x16x32_f16 over FP16 operand parts) on
    workload 2's 4 IF x 64 channel layout: the model against the FP64
    oracle, and subtract / divide against the VALU kernel; a ragged last
    row quad."""
    uv, sky, _ = synth.workload(2, rows=1027, ncomp=700)
    out = _run(uv, sky, op=op, variant=(64, 64, 2))
    if op == "repl":
        ref = oracle.dft_model_fp64(uv, sky)
        err = np.abs(_model_from(uv, out) - ref).max() / np.abs(ref).max()
        print(f"matrix-core kernel: max {err:.3e}", flush=True)
        assert err <= TOL
        return
    ref_run = _run(uv, sky, op=op, variant=(32, 64, 2))
    peak = np.abs(oracle.dft_model_fp64(uv, sky)).max()
    da = out[:, uv.nrparm:].reshape(uv.nrow, -1, 3)
    db = ref_run[:, uv.nrparm:].reshape(uv.nrow, -1, 3)
    if op == "sub":
        assert np.abs(da[..., :2] - db[..., :2]).max() <= 2 * TOL * peak
        assert np.array_equal(da[..., 2], db[..., 2])
    else:
        # weight x |model|: the kernels' models agree within 2 TOL of the peak
        w0 = uv.buffer[:, uv.nrparm:].reshape(uv.nrow, -1, 3)[..., 2]
        assert np.all(np.abs(da[..., 2] - db[..., 2]) <= 2 * TOL * peak * np.abs(w0) + 1e-30)


def test_matrix_core_amplitude_scale(oracle):
    """The matrix-core kernel scales its FP16 operand parts by the table's
    largest amplitude, taken again at every model load: through one
    SkyModelDFT (replace), a table of tiny amplitudes, then large ones
    with a 1e6 dynamic range, against the FP64 oracle; then an all-zero
    table, whose model is zero."""
    import dataclasses
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, _ = synth.workload(2, rows=258, ncomp=500)
    tiny = sky.comps.copy()
    tiny[:, 0] *= 1e-7
    wide = sky.comps.copy()
    wide[:, 0] *= 3e3
    wide[::7, 0] *= 1e-6
    zero = sky.comps.copy()
    zero[:, 0] = 0.0
    with SkyModelDFT(uv, dataclasses.replace(sky, comps=tiny), op="repl") as sm:
        sm.set_variant(64, 64, 2)
        for name, comps in (("tiny", tiny), ("wide", wide), ("zero", zero)):
            sm.set_model(comps)
            sm.set_vis(uv.buffer)
            sm.calc_resident()
            got = _model_from(uv, sm.get_vis())
            if name == "zero":
                assert np.all(got == 0)
                continue
            ref = oracle.dft_model_fp64(uv, dataclasses.replace(sky, comps=comps))
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"{name}: max {err:.3e}", flush=True)
            assert err <= TOL


def test_buffer_matrix_core_single_part(oracle):
    """A per-buffer call whose table is too short to cut into parts (100
    point components) runs the matrix-unit kernel on each piece directly:
    replace against the FP64 oracle, subtract against the resident launch."""
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, _ = synth.workload(2, rows=700, ncomp=100)
    with SkyModelDFT(uv, sky, op="repl", nVisPIO=256) as sm:
        out = sm.calc(uv.buffer.copy())
        cpt, nsplit = sm.buffer_shape()
    assert (cpt, nsplit) == (64, 1)
    ref = oracle.dft_model_fp64(uv, sky)
    peak = np.abs(ref).max()
    err = np.abs(_model_from(uv, out) - ref).max() / peak
    print(f"single part: max {err:.3e}", flush=True)
    assert err <= TOL
    a = _run(uv, sky, op="sub", path="buffer", nVisPIO=256)
    b = _run(uv, sky, op="sub", path="resident", variant=(64, 64, 2))
    assert np.array_equal(a, b)


def test_mixed_table_subtract(oracle):
    """A Gaussian-type table with point components (workload 3's 40
    facets, 20 % Gaussians), subtracted with variant cpt 64: its points run
    on the matrix-unit kernel and its Gaussians on the VALU kernel.  Against
    the FP64 oracle and the single VALU launch (cpt 32); then the same
    table with every Gaussian term zeroed (all points, one matrix-unit
    launch)."""
    import dataclasses
    uv, sky, _ = synth.workload(3, rows=300)
    assert (sky.comps[:, 4:7] != 0).any(axis=1).sum() > 0
    for name, s in (("mixed", sky),
                    ("points", dataclasses.replace(sky, comps=np.concatenate(
                        [sky.comps[:, :4], np.zeros_like(sky.comps[:, 4:7])], axis=1)))):
        ref = oracle.dft_model_fp64(uv, s)
        peak = np.abs(ref).max()
        a = _run(uv, s, op="sub", variant=(64, 64, 2))
        b = _run(uv, s, op="sub", variant=(32, 64, 2))
        d0 = uv.buffer[:, uv.nrparm:].reshape(uv.nrow, uv.nif, uv.nchan, uv.nstok, 3)
        da = a[:, uv.nrparm:].reshape(d0.shape)
        db = b[:, uv.nrparm:].reshape(d0.shape)
        w = d0[..., 0, 2] > 0
        got = (d0[..., 0, 0] - da[..., 0, 0]) + 1j * (d0[..., 0, 1] - da[..., 0, 1])
        err = np.abs(got - ref)[w].max()
        print(f"{name}: max {err / peak:.3e} of peak vs FP64", flush=True)
        assert err <= TOL * peak + 4e-7 * np.abs(d0[..., 0, :2]).max()
        assert np.abs(da[..., :2] - db[..., :2]).max() <= 2 * TOL * peak
        assert np.array_equal(da[..., 2], db[..., 2])
        # another kernel ran (a fallback to the VALU kernel gives cpt 32's
        # result bit for bit)
        assert not np.array_equal(da[..., :2], db[..., :2])


def test_matrix_core_ragged_channels(oracle):
    """A channel selection that is not whole 64-channel groups (channels
    9..56 of each IF's 64): the matrix-unit kernel runs each IF's 48
    channels as one group and writes only those; replace against the FP64
    oracle and the VALU kernel, subtract against the VALU kernel, and the
    channels outside the selection untouched."""
    uv, sky, _ = synth.workload(2, rows=300, ncomp=300)
    kw = dict(startChannel=9, numberChannel=48)
    a = _run(uv, sky, op="repl", variant=(64, 64, 2), **kw)
    b = _run(uv, sky, op="repl", variant=(32, 64, 2), **kw)
    assert not np.array_equal(a, b)  # the matrix-unit kernel ran
    d0 = uv.buffer[:, uv.nrparm:].reshape(uv.nrow, uv.nif, uv.nchan, uv.nstok, 3)
    da = a[:, uv.nrparm:].reshape(d0.shape)
    db = b[:, uv.nrparm:].reshape(d0.shape)
    sel = np.zeros(uv.nchan, bool)
    sel[8:56] = True  # startChannel is 1-relative
    assert np.array_equal(da[:, :, ~sel], d0[:, :, ~sel])
    ref = oracle.dft_model_fp64(uv, sky)
    peak = np.abs(ref).max()
    got = da[..., 0, 0] + 1j * da[..., 0, 1]
    err = np.abs(got - ref)[:, :, sel].max() / peak
    print(f"ragged channels: max {err:.3e}", flush=True)
    assert err <= TOL
    assert np.abs(da[..., :2] - db[..., :2]).max() <= 2 * TOL * peak
    s1 = _run(uv, sky, op="sub", variant=(64, 64, 2), **kw)
    s2 = _run(uv, sky, op="sub", variant=(32, 64, 2), **kw)
    assert np.abs(s1[:, uv.nrparm:] - s2[:, uv.nrparm:]).max() <= 2 * TOL * peak + 1e-6 * np.abs(uv.buffer).max()


def test_matrix_core_kernel_falls_back(oracle):
    """A Gaussian model with replace: the cpt 64 variant runs the VALU
    kernel, with the same result as cpt 32."""
    uv3, sky3, _ = synth.workload(3, rows=64)
    a = _run(uv3, sky3, op="repl", variant=(64, 64, 2))
    b = _run(uv3, sky3, op="repl", variant=(32, 64, 2))
    assert np.array_equal(a, b)
