"""GPU gridding parity: device sample arithmetic vs the CPU restatement
(bit-exact indices, taps and weighted values), gridded planes vs the FP64
oracle on the same samples, flip and trim/centre vs Obit's CPU routines."""
import numpy as np
import pytest

from obit_amd import synth

pytestmark = pytest.mark.gpu


def _facets_cfg4(n=3, nx=1024):
    return synth.make_facets(n, nx, 0.4 / 3600.0, 0.3 * np.pi / 180.0)


@pytest.fixture(scope="module")
def uv4():
    uv, _, _ = synth.workload(4, rows=384)
    return uv


def _gpu_grid(uv, facets, **kw):
    from obit_amd.gridder import GPUGrid
    return GPUGrid(uv, facets, **kw)


@pytest.mark.parametrize("ifacet", [0, 1, 2])
def test_prep_bit_exact(oracle, uv4, ifacet):
    """Per-sample gridding arithmetic against the CPU gridder's: first
    column and row, both GCF tap blocks bit-identical; the weighted,
    phase-shifted value within 1e-6 of its magnitude (the CPU rounds the
    reduced phase to FP32 radians for a Cephes sincos, the device to FP32
    turns for a quadrant-reduced series: each within ~4e-7 rad)."""
    facets = _facets_cfg4()
    with _gpu_grid(uv4, facets) as gg:
        iu, iv, itu, itv, vis = gg.prep_debug(ifacet, uv4.buffer)
    oiu, oiv, oitu, oitv, ovis = oracle.grid_prep(uv4, facets[ifacet])
    ok = ovis.reshape(-1, 2).any(axis=1)
    print(f"facet {ifacet}: {ok.sum()} valid samples of {ok.size}", flush=True)
    assert ok.sum() > 0.5 * ok.size
    for name, a, b in (("iu", iu, oiu), ("iv", iv, oiv), ("itu", itu, oitu), ("itv", itv, oitv)):
        bad = np.flatnonzero(a[ok] != b[ok])
        assert bad.size == 0, f"{name}: {bad.size} mismatches, first {a[ok][bad[:5]]} vs {b[ok][bad[:5]]}"
    g = vis.reshape(-1, 2)[ok].astype(np.float64)
    o = ovis.reshape(-1, 2)[ok].astype(np.float64)
    mag = np.hypot(o[:, 0], o[:, 1])
    dv = np.hypot(g[:, 0] - o[:, 0], g[:, 1] - o[:, 1])
    print(f"weighted values: max |d|/|v| {(dv / mag).max():.2e}", flush=True)
    assert (dv <= 1e-6 * mag).all(), "weighted visibilities differ"
    # samples the CPU drops are dropped here too (zero value)
    assert not vis.reshape(-1, 2)[~ok].any()


def _planes_vs_oracle(oracle, uv, facets, gg, plane_bins=None, **okw):
    worst = 0.0
    for i, f in enumerate(facets):
        g = gg.to_host(i)
        nplane = g.shape[0]
        for p in range(nplane):
            kw = dict(okw)
            if plane_bins is not None:
                kw["bChan"], kw["eChan"] = plane_bins[p]
            ref = oracle.trim_centre(oracle.grid_fp64(uv, f, **kw), f)
            err = np.abs(g[p] - ref).max() / np.abs(ref).max()
            worst = max(worst, err)
    return worst


def test_grid_single_plane(oracle, uv4):
    facets = _facets_cfg4()
    with _gpu_grid(uv4, facets) as gg:
        gg.grid(uv4.buffer)
        gg.flip()
        err = _planes_vs_oracle(oracle, uv4, facets, gg)
    print(f"max |dG|/peak {err:.3e}", flush=True)
    assert err < 2e-6


def test_grid_unflipped_and_flip(oracle, uv4):
    """Planes before and after the device flip against the oracle grid
    before and after ThreadFlip."""
    f = _facets_cfg4(1)
    with _gpu_grid(uv4, f) as gg:
        gg.grid(uv4.buffer)
        raw = gg.to_host(0)[0]
        gg.flip()
        flipped = gg.to_host(0)[0]
    ref_raw = oracle.trim_centre(oracle.grid_fp64(uv4, f[0], flip=False), f[0])
    ref_flip = oracle.trim_centre(oracle.grid_fp64(uv4, f[0], flip=True), f[0])
    peak = np.abs(ref_flip).max()
    assert np.abs(raw - ref_raw).max() < 2e-6 * peak
    assert np.abs(flipped - ref_flip).max() < 2e-6 * peak
    assert np.abs(ref_flip - ref_raw).max() > 1e-3 * peak  # the flip did something


def test_grid_mf_bins_cpu_parity(oracle):
    """8 coarse bins of 32 channels, the CPU's exclusive upper bound (last
    channel of each bin left out), 2 facets."""
    from obit_amd.gridder import plane_of_channel
    uv, _, _ = synth.workload(4, rows=192)
    facets = _facets_cfg4(2)
    b, e = synth.coarse_bins(256, 8)
    poc = plane_of_channel(256, b, e, cpu_parity=True)
    assert (poc[e] == 8 + 100).all() and (poc[b] == np.arange(8)).all()
    with _gpu_grid(uv, facets, nplane=8, plane_of_chan=poc) as gg:
        gg.grid(uv.buffer)
        gg.flip()
        err = _planes_vs_oracle(oracle, uv, facets, gg, plane_bins=[(int(b[p]), int(e[p])) for p in range(8)])
    print(f"mf max |dG|/peak {err:.3e}", flush=True)
    assert err < 2e-6


def test_grid_beam_and_taper(oracle, uv4):
    f = _facets_cfg4(1)
    taper = -1.0e-10
    with _gpu_grid(uv4, f, beam=True, bmTaper=taper) as gg:
        gg.grid(uv4.buffer)
        gg.flip()
        g = gg.to_host(0)[0]
    ref = oracle.trim_centre(oracle.grid_fp64(uv4, f[0], isBeam=True, bmTaper=taper), f[0])
    err = np.abs(g - ref).max() / np.abs(ref).max()
    print(f"beam+taper max |dG|/peak {err:.3e}", flush=True)
    # expf on the device and in libm may differ by an ulp
    assert err < 5e-6


def test_grid_device_path_matches_buffers(uv4):
    """One resident-device pass == the per-buffer (nVisPIO) path."""
    from obit_amd import _lib
    lib = _lib.load()
    f = _facets_cfg4(2)
    with _gpu_grid(uv4, f, nVisPIO=100) as gg:
        gg.grid(uv4.buffer)
        gg.flip()
        a = [gg.to_host(i) for i in range(2)]
    import ctypes as C
    dptr = C.c_void_p()
    nbytes = uv4.buffer.nbytes
    _lib.check(lib.ObitHIPMalloc(C.byref(dptr), nbytes))
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(dptr, _lib.ptr(uv4.buffer), nbytes))
        with _gpu_grid(uv4, f) as gg:
            gg.grid_device(dptr.value, uv4.nrow)
            gg.flip()
            b = [gg.to_host(i) for i in range(2)]
            c = [gg.to_host_device_path(i) for i in range(2)]
    finally:
        lib.ObitHIPFree(dptr)
    for x, y, z in zip(a, b, c):
        assert np.abs(x - y).max() <= 1e-6 * np.abs(x).max()
        assert np.array_equal(y, z)


@pytest.mark.parametrize("nx", [1024, 2048, 512])
def test_fft2image_matches_oracle(oracle, uv4, nx):
    """Beam and image of an nx^2 facet through the device FFT, quadrant
    swap, correction and normalisation (ObitUVGridFFT2Im) against the
    oracle's FP64 restatement applied to the device's own planes, with the
    library's correction functions (pinned separately against their
    definition).  1024 and 2048 take the column kernel and a batched 1-D
    C2R, 512 the batched 2-D C2R."""
    import ctypes as C
    from obit_amd import _lib
    f = _facets_cfg4(1, nx)
    cf = oracle.convfunc()[0]
    n = f[0].nx
    corr = np.zeros(n, np.float32)
    _lib.load().ObitHostGridCorrFn(n, n // 2 + 1, cf.ctypes.data_as(C.POINTER(C.c_float)), 7, 100,
                                   corr.ctypes.data_as(C.POINTER(C.c_float)))
    # the transform consumes the grid planes: take them first
    with _gpu_grid(uv4, f, beam=True) as gg:
        gg.grid(uv4.buffer)
        gg.flip()
        planes_b = gg.to_host(0)
        beam, norm = gg.fft_image(0, True)
    with _gpu_grid(uv4, f) as gg:
        gg.grid(uv4.buffer)
        gg.flip()
        planes_i = gg.to_host(0)
        img, _ = gg.fft_image(0, False, norm)
    rb, rnorm = oracle.fft2image(planes_b, cf, True, corr=(corr, corr))
    ri, _ = oracle.fft2image(planes_i, cf, False, rnorm, corr=(corr, corr))
    assert abs(norm[0] / rnorm[0] - 1) < 1e-5
    assert abs(beam[0, n // 2, n // 2] - 1.0) < 1e-5
    # The correction grows to ~230 per axis at the field edge, so FP32 FFT
    # rounding (~1e-7 of the peak) is amplified ~5e4 times in the corners,
    # as in the reference's FP32 FFTW path: compare on the uncorrected scale.
    taper = (corr[:, None] * corr[None, :]).astype(np.float64)
    eb = np.abs((beam[0] - rb[0]) / taper).max() / np.abs(rb[0] / taper).max()
    ei = np.abs((img[0] - ri[0]) / taper).max() / np.abs(ri[0] / taper).max()
    print(f"uncorrected scale: beam max |d|/peak {eb:.2e}, image {ei:.2e}; beam norm {norm[0]:.6g}",
          flush=True)
    assert eb < 1e-6 and ei < 1e-6


def test_grid_split_lists(oracle):
    """A small facet with a fine cell (the uv coverage within 45 cells of
    the centre) collects up to about 2,000 tracks in each of its central
    tiles, so those lists are split over several work items that add into
    the grid with device atomics.  The plane still matches the oracle."""
    import ctypes as C
    from obit_amd import _lib
    from obit_amd.gridder import GPUGrid
    lib = _lib.load()
    uv, _, _ = synth.workload(4, rows=4000)
    f = synth.make_facets(1, 256, 0.5 * 0.4 / 3600.0, 0.0)
    dptr = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(dptr), uv.buffer.nbytes))
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(dptr, _lib.ptr(uv.buffer), uv.buffer.nbytes))
        with GPUGrid(uv, f) as gg:
            gg.grid_device(dptr.value, uv.nrow)
            st = (C.c_longlong * 8)()
            _lib.check(lib.ObitHIPGridListStats(gg.c, 0, st))
            gg.flip()
            g = gg.to_host(0)[0]
    finally:
        lib.ObitHIPFree(dptr)
    print(f"lists {st[0]}, entries {st[2]}, items {st[3]}, longest {st[4]}, split {st[5]}", flush=True)
    assert st[5] > 0
    ref = oracle.trim_centre(oracle.grid_fp64(uv, f[0]), f[0])
    err = np.abs(g - ref).max() / np.abs(ref).max()
    # up to ~10^4 samples a cell here: FP32 accumulation rounding alone
    # reaches ~1e-5 of the peak, as Obit's CPU gridder shows on the same
    # samples (a lost or doubled sample would cost ~3e-3 of the peak)
    cpu = oracle.trim_centre(oracle.grid_obit(uv, f[0]), f[0])
    err_cpu = np.abs(cpu - ref).max() / np.abs(ref).max()
    print(f"split lists: max |dG|/peak {err:.3e} (Obit's CPU gridder {err_cpu:.3e})", flush=True)
    assert err < max(2e-6, 1.5 * err_cpu)


def test_grid_reuse_lists(uv4):
    """A second pass over rows with the same uvw and new visibilities that
    reuses the first pass's track lists (ObitHIPGridGridDeviceReuse) adds
    what a full pass adds; a reuse without lists, or over a different row
    count, is refused."""
    import ctypes as C
    from obit_amd import _lib
    lib = _lib.load()
    f = _facets_cfg4(2)
    rng = np.random.default_rng(3)
    b2 = uv4.buffer.copy()
    d = b2[:, uv4.nrparm:].reshape(uv4.nrow, -1, 3)
    d[..., 0] = rng.normal(size=d.shape[:-1])
    d[..., 1] = rng.normal(size=d.shape[:-1])
    nbytes = uv4.buffer.nbytes
    dA, dB = C.c_void_p(), C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(dA), nbytes))
    _lib.check(lib.ObitHIPMalloc(C.byref(dB), nbytes))
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(dA, _lib.ptr(uv4.buffer), nbytes))
        _lib.check(lib.ObitHIPMemcpyH2D(dB, _lib.ptr(b2), nbytes))
        with _gpu_grid(uv4, f) as gg:
            gg.grid_device(dA.value, uv4.nrow)
            gg.grid_device(dB.value, uv4.nrow)
            gg.flip()
            full = [gg.to_host(i) for i in range(2)]
        with _gpu_grid(uv4, f) as gg:
            with pytest.raises(Exception, match="reuse"):
                gg.grid_device_reuse(dB.value, uv4.nrow)
            gg.grid_device(dA.value, uv4.nrow)
            gg.grid_device_reuse(dB.value, uv4.nrow)
            with pytest.raises(Exception, match="reuse"):
                gg.grid_device_reuse(dB.value, uv4.nrow - 1)
            gg.flip()
            reused = [gg.to_host(i) for i in range(2)]
        # as many rows but other u, v, w (another chunk of the data set):
        # refused, and the grids are left as they were
        b3 = b2.copy()
        b3[:, 0] *= np.float32(1.001)
        _lib.check(lib.ObitHIPMemcpyH2D(dB, _lib.ptr(b3), nbytes))
        with _gpu_grid(uv4, f) as gg:
            gg.grid_device(dA.value, uv4.nrow)
            before = gg.to_host(0)
            with pytest.raises(Exception, match="differ"):
                gg.grid_device_reuse(dB.value, uv4.nrow)
            assert np.array_equal(gg.to_host(0), before)
    finally:
        lib.ObitHIPFree(dA)
        lib.ObitHIPFree(dB)
    for x, y in zip(full, reused):
        assert np.abs(x - y).max() <= 1e-6 * np.abs(x).max()


def test_workload4_facets_at_size(oracle):
    """Workload 4 at its facet size: 8 coarse bins (32 channels each, all
    channels gridded) of 2048^2 for the centre facet and the outermost
    facet of the 40-facet hex layout (the largest shift and rotation), on a
    seeded 2,000-row sample of the full geometry, against the FP64 oracle
    plane by plane."""
    from obit_amd.gridder import GPUGrid, plane_of_channel
    uv, _, facets = synth.workload(4, rows=2000)
    fs = [facets[0], facets[39]]
    assert fs[0].nx == 2048 and len(facets) == 40
    b, e = synth.coarse_bins(256, 8)
    poc = plane_of_channel(256, b, e, cpu_parity=False)
    worst = 0.0
    with GPUGrid(uv, fs, nplane=8, plane_of_chan=poc) as gg:
        gg.grid(uv.buffer)
        gg.flip()
        for i, f in enumerate(fs):
            g = gg.to_host(i)
            assert g.shape == (8, 2048, 1025)
            for p in range(8):
                ref = oracle.trim_centre(oracle.grid_fp64(uv, f, bChan=int(b[p]), eChan=int(e[p]) + 1), f)
                err = np.abs(g[p] - ref).max() / np.abs(ref).max()
                worst = max(worst, err)
    print(f"2048^2 facets 0 and 39 x 8 planes: max |dG|/peak {worst:.3e}", flush=True)
    assert worst < 2e-6


def test_workload4_full_pass_linearity():
    """The full workload 4 (210,600 rows x 256 channels into 40 facets x 8
    planes of 2048^2) in one resident pass equals, cell by cell, the sum of
    the passes over its even and its odd rows: a size-independent check
    that no sample is lost or gridded twice at the full list sizes (split
    lists, halo atomics), on four facets' planes."""
    import ctypes as C
    from obit_amd import _lib
    from obit_amd.gridder import GPUGrid, plane_of_channel
    lib = _lib.load()
    uv, _, facets = synth.workload(4)
    assert uv.nrow == 210600
    b, e = synth.coarse_bins(256, 8)
    poc = plane_of_channel(256, b, e, cpu_parity=False)
    halves = [np.ascontiguousarray(uv.buffer[0::2]), np.ascontiguousarray(uv.buffer[1::2])]
    check = [0, 7, 20, 39]
    dptr = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(dptr), uv.buffer.nbytes))
    try:
        out = []
        for buf in [uv.buffer] + halves:
            _lib.check(lib.ObitHIPMemcpyH2D(dptr, _lib.ptr(buf), buf.nbytes))
            with GPUGrid(uv, facets, nplane=8, plane_of_chan=poc) as gg:
                gg.grid_device(dptr.value, buf.shape[0])
                st = (C.c_longlong * 8)()
                _lib.check(lib.ObitHIPGridListStats(gg.c, 0, st))
                out.append([gg.to_host(i) for i in check])
            print(f"{buf.shape[0]} rows: entries {st[2]}, items {st[3]}, split lists {st[5]}, "
                  f"tile samples {st[6]}", flush=True)
    finally:
        lib.ObitHIPFree(dptr)
    worst = 0.0
    for k in range(len(check)):
        full, a, bb = out[0][k], out[1][k], out[2][k]
        for p in range(8):
            pk = np.abs(full[p]).max()
            assert pk > 0
            worst = max(worst, float(np.abs(full[p] - (a[p] + bb[p])).max() / pk))
    print(f"full pass vs even + odd rows: max |d|/peak {worst:.3e}", flush=True)
    assert worst < 1e-5
