"""Benchmark: DFT model subtract (and uv gridding) on one MI355X.

Headline (value): DFT evaluations per second, one evaluation = one
(visibility row, channel, component), of the split-precision HIP DFT
subtracting a 10,000-component point model from the resident 64-antenna
data set (2016 baselines x 450 integrations, 4 IF x 64 channels, RR/LL).
A step is one subtract pass over every row (one persistent launch replayed
from a hipGraph); inputs are resident in HBM before the timed region.

Also reported on the same line: the plain all-FP64 GPU DFT on a row sample
(the comparator), the split-vs-FP64 error, gridding of the 40-facet x
8-plane wideband workload (samples/s and per-facet-stream GB/s), the
roofline of the DFT kernel (FP32 VALU), and Obit's threaded CPU DFT
restated in C, timed on this host's cores on a bounded row sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 vector (AMD published)
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FLOP_PER_EVAL = 10         # complex phasor step (6) + amplitude MAC (4)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _events(lib):
    return lib.ObitCUDAEventCreate(), lib.ObitCUDAEventCreate()


def bench_dft(lib, steps, warmup, sample_rows, variant=None, accuracy=True):
    from obit_amd import synth
    from obit_amd.skymodel import SkyModelDFT
    t0 = time.time()
    uv, sky, _ = synth.workload(2)
    log(f"workload 2: {uv.nrow} rows x {uv.nif * uv.nchan} ch x {sky.comps.shape[0]} comps "
        f"({uv.buffer.nbytes / 1e9:.2f} GB) built in {time.time() - t0:.1f}s")
    evals = uv.nrow * uv.nif * uv.nchan * sky.comps.shape[0]
    res = {"rows": uv.nrow, "channels": uv.nif * uv.nchan, "components": int(sky.comps.shape[0]),
           "evals_per_step": evals}
    sm = SkyModelDFT(uv, sky, op="sub")
    try:
        sm.set_vis(uv.buffer)
        t0 = time.time()
        if variant:
            sm.set_variant(*variant)
            res["variant"] = dict(zip(("cpt", "tile", "wg_per_cu"), sm.variant()))
            log(f"variant {res['variant']} (fixed)")
        else:
            res["variant"] = dict(zip(("cpt", "tile", "wg_per_cu"), sm.autotune(sample_rows=32768, reps=1)))
            log(f"autotune -> {res['variant']} in {time.time() - t0:.1f}s")
        st = sm.stream()
        for _ in range(warmup):
            sm.graph_launch()
        lib.ObitHIPDeviceSynchronize()
        e0, e1 = _events(lib)
        lib.ObitHIPDeviceSynchronize()
        w0 = time.perf_counter()
        lib.ObitCUDAEventRecord(e0, st)
        for _ in range(steps):
            sm.graph_launch()
        lib.ObitCUDAEventRecord(e1, st)
        lib.ObitHIPDeviceSynchronize()
        wall = time.perf_counter() - w0
        kms = lib.ObitHIPEventElapsedMs(e0, e1) / steps
        res["kernel_ms"] = kms
        res["wall_ms_per_step"] = 1e3 * wall / steps
        res["evals_per_s"] = evals * steps / wall
        res["kernel_evals_per_s"] = evals / (kms * 1e-3)
        log(f"split DFT: {kms:.2f} ms/launch, {res['kernel_evals_per_s'] / 1e12:.3f} Tevals/s")
    finally:
        sm.close()
    if not accuracy:
        return res, uv, sky

    # accuracy and the FP64 comparator on a row sample
    rng = np.random.default_rng(5)
    rows = np.sort(rng.choice(uv.nrow, sample_rows, replace=False))
    sub = uv.subset(rows)
    out = {}
    for kind in ("split", "fp64"):
        with SkyModelDFT(sub, sky, op="repl") as s2:
            s2.set_vis(sub.buffer)
            if kind == "split":
                # the benched kernel variant
                s2.set_variant(*(res["variant"][k] for k in ("cpt", "tile", "wg_per_cu")))
                s2.calc_resident()
            else:
                st2 = s2.stream()
                e0, e1 = _events(lib)
                s2.calc_fp64()  # warm
                lib.ObitHIPDeviceSynchronize()
                lib.ObitCUDAEventRecord(e0, st2)
                s2.calc_fp64()
                lib.ObitCUDAEventRecord(e1, st2)
                lib.ObitHIPDeviceSynchronize()
                ms = lib.ObitHIPEventElapsedMs(e0, e1)
                ev = sub.nrow * sub.nif * sub.nchan * sky.comps.shape[0]
                res["fp64_evals_per_s"] = ev / (ms * 1e-3)
            d = s2.get_vis()[:, 5:].reshape(sub.nrow, sub.nif, sub.nchan, sub.nstok, 3)
            out[kind] = d[..., 0, 0].astype(np.float64) + 1j * d[..., 0, 1]
    peak = np.abs(out["fp64"]).max()
    diff = np.abs(out["split"] - out["fp64"])
    res["err_max_rel_peak"] = float(diff.max() / peak)
    res["err_rms_rel_peak"] = float(np.sqrt(np.mean(diff ** 2)) / peak)
    res["speedup_vs_fp64_gpu"] = res["kernel_evals_per_s"] / res["fp64_evals_per_s"]
    log(f"fp64 comparator: {res['fp64_evals_per_s'] / 1e12:.4f} Tevals/s; split/fp64 "
        f"{res['speedup_vs_fp64_gpu']:.1f}x; max|dV|/peak {res['err_max_rel_peak']:.2e}")
    return res, uv, sky


def bench_batched(lib, steps=3, sample_rows=2048):
    """Workload 3: the 40-facet model (point and Gaussian components, 10 to
    20,000 per facet) subtracted from workload 1's visibilities in one
    launch per step (graph replays, autotuned), and its error against the
    FP64 comparator on a row sample."""
    from obit_amd import synth
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, _ = synth.workload(3)
    evals = uv.nrow * uv.nif * uv.nchan * sky.comps.shape[0]
    res = {"workload": "workload 3: 40 facets, 10-20,000 components each, 20 % Gaussian; one launch",
           "rows": uv.nrow, "channels": uv.nif * uv.nchan, "components": int(sky.comps.shape[0]),
           "gaussian_components": int((sky.comps[:, 4:7] != 0).any(axis=1).sum())}
    with SkyModelDFT(uv, sky, op="sub") as sm:
        sm.set_vis(uv.buffer)
        res["variant"] = dict(zip(("cpt", "tile", "wg_per_cu"), sm.autotune(sample_rows=16384, reps=1)))
        sm.graph_launch()
        lib.ObitHIPDeviceSynchronize()
        st = sm.stream()
        e0, e1 = _events(lib)
        lib.ObitCUDAEventRecord(e0, st)
        for _ in range(steps):
            sm.graph_launch()
        lib.ObitCUDAEventRecord(e1, st)
        lib.ObitHIPDeviceSynchronize()
        ms = lib.ObitHIPEventElapsedMs(e0, e1) / steps
    res["kernel_ms"] = ms
    res["evals_per_s"] = evals / (ms * 1e-3)
    rows = np.sort(np.random.default_rng(6).choice(uv.nrow, sample_rows, replace=False))
    sub = uv.subset(rows)
    out = {}
    for kind in ("split", "fp64"):
        with SkyModelDFT(sub, sky, op="repl") as s2:
            s2.set_vis(sub.buffer)
            s2.calc_resident() if kind == "split" else s2.calc_fp64()
            d = s2.get_vis()[:, 5:].reshape(sub.nrow, sub.nif, sub.nchan, sub.nstok, 3)
            out[kind] = d[..., 0, 0].astype(np.float64) + 1j * d[..., 0, 1]
    peak = np.abs(out["fp64"]).max()
    diff = np.abs(out["split"] - out["fp64"])
    res["err_max_vs_fp64_rel_peak"] = float(diff.max() / peak)
    res["err_rms_vs_fp64_rel_peak"] = float(np.sqrt(np.mean(diff ** 2)) / peak)
    log(f"workload 3: {evals:.3e} evaluations in {ms:.1f} ms = {res['evals_per_s']:.3e}/s; "
        f"max|dV|/peak {res['err_max_vs_fp64_rel_peak']:.2e}")
    return res


def bench_pcie(lib, uv, sky, variant, rows=100000, nvispio=10000):
    """Obit's per-buffer path: ObitCUDASkyModelDFTCalc on host buffers of
    nVisPIO rows (copy in to pinned memory, H2D, kernel, D2H, copy out),
    over the first `rows` rows: the rate a caller that hands over host
    buffers sees, PCIe included.  The host class alternates two pinned
    buffers so one piece's host copies overlap the next piece's device
    work; the shares reported are the time inside ObitCUDASkyModelDFTCalc
    (H2D, kernels, D2H on the caller's streams), the host copies' own time,
    and the resident launch (the bench's variant) on one nVisPIO-row piece.
    The best of three passes over the rows is reported."""
    from obit_amd.skymodel import SkyModelDFT
    sub = uv.subset(np.arange(min(rows, uv.nrow)))
    ev = sub.nrow * sub.nif * sub.nchan * sky.comps.shape[0]
    best = None
    with SkyModelDFT(sub, sky, op="sub", nVisPIO=nvispio) as sm:
        if variant:
            sm.set_variant(*variant)
        sm.calc(sub.buffer[:nvispio].copy())  # warm
        for _ in range(3):
            buf = sub.buffer.copy()
            times = np.zeros(3)
            sm.calc(buf, times)
            if best is None or times[0] < best[0]:
                best = times.copy()
        shape = sm.buffer_shape()
    dt = best[0]
    # the resident launch on one piece of nVisPIO rows
    one = sub.subset(np.arange(nvispio))
    with SkyModelDFT(one, sky, op="sub") as sk:
        if variant:
            sk.set_variant(*variant)
        sk.set_vis(one.buffer)
        sk.calc_resident()
        lib.ObitCUDADeviceSynchronize(None)
        t0 = time.perf_counter()
        for _ in range(3):
            sk.calc_resident()
        lib.ObitCUDADeviceSynchronize(None)
        kern = (time.perf_counter() - t0) / 3
    npieces = (sub.nrow + nvispio - 1) // nvispio
    # the reference host class's own share, bounded: the same bytes copied
    # into and out of a pinned buffer by one thread (ObitGPUSkyModelDFTCalc's
    # memcpy)
    chunk = nvispio * sub.lrec * 4
    pinned = lib.ObitCUDAUtilAllocHost(chunk)
    try:
        src = sub.buffer.copy()
        t0 = time.perf_counter()
        for r0 in range(0, sub.nrow, nvispio):
            n = min(nvispio, sub.nrow - r0) * sub.lrec * 4
            p = src.ctypes.data + r0 * sub.lrec * 4
            C.memmove(pinned, p, n)
            C.memmove(p, pinned, n)
        dt_mem = time.perf_counter() - t0
    finally:
        lib.ObitCUDAUtilFreeHost(pinned)
    log(f"per-buffer path (host buffers, PCIe included): {sub.nrow} rows in {dt:.4f}s -> "
        f"{ev / dt / 1e12:.3f} Tevals/s with {shape[0]} channels per thread and {shape[1]} component "
        f"parts; per {nvispio}-row piece: {dt / npieces * 1e3:.2f} ms wall, "
        f"{best[1] / npieces * 1e3:.2f} ms in DFTCalc, {best[2] / npieces * 1e3:.2f} ms host copies "
        f"(overlapped), resident launch on one piece {kern * 1e3:.2f} ms; one-thread copies "
        f"{dt_mem / npieces * 1e3:.2f} ms")
    return {"pcie_inclusive_evals_per_s": ev / dt, "pcie_rows": sub.nrow, "nVisPIO": nvispio,
            "pcie_s": dt, "pieces": npieces, "dftcalc_s": best[1], "host_copies_s": best[2],
            "buffer_shape": {"cpt": shape[0], "component_parts": shape[1]},
            "resident_launch_per_piece_s": kern, "host_memcpy_one_thread_s": dt_mem}


def grid_accuracy(uv, facets, b, e, rows=2000, nfacet=2):
    """Gridded planes of a seeded row sample (2 facets x 8 planes) against
    the oracle's FP64 gridder on the same samples: max and RMS |dG| over
    the peak of each plane, worst plane reported."""
    from obit_amd.gridder import GPUGrid, plane_of_channel
    from oracle import oracle as orc
    rng = np.random.default_rng(17)
    sub = uv.subset(np.sort(rng.choice(uv.nrow, rows, replace=False)))
    nci = uv.nchan * uv.nif
    poc = plane_of_channel(nci, b, e, cpu_parity=False)
    fs = facets[:nfacet]
    emax = erms = 0.0
    with GPUGrid(sub, fs, nplane=len(b), plane_of_chan=poc) as gg:
        gg.grid(sub.buffer)
        gg.flip()
        for i, f in enumerate(fs):
            g = gg.to_host(i)
            for p in range(len(b)):
                ref = orc.trim_centre(orc.grid_fp64(sub, f, bChan=int(b[p]), eChan=int(e[p]) + 1), f)
                pk = np.abs(ref).max()
                d = np.abs(g[p] - ref)
                emax = max(emax, float(d.max() / pk))
                erms = max(erms, float(np.sqrt(np.mean(d[ref != 0] ** 2)) / pk))
    log(f"gridding accuracy ({rows} rows, {nfacet} facets x {len(b)} planes): max |dG|/peak {emax:.2e}, "
        f"rms {erms:.2e}")
    return {"rows": rows, "facets": nfacet, "planes": len(b), "err_max_vs_fp64_rel_peak": emax,
            "err_rms_vs_fp64_rel_peak": erms}


def bench_fft2image(lib, gg, facets, regrid):
    """Flip and FFT-to-image (ObitHIPGridFFT2Image: C2R, quadrant swap,
    gridding correction, normalisation) of every facet x plane of the
    gridded workload, timed with HIP events on the gridder's stream.  The
    transform consumes its facet's grid, so after the warm-up call (plans,
    tables) the grids are cleared and gridded again (regrid) before the
    timed calls."""
    from obit_amd import _lib
    f = facets[0]
    nplane = gg.nplane
    nbytes = nplane * f.ny * f.nx * 4
    d = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d), nbytes), "malloc image")
    norm = np.zeros(nplane, np.float32)
    pn = norm.ctypes.data_as(C.POINTER(C.c_float))
    try:
        gg.flip()
        _lib.check(lib.ObitHIPGridFFT2Image(gg.c, 0, 1, pn, d), "FFT2Image")  # plans, tables
        regrid()
        gg.flip()
        st = gg.stream()
        e0, e1 = _events(lib)
        lib.ObitCUDAEventRecord(e0, st)
        for i in range(len(facets)):
            _lib.check(lib.ObitHIPGridFFT2Image(gg.c, i, 1, pn, d), "FFT2Image")
        lib.ObitCUDAEventRecord(e1, st)
        lib.ObitHIPDeviceSynchronize()
        ms = lib.ObitHIPEventElapsedMs(e0, e1)
    finally:
        lib.ObitHIPFree(d)
    planes = len(facets) * nplane
    # per plane at least: the half-plane grid read and written by the
    # column transform, read by the row transform, and the image written
    mib = planes * f.ny * (f.nx // 2 + 1) * 8 * 3 / 2**20 + planes * f.ny * f.nx * 4 / 2**20
    log(f"fft to image: {planes} planes of {f.nx}x{f.ny} in {ms:.1f} ms")
    return {"planes": planes, "ms": ms, "planes_per_s": planes / (ms * 1e-3),
            "min_traffic_GBps": mib * 2**20 / (ms * 1e-3) / 1e9}


def bench_beam(lib, uv, sky, reps=2, check_rows=16, check=True):
    """The Jones-corrected DFT (ObitHIPBeamDFT*) on workload 2 at full size:
    every resident row, XX/YY, the 10,000-point table, one antenna type
    (a homogeneous array), 8 beam channels of 32 channels each, seeded
    near-unit Jones matrices, Stokes I, subtract; HIP events on the default
    stream.  With check (the CPU leg), a seeded row sample of the first
    launch is compared with the oracle's FP64 model (checker only, outside
    the timing)."""
    from obit_amd import _lib
    from obit_amd.beam import BeamDFT
    cb = uv.buffer[:, uv.ilocb].astype(np.int64)
    nant = int(max((cb // 256).max(), (cb % 256).max()))
    nf = uv.nif * uv.nchan
    njchan = 8
    jmap = (np.arange(nf) * njchan) // nf
    comps = np.ascontiguousarray(sky.comps[:, :4], np.float32)
    rng = np.random.default_rng(11)
    jones = np.zeros((1, comps.shape[0], njchan, 2, 2), np.complex64)
    jones[..., 0, 0] = jones[..., 1, 1] = 1.0
    jones += (0.1 * (rng.normal(size=jones.shape) + 1j * rng.normal(size=jones.shape))).astype(np.complex64)
    evals = uv.nrow * nf * comps.shape[0]
    rows = np.sort(np.random.default_rng(12).choice(uv.nrow, check_rows, replace=False))
    d = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d), uv.buffer.nbytes), "malloc vis")
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(d, _lib.ptr(uv.buffer), uv.buffer.nbytes), "upload vis")
        with BeamDFT(uv, comps, np.zeros(nant, np.int32), jmap, njchan, ntype=1, stok_type=1, op="sub") as bm:
            bm.set_jones(jones)
            bm.calc_device(d.value, uv.nrow, 0.3)            # warm; its rows are checked
            lib.ObitHIPDeviceSynchronize()
            got = np.empty((rows.size, uv.lrec), np.float32)
            for i, r in enumerate(rows):
                _lib.check(lib.ObitHIPMemcpyD2H(got[i].ctypes.data, d.value + int(r) * uv.lrec * 4, uv.lrec * 4),
                           "D2H")
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, None)
            for _ in range(reps):
                bm.calc_device(d.value, uv.nrow, 0.3)
            lib.ObitCUDAEventRecord(e1, None)
            lib.ObitHIPDeviceSynchronize()
            ms = lib.ObitHIPEventElapsedMs(e0, e1) / reps
    finally:
        lib.ObitHIPFree(d)
    flop = 24.0   # phasor step 6, amplitude 2, two complex multiply-adds 16
    res = {"workload": "workload 2 rows (907,200 x 256 ch, XX/YY) x 10,000 points, 1 antenna type, "
                       "8 beam channels, Stokes I, subtract",
           "rows": uv.nrow, "channels": nf, "components": int(comps.shape[0]), "beam_channels": njchan,
           "ms": ms, "evals_per_s": evals / (ms * 1e-3),
           "roofline": {"bound": "valu", "achieved": evals * flop / (ms * 1e-3) / 1e12, "peak": 157.3,
                        "unit": "TFLOP/s", "frac": evals * flop / (ms * 1e-3) / 1e12 / 157.3,
                        "flop_per_eval": flop}}
    if check:
        from oracle import oracle as orc
        sub = uv.buffer[rows]
        m, live = orc.beam_model_fp64(sub, comps, uv.freqArr / uv.freq, np.zeros(nant, int), jmap,
                                      jones.astype(np.complex128), 1, 0.3, ilocb=uv.ilocb)
        d0 = sub[:, uv.nrparm:].reshape(rows.size, nf, uv.nstok, 3)
        d1 = got[:, uv.nrparm:].reshape(d0.shape)
        sel = d0[..., 2] > 0
        diff = 0.0
        for h in range(2):
            mod = (d0[:, :, h, 0] - d1[:, :, h, 0]).astype(np.float64) + 1j * (d0[:, :, h, 1] - d1[:, :, h, 1])
            diff = max(diff, float(np.abs(mod - m[:, :, h, h])[sel[..., h]].max()))
        res["err_max_vs_fp64_rel_peak"] = diff / float(np.abs(m).max())
        res["check_rows"] = check_rows
    log(f"beam DFT: {evals:.3e} evaluations in {ms:.1f} ms = {res['evals_per_s']:.3e}/s "
        f"({res['roofline']['frac'] * 100:.1f} % of FP32 at {flop:.0f} FLOP each)"
        + (f"; max|dV|/peak {res['err_max_vs_fp64_rel_peak']:.2e}" if check else ""))
    return res


def bench_beam4(lib, uv, sky, rows=226800, reps=2):
    """The beam-corrected DFT with 4 correlations (XX, YY, XY, YX), Stokes Q,
    over the first quarter of workload 2's rows (the cross hands copied
    from the parallel ones), otherwise as bench_beam; 40 FLOP per
    evaluation (phasor step 6, amplitude 2, four complex multiply-adds 32)."""
    from obit_amd import _lib, synth
    from obit_amd.beam import BeamDFT
    nf = uv.nif * uv.nchan
    src = uv.buffer[:rows]
    d2 = src[:, uv.nrparm:].reshape(rows, nf, 2, 3)
    buf = np.empty((rows, uv.nrparm + nf * 12), np.float32)
    buf[:, :uv.nrparm] = src[:, :uv.nrparm]
    d4 = buf[:, uv.nrparm:].reshape(rows, nf, 4, 3)
    d4[:, :, :2] = d2
    d4[:, :, 2:] = d2
    uv4 = synth.UVData(buffer=buf, nchan=uv.nchan, nif=uv.nif, nstok=4, freq=uv.freq, freqArr=uv.freqArr,
                       crval_stokes=-5.0, nrparm=uv.nrparm)
    cb = src[:, uv.ilocb].astype(np.int64)
    nant = int(max((cb // 256).max(), (cb % 256).max()))
    njchan = 8
    jmap = (np.arange(nf) * njchan) // nf
    comps = np.ascontiguousarray(sky.comps[:, :4], np.float32)
    rng = np.random.default_rng(11)
    jones = np.zeros((1, comps.shape[0], njchan, 2, 2), np.complex64)
    jones[..., 0, 0] = jones[..., 1, 1] = 1.0
    jones += (0.1 * (rng.normal(size=jones.shape) + 1j * rng.normal(size=jones.shape))).astype(np.complex64)
    evals = rows * nf * comps.shape[0]
    d = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d), buf.nbytes), "malloc vis")
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(d, _lib.ptr(buf), buf.nbytes), "upload vis")
        with BeamDFT(uv4, comps, np.zeros(nant, np.int32), jmap, njchan, ntype=1, stok_type=2, op="sub") as bm:
            bm.set_jones(jones)
            bm.calc_device(d.value, rows, 0.3)
            lib.ObitHIPDeviceSynchronize()
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, None)
            for _ in range(reps):
                bm.calc_device(d.value, rows, 0.3)
            lib.ObitCUDAEventRecord(e1, None)
            lib.ObitHIPDeviceSynchronize()
            ms = lib.ObitHIPEventElapsedMs(e0, e1) / reps
    finally:
        lib.ObitHIPFree(d)
    flop = 40.0
    res = {"workload": f"first {rows} rows of workload 2 x 256 ch, XX/YY/XY/YX, 10,000 points, 1 antenna type, "
                       "8 beam channels, Stokes Q, subtract",
           "rows": rows, "channels": nf, "components": int(comps.shape[0]), "ms": ms,
           "evals_per_s": evals / (ms * 1e-3),
           "roofline": {"bound": "valu", "achieved": evals * flop / (ms * 1e-3) / 1e12, "peak": 157.3,
                        "unit": "TFLOP/s", "frac": evals * flop / (ms * 1e-3) / 1e12 / 157.3,
                        "flop_per_eval": flop}}
    log(f"beam DFT, 4 correlations: {evals:.3e} evaluations in {ms:.1f} ms = {res['evals_per_s']:.3e}/s "
        f"({res['roofline']['frac'] * 100:.1f} % of FP32 at {flop:.0f} FLOP each)")
    return res


def bench_stokes(lib, uv, reps=5, check_rows=2000, check=True):
    """Stokes I records of workload 2's resident RR/LL rows
    (ObitHIPStokesI, the gridder's input), timed with HIP events on the
    default stream.  With check (the CPU leg), a seeded row sample is
    compared with the oracle's restatement of ObitUVCalSelect (checker
    only, outside the timing)."""
    from obit_amd import _lib
    from obit_amd.stokes import stokes_i_device
    lout = uv.nrparm + 3 * uv.nif * uv.nchan
    nin, nout = uv.buffer.nbytes, uv.nrow * lout * 4
    d_in, d_out = C.c_void_p(), C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d_in), nin), "malloc vis")
    try:
        _lib.check(lib.ObitHIPMalloc(C.byref(d_out), nout), "malloc I")
        try:
            _lib.check(lib.ObitHIPMemcpyH2D(d_in, _lib.ptr(uv.buffer), nin), "upload vis")
            stokes_i_device(lib, d_in.value, d_out.value, uv)  # warm
            lib.ObitHIPDeviceSynchronize()
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, None)
            for _ in range(reps):
                stokes_i_device(lib, d_in.value, d_out.value, uv)
            lib.ObitCUDAEventRecord(e1, None)
            lib.ObitHIPDeviceSynchronize()
            ms = lib.ObitHIPEventElapsedMs(e0, e1) / reps
            rows = np.sort(np.random.default_rng(7).choice(uv.nrow, min(check_rows, uv.nrow), replace=False))
            got = np.empty((rows.size, lout), np.float32)
            for i, r in enumerate(rows):
                _lib.check(lib.ObitHIPMemcpyD2H(got[i].ctypes.data, d_out.value + int(r) * lout * 4, lout * 4), "D2H")
        finally:
            lib.ObitHIPFree(d_out)
    finally:
        lib.ObitHIPFree(d_in)
    exact = None
    if check:
        from oracle import oracle as orc
        ref = orc.stokes_i(uv.buffer[rows], uv.nrparm, uv.nif, uv.nchan, uv.nstok, uv.incs, uv.incf,
                           uv.incif, int(round(uv.crval_stokes)))
        exact = bool(np.array_equal(got, ref, equal_nan=True))
    # algorithmic bytes: the record's random parameters and first two
    # correlations read, the Stokes I record written
    nbytes = uv.nrow * 4 * (2 * uv.nrparm + uv.nif * uv.nchan * (6 + 3))
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"stokes I: {uv.nrow} rows in {ms:.2f} ms = {gbs:.0f} GB/s; sample of {rows.size} rows bit-identical "
        f"to the oracle: {exact}")
    return {"workload": "workload 2 RR/LL rows -> Stokes I records", "rows": uv.nrow, "ms": ms,
            "rows_chan_per_s": uv.nrow * uv.nif * uv.nchan / (ms * 1e-3),
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": nbytes},
            "sample_rows_bit_identical": exact, "sample_rows": int(rows.size)}


def bench_grid(lib, steps):
    from obit_amd import synth, _lib
    from obit_amd.gridder import GPUGrid, plane_of_channel
    uv, _, facets = synth.workload(4)
    nci = uv.nchan * uv.nif
    b, e = synth.coarse_bins(nci, 8)
    poc = plane_of_channel(nci, b, e, cpu_parity=False)
    dptr = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(dptr), uv.buffer.nbytes), "malloc vis")
    res = {}
    try:
        _lib.check(lib.ObitHIPMemcpyH2D(dptr, _lib.ptr(uv.buffer), uv.buffer.nbytes), "upload vis")
        gg = GPUGrid(uv, facets, nplane=8, plane_of_chan=poc)
        try:
            st = gg.stream()
            gg.grid_device(dptr.value, uv.nrow)  # warm
            lib.ObitHIPDeviceSynchronize()
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, st)
            for _ in range(steps):
                gg.grid_device(dptr.value, uv.nrow)
            lib.ObitCUDAEventRecord(e1, st)
            lib.ObitHIPDeviceSynchronize()
            ms = lib.ObitHIPEventElapsedMs(e0, e1) / steps
            # a major cycle's later passes: same uvw, the binning pass reused
            lib.ObitCUDAEventRecord(e0, st)
            for _ in range(steps):
                gg.grid_device_reuse(dptr.value, uv.nrow)
            lib.ObitCUDAEventRecord(e1, st)
            lib.ObitHIPDeviceSynchronize()
            res["ms_per_pass_lists_reused"] = lib.ObitHIPEventElapsedMs(e0, e1) / steps
            log(f"gridding with the track lists reused: {res['ms_per_pass_lists_reused']:.2f} ms/pass")
            st8 = (C.c_longlong * 8)()
            _lib.check(lib.ObitHIPGridListStats(gg.c, 0, st8), "list stats")
            res["lists"] = dict(zip(("lists", "non_empty", "entries", "items", "longest", "split",
                                     "tile_samples", "cell_moves"), list(st8)))
            log(f"gridding lists: {res['lists']}")
            def regrid():
                lib.ObitCUDAGridInit(gg.c, 0)
                gg.grid_device(dptr.value, uv.nrow)
                lib.ObitHIPDeviceSynchronize()
            res["fft2image"] = bench_fft2image(lib, gg, facets, regrid)
        finally:
            gg.close()
    finally:
        lib.ObitHIPFree(dptr)
    nf = len(facets)
    samples = uv.nrow * nci * nf
    flop = 245.0 * samples  # 49 weight products + 98 complex-MAC FMAs per splat
    lrec_bytes = uv.lrec * 4
    grid_bytes = nf * 8 * facets[0].ny * (facets[0].nx // 2 + 1) * 8
    stream_bytes = nf * uv.nrow * lrec_bytes + grid_bytes
    res.update({
        "workload": "40 facets x 8 coarse bins of 2048^2, 210,600 rows x 256 ch Stokes I",
        "ms_per_pass": ms,
        "samples_per_s": samples / (ms * 1e-3),
        "rows_chan_per_s": uv.nrow * nci / (ms * 1e-3),
        "per_facet_stream_GBps": stream_bytes / (ms * 1e-3) / 1e9,
        "per_facet_stream_frac_hbm": stream_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm", "achieved": stream_bytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": stream_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": read_traffic("grid_traffic.json"),
            "bytes_model": "per-facet stream: facets x rows x record bytes + grids",
            "valu_TFLOPs": flop / (ms * 1e-3) / 1e12,
            "valu_frac": flop / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
        },
    })
    log(f"gridding: {ms:.2f} ms/pass, {res['samples_per_s'] / 1e9:.2f} Gsamples/s, "
        f"{res['per_facet_stream_GBps']:.0f} GB/s per-facet-stream model")
    return res, (uv, facets, b, e)


def cpu_grid_baseline(seconds):
    """Obit's threaded CPU gridder (restated in C: fast_prep_grid + 7x7
    scatter, rows split over threads into nGpI private grids, merged) on
    whole facets of the wideband workload, all 256 channels into one
    plane, facet after facet until about `seconds` have passed."""
    from obit_amd import synth
    from oracle import oracle as orc
    cores = _cores()
    uv, _, facets = synth.workload(4)
    ngpi = min(cores, 16)
    t0 = time.perf_counter()
    done = 0
    for f in facets:
        orc.grid_obit_threaded(uv, f, nGpI=ngpi, nthreads=cores, cpu_build=True)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    n = done * uv.nrow * uv.nchan * uv.nif
    log(f"cpu gridding: {done} facets x {uv.nrow} rows x 256 ch in {dt:.1f}s on {cores} threads "
        f"-> {n / dt / 1e9:.3f} Gsamples/s")
    return {"value": n / dt, "unit": "samples/s", "cores": cores, "kind": "port",
            "sample": f"{done} of the 40 facets, every row and channel ({n:.3g} samples, {dt:.1f} s)",
            "build": orc.CPU_BUILD_FLAGS}


def _cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 64))


def cpu_baseline(uv, sky, seconds):
    """Obit's threaded CPU DFT (restated in C: FP32 phase, SSE Cephes
    sincos, FP64 sums, ObitSkyModelFTDFT row split) on a row sample sized
    for about `seconds` of CPU time, on this host's cores."""
    from oracle import oracle as orc
    cores = _cores()
    rng = np.random.default_rng(9)
    n = 64
    per_comp = uv.nif * uv.nchan * sky.comps.shape[0]
    while True:
        rows = np.sort(rng.choice(uv.nrow, n, replace=False))
        sub = uv.subset(rows)
        t0 = time.perf_counter()
        orc.dft_obit(sub, sky, op="sub", nthreads=cores, cpu_build=True)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * seconds or n >= uv.nrow:
            break
        n = min(uv.nrow, max(n * 2, int(n * seconds / max(dt, 1e-3))))
        n = max(n, 1000)  # the reference threads only from 1000 rows up
    rate = n * per_comp / dt
    log(f"cpu baseline: {n} rows in {dt:.1f}s on {cores} threads -> {rate / 1e9:.3f} Gevals/s")
    return {"value": rate, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": f"{n} seeded rows of the 907,200-row workload, all 256 channels and "
                      f"10,000 components ({n * per_comp:.3g} evals, {dt:.1f} s)",
            "build": orc.CPU_BUILD_FLAGS}


def copy_calibration(lib, mib):
    """Three launches of the library's 16-byte copy kernel over `mib` MiB
    (reads and writes exactly 2 x mib MiB each): a profile of this run
    gives the byte counters' reading for a known amount of traffic."""
    from obit_amd import _lib
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(a), n), "malloc")
    _lib.check(lib.ObitHIPMalloc(C.byref(b), n), "malloc")
    try:
        for _ in range(3):
            _lib.check(lib.ObitHIPCopyKernel(b, a, n, None), "copy")
        _lib.check(lib.ObitHIPDeviceSynchronize(), "sync")
        log(f"copy kernel: 3 x {mib} MiB")
    finally:
        lib.ObitHIPFree(a)
        lib.ObitHIPFree(b)


def read_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel factors)."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def _virtual(nrow, lrec):
    """A read-only (nrow, lrec) float32 array of one element: the host-side
    descriptor of rows that live only on the device."""
    return np.lib.stride_tricks.as_strided(np.zeros(1, np.float32), shape=(nrow, lrec), strides=(0, 0),
                                           writeable=False)


def bench_workload5(lib, sample_rows=32):
    """Workload 5 at full size: 2016 baselines x 1,800 integrations x 1,024
    channels (4 IF) RR/LL, 89 GB resident; one batched subtract of the
    40-facet spectral model, Stokes I records, then one regrid into 40
    facets x 8 coarse bins of 2048^2.  Visibilities are N(0,1) noise filled
    on the device; a seeded row sample is kept on the host and checked
    against the FP64 oracle (model and gridding) after the timed steps."""
    from obit_amd import synth, _lib
    from obit_amd.gridder import GPUGrid, plane_of_channel
    from obit_amd.skymodel import SkyModelDFT
    from obit_amd.stokes import stokes_i_device
    rng = np.random.default_rng(synth.SEED0 + 5)
    enu = synth.core_arm_array(rng)
    f0 = 1.28e9
    freqArr = synth.freq_table(f0, 4, 256, 1e6)
    hdr = synth.uvw_tracks(enu, 34.08, 30.0, 1800, 8.0, f0).astype(np.float32)
    _, sky, facets = synth.workload(5, rows=24)  # the tests' 40-facet model
    nrow, nif, nchan, nstok = hdr.shape[0], 4, 256, 2
    lrec = 5 + 3 * nif * nchan * nstok
    lout = 5 + 3 * nif * nchan
    uv = synth.UVData(buffer=_virtual(nrow, lrec), nchan=nchan, nif=nif, nstok=nstok, freq=f0,
                      freqArr=freqArr, crval_stokes=-1.0)
    log(f"workload 5: {nrow} rows x {nif * nchan} ch x {nstok} Stokes ({nrow * lrec * 4 / 1e9:.1f} GB), "
        f"{sky.comps.shape[0]} spectral components in 40 facets")
    d_vis, d_i = C.c_void_p(), C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d_vis), nrow * lrec * 4), "malloc vis")
    _lib.check(lib.ObitHIPMalloc(C.byref(d_i), nrow * lout * 4), "malloc Stokes I")
    rows = np.sort(np.random.default_rng(9).choice(nrow, sample_rows, replace=False))
    before = np.empty((sample_rows, lrec), np.float32)
    try:
        # N(0,1) visibilities, unit weights, generated and uploaded in chunks
        step = 131072
        chunk = np.empty((step, lrec), np.float32)
        for b in range(0, nrow, step):
            e = min(nrow, b + step)
            c = chunk[:e - b]
            c[:, :5] = hdr[b:e]
            v = c[:, 5:].reshape(e - b, -1, 3)
            v[:, :, :2] = rng.standard_normal((e - b, v.shape[1], 2), dtype=np.float32)
            v[:, :, 2] = 1.0
            sel = (rows >= b) & (rows < e)
            before[sel] = c[rows[sel] - b]
            _lib.check(lib.ObitHIPMemcpyH2D(C.c_void_p(d_vis.value + b * lrec * 4), c.ctypes.data,
                                            c.nbytes), "upload")
            if (b // step) % 7 == 0:
                log(f"workload 5: uploaded {e} rows")
        res = {"workload": "workload 5: 2016 baselines x 1,800 integrations x 1,024 ch RR/LL, 40 facets, "
                           "2-term spectral model; subtract, Stokes I, regrid 40 x 8 x 2048^2",
               "rows": nrow, "channels": nif * nchan, "components": int(sky.comps.shape[0])}
        with SkyModelDFT(uv, sky, op="sub", nVisPIO=1024) as sm:
            sm.attach_vis(d_vis.value, nrow)
            res["dft_variant"] = sm.autotune(sample_rows=8192, reps=1)
            log(f"workload 5: variant {res['dft_variant']}; one launch over every row follows")
            st = sm.stream()
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, st)
            sm.calc_resident()
            lib.ObitCUDAEventRecord(e1, st)
            lib.ObitHIPDeviceSynchronize()
            ms = lib.ObitHIPEventElapsedMs(e0, e1)
        evals = nrow * nif * nchan * sky.comps.shape[0]
        res["dft_ms"] = ms
        res["dft_evals_per_s"] = evals / (ms * 1e-3)
        log(f"workload 5: subtract {evals:.3e} evaluations in {ms / 1e3:.1f} s = {evals / (ms * 1e-3):.3e}/s")
        after = np.empty_like(before)
        for i, r in enumerate(rows):
            _lib.check(lib.ObitHIPMemcpyD2H(after[i].ctypes.data, C.c_void_p(d_vis.value + int(r) * lrec * 4),
                                            lrec * 4), "D2H")
        e0, e1 = _events(lib)
        lib.ObitCUDAEventRecord(e0, None)
        stokes_i_device(lib, d_vis.value, d_i.value, uv)
        lib.ObitCUDAEventRecord(e1, None)
        lib.ObitHIPDeviceSynchronize()
        res["stokes_i_ms"] = lib.ObitHIPEventElapsedMs(e0, e1)
        lib.ObitHIPFree(d_vis)
        d_vis = C.c_void_p()
        uvI = synth.UVData(buffer=_virtual(nrow, lout), nchan=nchan, nif=nif, nstok=1, freq=f0,
                           freqArr=freqArr, crval_stokes=1.0)
        b, e = synth.coarse_bins(nif * nchan, 8)
        poc = plane_of_channel(nif * nchan, b, e, cpu_parity=False)
        with GPUGrid(uvI, facets, nplane=8, plane_of_chan=poc) as gg:
            st = gg.stream()
            e0, e1 = _events(lib)
            lib.ObitCUDAEventRecord(e0, st)
            gg.grid_device(d_i.value, nrow)
            lib.ObitCUDAEventRecord(e1, st)
            lib.ObitHIPDeviceSynchronize()
            gms = lib.ObitHIPEventElapsedMs(e0, e1)
            st8 = (C.c_longlong * 8)()
            _lib.check(lib.ObitHIPGridListStats(gg.c, 0, st8), "list stats")
            # the next major cycle: same uvw, the track lists reused
            lib.ObitCUDAEventRecord(e0, st)
            gg.grid_device_reuse(d_i.value, nrow)
            lib.ObitCUDAEventRecord(e1, st)
            lib.ObitHIPDeviceSynchronize()
            res["grid_ms_lists_reused"] = lib.ObitHIPEventElapsedMs(e0, e1)
    finally:
        if d_vis.value:
            lib.ObitHIPFree(d_vis)
        lib.ObitHIPFree(d_i)
    samples = nrow * nif * nchan * len(facets)
    res["grid_ms"] = gms
    res["grid_samples_per_s"] = samples / (gms * 1e-3)
    res["grid_lists"] = dict(zip(("lists", "non_empty", "entries", "items", "longest", "split",
                                  "tile_samples", "cell_moves"), list(st8)))
    log(f"workload 5: Stokes I {res['stokes_i_ms']:.1f} ms; regrid {samples:.3e} samples in {gms:.0f} ms "
        f"({res['grid_ms_lists_reused']:.0f} ms with the track lists reused)")
    # the checker: the sampled rows' model (before - after) against the
    # oracle's FP64 DFT, and their Stokes I records gridded on the device
    # against the oracle's FP64 gridder
    from oracle import oracle as orc
    sub = synth.UVData(buffer=np.ascontiguousarray(before), nchan=nchan, nif=nif, nstok=nstok, freq=f0,
                       freqArr=freqArr, crval_stokes=-1.0)
    ref = orc.dft_model_fp64(sub, sky, nthreads=_cores()).reshape(sample_rows, nif * nchan)
    d0 = before[:, 5:].reshape(sample_rows, nif * nchan, nstok, 3).astype(np.float64)
    d1 = after[:, 5:].reshape(sample_rows, nif * nchan, nstok, 3).astype(np.float64)
    got = (d0[..., 0, 0] - d1[..., 0, 0]) + 1j * (d0[..., 0, 1] - d1[..., 0, 1])
    peak = np.abs(ref).max()
    res["dft_err_max_vs_fp64_rel_peak"] = float(np.abs(got - ref).max() / peak)
    res["dft_err_rms_vs_fp64_rel_peak"] = float(np.sqrt(np.mean(np.abs(got - ref) ** 2)) / peak)
    res["check_rows"] = sample_rows
    log(f"workload 5: sampled model max |dV|/peak {res['dft_err_max_vs_fp64_rel_peak']:.2e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--sample-rows", type=int, default=4096,
                    help="rows of the FP64 comparator / accuracy sample")
    ap.add_argument("--variant", type=str, default=None,
                    help="cpt,tile,wg_per_cu: fixed DFT variant instead of the autotuner")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 comparator")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run the copy kernel over this many MiB three times "
                         "(calibrates byte counters)")
    ap.add_argument("--grid-steps", type=int, default=3)
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid-only", action="store_true", help="time the gridding workload alone")
    ap.add_argument("--beam-only", action="store_true", help="time the Jones-corrected DFT alone")
    ap.add_argument("--pcie-only", action="store_true", help="time the per-buffer DFT path alone")
    ap.add_argument("--workload5", action="store_true",
                    help="run workload 5 at full size (89 GB resident; about two minutes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the DFT shards by rows and the gridder by facets with no exchange)", file=sys.stderr)
        sys.exit(2)
    from obit_amd import _lib
    lib = _lib.load()
    name = C.create_string_buffer(64)
    ncu, clk, mem = C.c_int(), C.c_int(), C.c_longlong()
    _lib.check(lib.ObitHIPDeviceInfo(0, name, 64, C.byref(ncu), C.byref(clk), C.byref(mem)), "device")
    log(f"device {name.value.decode()} {ncu.value} CUs {clk.value / 1e3:.0f} MHz {mem.value / 2**30:.0f} GiB")

    if args.copy_mib:
        copy_calibration(lib, args.copy_mib)
    if args.workload5:
        print(json.dumps({"workload5": bench_workload5(lib)}), flush=True)
        return
    if args.beam_only:
        from obit_amd import synth
        uv, sky, _ = synth.workload(2)
        print(json.dumps({"beam_dft": bench_beam(lib, uv, sky, check=not args.no_cpu),
                          "beam_dft_4corr": bench_beam4(lib, uv, sky)}), flush=True)
        return
    if args.pcie_only:
        from obit_amd import synth
        uv, sky, _ = synth.workload(2)
        vt = tuple(int(x) for x in args.variant.split(",")) if args.variant else None
        print(json.dumps({"pcie": bench_pcie(lib, uv, sky, vt)}), flush=True)
        return
    if args.grid_only:
        grid, _ = bench_grid(lib, args.grid_steps)
        print(json.dumps({"gridding": grid}), flush=True)
        return
    variant = tuple(int(x) for x in args.variant.split(",")) if args.variant else None
    dft, uv, sky = bench_dft(lib, args.steps, args.warmup, args.sample_rows, variant,
                             accuracy=not args.no_accuracy)
    vt = tuple(dft["variant"][k] for k in ("cpt", "tile", "wg_per_cu"))
    pcie = None if args.no_accuracy else bench_pcie(lib, uv, sky, vt)
    stokes = None if args.no_grid else bench_stokes(lib, uv, check=not args.no_cpu)
    batched = None if args.no_grid else bench_batched(lib)
    beam = None if args.no_grid else bench_beam(lib, uv, sky, check=not args.no_cpu)
    grid, gwork = (None, None) if args.no_grid else bench_grid(lib, args.grid_steps)
    cpu = None if args.no_cpu else cpu_baseline(uv, sky, args.cpu_seconds)
    if grid is not None and not args.no_cpu:
        # the CPU leg: Obit's threaded gridder timed, and the oracle's FP64
        # gridder as the checker of a gridded row sample
        grid["cpu_baseline"] = cpu_grid_baseline(args.cpu_seconds / 2)
        grid["accuracy"] = grid_accuracy(*gwork)
    if not args.no_accuracy:
        dft_extra = {
            "err_max_vs_fp64_rel_peak": dft["err_max_rel_peak"],
            "err_rms_vs_fp64_rel_peak": dft["err_rms_rel_peak"],
            "fp64_gpu_evals_per_s": dft["fp64_evals_per_s"],
            "split_vs_fp64_gpu": dft["speedup_vs_fp64_gpu"],
            "accuracy_rows": args.sample_rows,
        }
    else:
        dft_extra = {}

    achieved = FLOP_PER_EVAL * dft["evals_per_step"] / (dft["kernel_ms"] * 1e-3) / 1e12
    mfma = dft["variant"]["cpt"] == 64
    line = {
        "metric": "vis x comp DFT evals/s (split precision, max|dV| vs FP64 stated), 1 MI355X",
        "value": dft["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dft["wall_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded array geometry, N(0,1) visibilities, power-law point model)",
        "config": {
            "workload": "64 antennas (2016 baselines) x 450 integrations, 4 IF x 64 ch, RR/LL, "
                        "10,000 power-law points within 0.7 deg; DFT subtract",
            "rows": dft["rows"], "channels": dft["channels"], "components": dft["components"],
            "precision": ("FP64 phases + range reduction, FP32 sincos and phasor powers, operands split "
                          "into FP16 hi + lo parts (hi*hi + hi*lo + lo*hi on the matrix cores), FP32 sums "
                          "over 32-component tiles folded into FP64" if mfma else
                          "FP64 phase + range reduction, FP32 sincos/recurrence/sums, FP64 tile folds"),
            "kernel": "k_dft_mfma (v_mfma_f32_16x16x32_f16, FP16 hi/lo parts)" if mfma else "k_dft_split (VALU)",
            "variant": dft["variant"],
            **dft_extra,
            **(pcie or {}),
            "kernel_ms": dft["kernel_ms"],
        },
        "roofline": {
            "bound": "mfma" if mfma else "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": read_traffic("dft_mfma_traffic.json" if mfma else "dft_traffic.json"),
        },
        "cpu_baseline": cpu,
        "gridding": grid,
        "stokes_i": stokes,
        "batched_multifacet": batched,
        "beam_dft": beam,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
