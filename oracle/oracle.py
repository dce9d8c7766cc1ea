"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
as the checker and by bench.py for the CPU-baseline timing.  The product
package (obit_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source built as Obit's stock "-O3 -msse" build, with the Cephes
# sincos/exp four lanes at a time: what bench.py times as the CPU baseline
LIB_CPU = os.path.join(HERE, "liboracle_cpu.so")
CPU_BUILD_FLAGS = "gcc -O3 -msse -msse2 -mno-fma, Cephes sincos/exp 4 lanes per SSE instruction"


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class DftDesc(C.Structure):
    _fields_ = [
        ("nrow", C.c_int), ("lrec", C.c_int), ("nrparm", C.c_int),
        ("ilocu", C.c_int), ("ilocv", C.c_int), ("ilocw", C.c_int),
        ("nchan", C.c_int), ("nif", C.c_int), ("nstok", C.c_int),
        ("incs", C.c_int), ("incf", C.c_int), ("incif", C.c_int),
        ("kincf", C.c_int), ("kincif", C.c_int),
        ("fscale", C.POINTER(C.c_float)), ("freqArr", C.POINTER(C.c_double)),
        ("startChan", C.c_int), ("numChan", C.c_int), ("startIF", C.c_int), ("numIF", C.c_int),
        ("startPoln", C.c_int), ("numPoln", C.c_int),
        ("modType", C.c_int), ("nSpecTerm", C.c_int),
        ("lcomp", C.c_int), ("ncomp", C.c_int), ("comps", C.POINTER(C.c_float)),
        ("smRefFreq", C.c_double),
        ("isMF", C.c_int), ("specIndex", C.POINTER(C.c_int)), ("specFreq", C.POINTER(C.c_double)),
        ("doDivide", C.c_int), ("doReplace", C.c_int), ("doFlip", C.c_int),
        ("stokFactor", C.c_float),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("nrow", C.c_int), ("lrec", C.c_int), ("nrparm", C.c_int), ("nchanIF", C.c_int),
        ("freqArr", C.POINTER(C.c_float)),
        ("convWidth", C.c_int), ("convNperCell", C.c_int),
        ("convgfn", C.POINTER(C.c_float)),
        ("nx", C.c_int), ("ny", C.c_int),
        ("uscale", C.c_float), ("vscale", C.c_float),
        ("shift", C.c_float * 3), ("rot", C.c_float * 9),
        ("guardu", C.c_float), ("guardv", C.c_float),
        ("maxBL", C.c_float), ("minBL", C.c_float), ("bmTaper", C.c_float),
        ("isBeam", C.c_int),
        ("bChan", C.c_int), ("eChan", C.c_int),
        ("sigma1", C.POINTER(C.c_float)), ("sigma2", C.POINTER(C.c_float)),
        ("sigma3", C.POINTER(C.c_float)),
    ]


_lib = None
_lib_cpu = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _declare(C.CDLL(LIB))
    return _lib


def load_cpu():
    """The CPU-baseline build (liboracle_cpu.so): same functions, same
    results bit for bit, built and vectorised as Obit's stock build."""
    global _lib_cpu
    if _lib_cpu is None:
        if not os.path.exists(LIB_CPU):
            build()
        _lib_cpu = _declare(C.CDLL(LIB_CPU))
    return _lib_cpu


def _declare(lib):
    if True:
        P = C.c_void_p
        lib.orc_sincos_cephes.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.orc_exp_cephes.argtypes = [C.c_float]
        lib.orc_exp_cephes.restype = C.c_float
        lib.orc_sincos_vec.argtypes = [C.c_int, P, P, P]
        lib.orc_exp_vec.argtypes = [C.c_int, P, P]
        lib.orc_dft_obit.argtypes = [C.POINTER(DftDesc), P, C.c_int, C.c_int]
        lib.orc_dft_obit_threaded.argtypes = [C.POINTER(DftDesc), P, C.c_int]
        lib.orc_dft_model_fp64.argtypes = [C.POINTER(DftDesc), P, P, P, C.c_int, C.c_int, C.c_int]
        lib.orc_convfunc_sph.argtypes = [P, P]
        lib.orc_sphfn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float]
        lib.orc_sphfn.restype = C.c_float
        lib.orc_grid_prep.argtypes = [C.POINTER(GridDesc), P, C.c_int, C.c_int, P, P, P, P, P]
        lib.orc_grid_obit.argtypes = [C.POINTER(GridDesc), P, C.c_int, C.c_int, P]
        lib.orc_grid_fp64.argtypes = [C.POINTER(GridDesc), P, C.c_int, C.c_int, P]
        lib.orc_grid_flip.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.orc_grid_flip_fp64.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.orc_grid_merge.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        lib.orc_grid_obit_threaded.argtypes = [C.POINTER(GridDesc), P, C.c_int, C.c_int, P]
        lib.orc_check_tap_fraction.argtypes = [C.c_float, C.c_int]
        lib.orc_check_tap_fraction.restype = C.c_long
    return lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ math

def sincos(x: float):
    s, c = C.c_float(), C.c_float()
    load().orc_sincos_cephes(C.c_float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def sincos_vec(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    load().orc_sincos_vec(len(x), _p(x), _p(s), _p(c))
    return s, c


def exp_vec(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    load().orc_exp_vec(len(x), _p(x), _p(y))
    return y


# ------------------------------------------------------------------ DFT

class _Keep(list):
    pass


def dft_desc(uv, sky, op="sub", stokFactor=1.0, doFlip=False, startChan=0, numChan=None,
             startIF=0, numIF=None, startPoln=0, numPoln=None, nrow=None):
    keep = _Keep()
    d = DftDesc()
    d.nrow = uv.nrow if nrow is None else nrow
    d.lrec, d.nrparm = uv.lrec, uv.nrparm
    d.ilocu, d.ilocv, d.ilocw = uv.ilocu, uv.ilocv, uv.ilocw
    d.nchan, d.nif, d.nstok = uv.nchan, uv.nif, uv.nstok
    d.incs, d.incf, d.incif = uv.incs, uv.incf, uv.incif
    d.kincf, d.kincif = 1, uv.nchan
    fs = np.ascontiguousarray(uv.fscale, np.float32)
    fa = np.ascontiguousarray(uv.freqArr, np.float64)
    keep += [fs, fa]
    d.fscale = fs.ctypes.data_as(C.POINTER(C.c_float))
    d.freqArr = fa.ctypes.data_as(C.POINTER(C.c_double))
    d.startChan = startChan
    d.numChan = uv.nchan - startChan if numChan is None else numChan
    d.startIF = startIF
    d.numIF = uv.nif - startIF if numIF is None else numIF
    d.startPoln = startPoln
    d.numPoln = min(2, uv.nstok) if numPoln is None else numPoln
    d.modType = sky.modType
    d.nSpecTerm = sky.nSpecTerm
    comps = np.ascontiguousarray(sky.comps, np.float32)
    keep.append(comps)
    d.lcomp, d.ncomp = comps.shape[1], comps.shape[0]
    d.comps = comps.ctypes.data_as(C.POINTER(C.c_float))
    d.smRefFreq = sky.imgRefFreq if sky.imgRefFreq > 0 else uv.freq
    if sky.specFreq is not None:
        sf = np.ascontiguousarray(sky.specFreq, np.float64)
        si = np.ascontiguousarray(sky.specIndex, np.int32)
        keep += [sf, si]
        d.isMF = 1
        d.specFreq = sf.ctypes.data_as(C.POINTER(C.c_double))
        d.specIndex = si.ctypes.data_as(C.POINTER(C.c_int))
    d.doDivide = int(op == "div")
    d.doReplace = int(op == "repl")
    d.doFlip = int(doFlip)
    d.stokFactor = stokFactor
    return d, keep


def dft_obit(uv, sky, buffer=None, nthreads=1, cpu_build=False, **kw):
    """Obit's CPU DFT (restated) applied in place to a copy of uv.buffer;
    cpu_build runs the CPU-baseline build of the same code."""
    buf = np.array(uv.buffer if buffer is None else buffer, dtype=np.float32, copy=True, order="C")
    d, keep = dft_desc(uv, sky, nrow=buf.shape[0], **kw)
    (load_cpu() if cpu_build else load()).orc_dft_obit_threaded(C.byref(d), _p(buf), nthreads)
    return buf


def dft_model_fp64(uv, sky, rows=None, nthreads=8, freq_ratio=None, **kw):
    """FP64 model visibility [nrow, nif, nchan] (complex128).  freq_ratio
    (nu/nu_uvref per channel, FP64) replaces freqArr/freq for u, v, w; the
    spectral terms still use freqArr."""
    buf = uv.buffer if rows is None else np.ascontiguousarray(uv.buffer[rows])
    buf = np.ascontiguousarray(buf, dtype=np.float32)
    d, keep = dft_desc(uv, sky, nrow=buf.shape[0], **kw)
    fr = np.ascontiguousarray(uv.freqArr / uv.freq if freq_ratio is None else freq_ratio, np.float64)
    out = np.zeros((buf.shape[0], uv.nif, uv.nchan, 2), np.float64)
    load().orc_dft_model_fp64(C.byref(d), _p(buf), _p(fr), _p(out), 0, buf.shape[0], nthreads)
    return out[..., 0] + 1j * out[..., 1]


# ------------------------------------------------------------------ gridding

def convfunc():
    cf = np.zeros(708, np.float32)
    cg = np.zeros(708, np.float32)
    load().orc_convfunc_sph(_p(cf), _p(cg))
    return cf, cg


def grid_desc(uv, facet, bChan=0, eChan=None, guardband=0.4, blmax=0.0, blmin=0.0,
              bmTaper=0.0, isBeam=False, sigma=None):
    keep = _Keep()
    g = GridDesc()
    g.nrow, g.lrec, g.nrparm = uv.nrow, uv.lrec, uv.nrparm
    g.nchanIF = uv.nchan * uv.nif
    fs = np.ascontiguousarray(uv.fscale, np.float32)
    _, cg = convfunc()
    keep += [fs, cg]
    g.freqArr = fs.ctypes.data_as(C.POINTER(C.c_float))
    g.convWidth, g.convNperCell = 7, 100
    g.convgfn = cg.ctypes.data_as(C.POINTER(C.c_float))
    g.nx, g.ny = facet.nx, facet.ny
    # ObitUVGridSetup (src/ObitUVGrid.c:275-280): cdelt is FP32, the cell in
    # radians is rounded to FP32, the scales are formed in double then FP32
    dg = np.pi / 180.0
    cell = np.float32(dg * float(np.float32(facet.cell_deg)))
    g.uscale = np.float32(facet.nx * abs(float(cell)))
    g.vscale = np.float32(-facet.ny * abs(float(cell)))
    for k in range(3):
        g.shift[k] = 0.0 if isBeam else float(facet.shift[k])
    for k in range(9):
        g.rot[k] = float(facet.rot[k])
    # guardband*nx in float (ObitThreadGrid.c:328-329)
    g.guardu = np.float32(np.float32(guardband) * np.float32(facet.nx))
    g.guardv = np.float32(np.float32(guardband) * np.float32(facet.ny))
    g.maxBL = blmax if blmax > 0 else 1.0e12
    g.minBL = blmin
    g.bmTaper = bmTaper
    g.isBeam = int(isBeam)
    g.bChan = bChan
    g.eChan = g.nchanIF if eChan is None else eChan
    if sigma is not None:
        s = [np.ascontiguousarray(x, np.float32) for x in sigma]
        keep += s
        g.sigma1, g.sigma2, g.sigma3 = (x.ctypes.data_as(C.POINTER(C.c_float)) for x in s)
    return g, keep


def grid_prep(uv, facet, rows=None, **kw):
    buf = uv.buffer if rows is None else np.ascontiguousarray(uv.buffer[rows])
    g, keep = grid_desc(uv, facet, **kw)
    g.nrow = buf.shape[0]
    n = buf.shape[0] * g.nchanIF
    iu, iv, itu, itv = (np.zeros(n, np.int32) for _ in range(4))
    vis = np.zeros(2 * n, np.float32)
    load().orc_grid_prep(C.byref(g), _p(np.ascontiguousarray(buf, np.float32)), 0, buf.shape[0],
                         _p(iu), _p(iv), _p(itu), _p(itv), _p(vis))
    return iu, iv, itu, itv, vis


def grid_size(facet):
    return 2 * (1 + 3 + facet.nx // 2) * facet.ny


def grid_fp64(uv, facet, flip=True, **kw):
    g, keep = grid_desc(uv, facet, **kw)
    buf = np.ascontiguousarray(uv.buffer, np.float32)
    grid = np.zeros(grid_size(facet), np.float64)
    load().orc_grid_fp64(C.byref(g), _p(buf), 0, buf.shape[0], _p(grid))
    if flip:
        load().orc_grid_flip_fp64(_p(grid), facet.nx, facet.ny, 7)
    return grid


def grid_obit(uv, facet, flip=True, **kw):
    g, keep = grid_desc(uv, facet, **kw)
    buf = np.ascontiguousarray(uv.buffer, np.float32)
    grid = np.zeros(grid_size(facet), np.float32)
    load().orc_grid_obit(C.byref(g), _p(buf), 0, buf.shape[0], _p(grid))
    if flip:
        load().orc_grid_flip(_p(grid), facet.nx, facet.ny, 7)
    return grid


def grid_obit_threaded(uv, facet, nGpI=4, nthreads=4, cpu_build=False, **kw):
    """ObitThreadGrid decomposition; returns complex64 [ny, nx/2+1] merged.
    cpu_build runs the CPU-baseline build of the same code."""
    g, keep = grid_desc(uv, facet, **kw)
    buf = np.ascontiguousarray(uv.buffer, np.float32)
    out = np.zeros((facet.ny, facet.nx // 2 + 1, 2), np.float32)
    (load_cpu() if cpu_build else load()).orc_grid_obit_threaded(C.byref(g), _p(buf), nGpI, nthreads, _p(out))
    return out[..., 0] + 1j * out[..., 1]


def trim_centre(grid, facet, dtype=np.complex128):
    """Padded grid -> [ny, nx/2+1] with v halves swapped (ThreadMerge)."""
    ny, nx = facet.ny, facet.nx
    lrow = 2 * (1 + nx // 2 + 3)
    g = grid[:lrow * ny].reshape(ny, lrow // 2, 2)
    g = g[:, 3:3 + nx // 2 + 1, 0] + 1j * g[:, 3:3 + nx // 2 + 1, 1]
    return np.roll(g, ny // 2, axis=0).astype(dtype)


# ------------------------------------------------------------ FFT to image

def grid_corr_fn(n, icent, convfn, width=7, nper=100):
    """GridCorrFn (src/ObitUVGrid.c:1612-1662) by its definition, in FP64:
    the cosine transform of the tabulated convolving function,
    sum_i convfn[i] cos(2 pi (i + 1 - bias)(x + 1 - icent) / (nper n)),
    normalised to 1 at pixel n/2 and inverted.  (The reference evaluates
    the same sum with an FP32 phase-ramp recurrence.)"""
    size = width * nper + 1
    bias = width * (nper // 2) + 1
    i = np.arange(size, dtype=np.float64)
    x = np.arange(n, dtype=np.float64)
    ph = 2.0 * np.pi * np.outer(x + 1.0 - icent, i + 1.0 - bias) / (nper * n)
    s = (np.asarray(convfn[:size], np.float64)[None, :] * np.cos(ph)).sum(axis=1)
    s = s / s[n // 2]
    return 1.0 / s


def fft2image(planes, convfn, is_beam, beam_norm=None, corr=None):
    """ObitUVGridFFT2Im (src/ObitUVGrid.c:673-834) in FP64 with numpy's FFT
    standing in for FFTW: planes [nplane, ny, nx/2+1] complex (trimmed and
    centred as ThreadMerge leaves them) -> unnormalised C2R, quadrant swap
    (ObitFArray2DCenter), correction applied as ObitFArrayMulColRow does
    (pixel (x, y) times xCorr[y] / norm times yCorr[x]) and the beam's
    centre-value normalisation.  corr=(xCorr, yCorr) replaces the FP64
    correction functions.  Returns (images [nplane, ny, nx], norms)."""
    planes = np.asarray(planes)
    nplane, ny, nxo = planes.shape
    nx = 2 * (nxo - 1)
    if corr is None:
        xc = grid_corr_fn(nx, nx // 2 + 1, convfn)
        yc = grid_corr_fn(ny, ny // 2 + 1, convfn)
    else:
        xc, yc = (np.asarray(c, np.float64) for c in corr)
    img = np.fft.irfft2(planes, s=(ny, nx), axes=(1, 2)) * (nx * ny)
    img = np.roll(img, (ny // 2, nx // 2), axis=(1, 2))
    if is_beam:
        norms = img[:, ny // 2, nx // 2] * xc[nx // 2] * yc[ny // 2]
    else:
        norms = np.asarray(beam_norm, np.float64)
    out = img * (xc[None, :, None] / norms[:, None, None]) * yc[None, None, :]
    return out, norms


# ---------------------------------------------------------------- Stokes I
def stokes_i_setup(kstoke0, nstok, formal_i=False):
    """ObitUVCalSelectIInit (src/ObitUVCalSelect.c:697-760): None when the
    data cannot give Stokes I, else (second correlation's offset in Stokes
    steps, factors (f1, f2), both hands required)."""
    missing = (kstoke0 > 1) or (kstoke0 < -2)
    if kstoke0 <= -5:
        missing = kstoke0 < -6
    if formal_i and kstoke0 < 0:
        missing = missing or nstok < 2 or (kstoke0 != -1 and kstoke0 != -5)
    if missing:
        return None
    if kstoke0 > 0 or nstok == 1:
        return 0, (np.float32(1.0), np.float32(0.0)), False
    return 1, (np.float32(0.5), np.float32(0.5)), bool(formal_i)


def stokes_i(buffer, nrparm, nif, nchan, nstok, incs, incf, incif, kstoke0, formal_i=False):
    """ObitUVCalSelect (src/ObitUVCalSelect.c:540-695) for a Stokes 'I'
    selection of every channel and IF: rows of nrparm + 3 nif nchan
    float32 (random parameters copied).  FP32 in the CPU's order."""
    setup = stokes_i_setup(kstoke0, nstok, formal_i)
    if setup is None:
        raise ValueError("Stokes I not available")
    j2, (f1, f2), both = setup
    buf = np.asarray(buffer, np.float32)
    nrow = buf.shape[0]
    out = np.zeros((nrow, nrparm + 3 * nif * nchan), np.float32)
    out[:, :nrparm] = buf[:, :nrparm]
    half, zero, two = np.float32(0.5), np.float32(0.0), np.float32(2.0)
    with np.errstate(invalid="ignore", over="ignore"):
        for iif in range(nif):
            for ch in range(nchan):
                p = nrparm + iif * incif + ch * incf
                q = p + j2 * incs
                a, b = buf[:, p:p + 3], buf[:, q:q + 3]
                # MAX (0.0, w) * 0.5: the comparison lets a NaN weight through
                w1 = np.where(zero > a[:, 2], zero, a[:, 2]) * half
                w2 = np.where(zero > b[:, 2], zero, b[:, 2]) * half
                take = ~(both & ((w1 <= 0) | (w2 <= 0)))
                wt = np.where(take, w1 + w2, zero).astype(np.float32)
                live = take & (wt > 0)
                x1 = np.where(w1 > 0, f1, zero).astype(np.float32)
                x2 = np.where(w2 > 0, f2, zero).astype(np.float32)
                vr = (x1 * a[:, 0] + x2 * b[:, 0]).astype(np.float32)
                vi = (x1 * a[:, 1] + x2 * b[:, 1]).astype(np.float32)
                one = (w1 <= 0) | (w2 <= 0)
                vr = np.where(one, two * vr, vr)
                vi = np.where(one, two * vi, vi)
                o = nrparm + 3 * (iif * nchan + ch)
                out[:, o] = np.where(live, vr, zero)
                out[:, o + 1] = np.where(live, vi, zero)
                out[:, o + 2] = wt
    return out


# ------------------------------------------------------- Jones-corrected DFT

def beam_ants(buffer, ilocb=3, iloca1=-1, iloca2=-1):
    """0-rel antenna numbers of each row: BASELINE = 256 a1 + a2 (+ subarray
    fraction), truncated as GetAnts does (reference
    src/ObitCUDABeamModel.cu:82-98), or ANTENNA1/ANTENNA2 rounded."""
    if ilocb >= 0:
        cb = buffer[:, ilocb].astype(np.float32).astype(np.int64)   # float -> int truncation
        a1 = (cb / 256.0 + 0.001).astype(np.int64)
        a2 = ((cb - a1 * 256) + 0.001).astype(np.int64)
    else:
        a1 = (buffer[:, iloca1] + np.float32(0.5)).astype(np.int64)
        a2 = (buffer[:, iloca2] + np.float32(0.5)).astype(np.int64)
    return a1 - 1, a2 - 1


def beam_coherency(stok_type, par_ang):
    """2x2 coherency of a unit model in Stokes I, Q, U or V for linear feeds
    at parallactic angle par_ang, as Obit's CPU model sets it
    (src/ObitSkyModelVMBeamMF.c:2722-2747, TMS2 eq. 4.41 order XX, XY, YX, YY)."""
    c, s = np.cos(2.0 * par_ang), np.sin(2.0 * par_ang)
    if stok_type == 2:
        return np.array([[c, -s], [-s, -c]], np.complex128)
    if stok_type == 3:
        return np.array([[s, c], [c, -s]], np.complex128)
    if stok_type == 4:
        return np.array([[0, 1j], [-1j, 0]], np.complex128)
    return np.eye(2, dtype=np.complex128)


def beam_products(jones, ant_types_pairs, stok_type, par_ang):
    """J1 P J2^H per (pair, component, beam channel): beam1 * model *
    beam2^H (src/ObitSkyModelVMBeamMF.c:2752-2755).  jones: complex
    [ntype, ncomp, njchan, 2, 2]."""
    P = beam_coherency(stok_type, par_ang)
    out = {}
    for (t1, t2) in ant_types_pairs:
        J1, J2 = jones[t1], jones[t2]
        out[(t1, t2)] = J1 @ P @ np.conj(np.swapaxes(J2, -1, -2))
    return out


def beam_model_fp64(buffer, comps, freq_scale, ant_type, vis_jchan, jones, stok_type=1, par_ang=0.0,
                    ilocu=0, ilocv=1, ilocw=2, ilocb=3, iloca1=-1, iloca2=-1, stok_fact=(1, 1, 1, 1),
                    amps=None):
    """FP64 model [nrow, nif*nchan, 2, 2] (XX XY / YX YY) of the
    Jones-corrected DFT, and the mask of rows whose antennas are in the
    type table.  Phase u x + v y + w z scaled by nu/nu0 (radians, x in
    radians per wavelength); a Gaussian multiplies the amplitude by
    exp((nu/nu0)^2 (g0 u^2 + g1 v^2 + g2 uv)) (reference
    src/ObitCUDABeamModel.cu:197-203, 761-767).  freq_scale: nu/nu0 per
    (IF, channel), IF-major.  amps [ncomp, nif*nchan], when given, replaces
    the amplitude column per channel (spectral component types)."""
    buf = np.asarray(buffer, np.float32)
    comps = np.asarray(comps, np.float32).astype(np.float64)
    a1, a2 = beam_ants(buf, ilocb, iloca1, iloca2)
    nant = len(ant_type)
    live = (a1 >= 0) & (a1 < nant) & (a2 >= 0) & (a2 < nant)
    at = np.asarray(ant_type)
    t1 = np.where(live, at[np.clip(a1, 0, nant - 1)], 0)
    t2 = np.where(live, at[np.clip(a2, 0, nant - 1)], 0)
    pairs = sorted(set(zip(t1.tolist(), t2.tolist())))
    prods = beam_products(jones, pairs, stok_type, par_ang)
    u = buf[:, ilocu].astype(np.float64)
    v = buf[:, ilocv].astype(np.float64)
    w = buf[:, ilocw].astype(np.float64)
    psi = np.outer(u, comps[:, 1]) + np.outer(v, comps[:, 2]) + np.outer(w, comps[:, 3])
    gauss = comps.shape[1] >= 7 and np.any(comps[:, 4:7] != 0)
    G = (np.outer(u * u, comps[:, 4]) + np.outer(v * v, comps[:, 5]) + np.outer(u * v, comps[:, 6])) if gauss else None
    nrow, nf = buf.shape[0], len(freq_scale)
    out = np.zeros((nrow, nf, 2, 2), np.complex128)
    rows_of = {p: np.nonzero((t1 == p[0]) & (t2 == p[1]))[0] for p in pairs}
    for f in range(nf):
        fs = float(freq_scale[f])
        a0 = (np.asarray(amps, np.float32)[:, f].astype(np.float64) if amps is not None else comps[:, 0])
        amp = a0[None, :] * (np.exp(fs * fs * G) if gauss else 1.0)
        V = amp * np.exp(1j * fs * psi)                       # [nrow, ncomp]
        jc = int(vis_jchan[f])
        for p, r in rows_of.items():
            if len(r):
                out[r, f] = np.einsum("rk,kab->rab", V[r], prods[p][:, jc])
    sf = np.asarray(stok_fact, np.float64).reshape(2, 2)
    return out * sf[None, None], live


def beam_apply(buffer, model, live, nrparm, nif, nchan, incf, incif, ncorr, op="sub", chan_sel=None,
               if_sel=None):
    """Apply a [nrow, nif*nchan, 2, 2] model to a copy of the rows as
    Obit's CPU beam DFT does (src/ObitSkyModelVMBeamMF.c:3711-3800), in
    FP32: channels with every correlation flagged are skipped unless
    replacing; each correlation changes where its weight is positive (or
    always when replacing); divide acts on XX and YY only; cross hands only
    with 4 correlations; only channels and IFs inside the selection
    (reference ObitCUDABeamModel.cu:180-184)."""
    buf = np.array(buffer, np.float32, copy=True)
    cb, ce = chan_sel if chan_sel is not None else (0, nchan - 1)
    ib, ie = if_sel if if_sel is not None else (0, nif - 1)
    m = model.astype(np.complex64)
    repl = op == "repl"
    hands = [(0, 0, 0), (1, 1, 3)] + ([(0, 1, 6), (1, 0, 9)] if ncorr == 4 and op != "div" else [])
    for i in range(ib, ie + 1):
        for c in range(cb, ce + 1):
            f = i * nchan + c
            o = nrparm + i * incif + c * incf
            wts = [buf[:, o + 2], buf[:, o + 5]] + ([buf[:, o + 8], buf[:, o + 11]] if ncorr == 4 else [])
            bad = np.all(np.stack(wts) <= 0, axis=0)
            act = live & (~bad | repl)
            for (a, b, off) in hands:
                q = o + off
                mr = np.real(m[:, f, a, b]).astype(np.float32)
                mi = np.imag(m[:, f, a, b]).astype(np.float32)
                sel = act & ((buf[:, q + 2] > 0) | repl)
                if op == "div":
                    with np.errstate(divide="ignore", invalid="ignore"):
                        wm = mr * mr + mi * mi
                        dr, di = mr / wm, mi / wm
                        re, im = buf[:, q].copy(), buf[:, q + 1].copy()
                        nr = dr * re + di * im
                        ni = dr * im - di * re
                        nw = buf[:, q + 2] * np.sqrt(wm)
                    buf[sel, q], buf[sel, q + 1], buf[sel, q + 2] = nr[sel], ni[sel], nw[sel]
                elif repl:
                    buf[sel, q], buf[sel, q + 1] = mr[sel], mi[sel]
                else:
                    buf[sel, q] -= mr[sel]
                    buf[sel, q + 1] -= mi[sel]
    return buf


def beam_lagrange(p, n, hw):
    """Start pixel (0-rel) and the 2 hw + 1 Lagrange weights for 1-rel
    pixel p on an axis of n pixels (SetConvKernal, reference
    include/CUDABeamInterp.h:233-258), FP32 as the device computes them."""
    iwid = 2 * hw + 1
    p = np.float32(p)
    cen = int(p + np.float32(0.5)) - hw
    cen = max(1, min(cen, n - iwid + 1)) - 1
    xx = np.float32(p - np.float32(cen))
    denom = []
    for j in range(1, iwid + 1):
        prod = np.float32(1.0)
        for i in range(1, iwid + 1):
            if i != j:
                prod = np.float32(prod * np.float32(j - i))
        denom.append(np.float32(1.0) / prod)
    k = []
    for j in range(iwid):
        prod = denom[j]
        for i in range(iwid):
            if i != j:
                prod = np.float32(prod * np.float32(xx - np.float32(i + 1)))
        k.append(prod)
    return cen, np.array(k, np.float32)


def beam_interp(plane, px, py, hw, fblank):
    """Interpolated value of an image plane [ny, nx] at 1-rel (px, py)
    (CUDABeamInterpPixel, include/CUDABeamInterp.h:81-145): blank outside
    the image or when the unblanked weights sum to 0.9 or less."""
    ny, nx = plane.shape
    if px < 1.0 or px > nx or py < 1.0 or py > ny:
        return np.float32(fblank)
    xs, xk = beam_lagrange(px, nx, hw)
    ys, yk = beam_lagrange(py, ny, hw)
    w = np.outer(yk.astype(np.float64), xk.astype(np.float64))
    blk = plane[ys:ys + len(yk), xs:xs + len(xk)].astype(np.float64)
    ok = blk != np.float32(fblank)
    sw = float(np.sum(w[ok]))
    if sw > 0.90:
        return np.float32(np.sum(blk[ok] * w[ok]) / sw)
    return np.float32(fblank)


def beam_pixel(desc, x, y, rot):
    """1-rel pixel of position (x, y) deg in a beam image rotated by rot
    (CUDAPositionXYpix / CUDAPositionXYpixLM, include/CUDASkyGeom.h:90-312)
    for linear, -SIN, -TAN and -ARC geometry; None where the projection
    cannot place it."""
    crval, crpix, cdelt, proj = desc["crval"], desc["crpix"], desc["cdelt"], desc.get("proj", 0)
    dt = x - crval[0]
    if dt > 180.0:
        x -= 360.0
    if dt < -180.0:
        x += 360.0
    d0, d1 = x - crval[0], y - crval[1]
    if cdelt[0] == 0 or cdelt[1] == 0:
        return None
    if proj:
        d2r = 1.745329252e-2
        ra0, dec0, ra, dec = crval[0] * d2r, crval[1] * d2r, x * d2r, y * d2r
        l = math.sin(ra - ra0) * math.cos(dec)
        n = math.sin(dec) * math.sin(dec0) + math.cos(dec) * math.cos(dec0) * math.cos(ra - ra0)
        m = math.sin(dec) * math.cos(dec0) - math.cos(dec) * math.sin(dec0) * math.cos(ra - ra0)
        if proj == 1:
            if n < 0:
                return None
        elif proj == 2:
            if n <= 0:
                return None
            l, m = l / n, m / n
        else:
            c = math.acos(min(1.0, max(-1.0, n)))
            c = c / math.sin(c) if c != 0 else 1.0
            l, m = l * c, m * c
        d0, d1 = l / d2r, m / d2r
    rot = float(np.float32(rot))
    sr, cr = float(np.float32(math.sin(rot))), float(np.float32(math.cos(rot)))
    dz = d0 * cr + d1 * sr
    d1 = d1 * cr - d0 * sr
    return np.float32(dz / np.float32(cdelt[0]) + np.float32(crpix[0])), \
        np.float32(d1 / np.float32(cdelt[1]) + np.float32(crpix[1]))


def inv_beam(fact, freq, ang):
    """1/sqrt of the cos^2 reference beam gain (InvBeam,
    ObitCUDABeamModel.cu:66-71); 100 beyond 1.02 of the argument."""
    arg = np.float32(np.float32(fact) * np.float32(freq) * np.float32(ang))
    if arg > 1.02:
        return np.float32(100.0)
    div = np.float32(1.0 - 4.0 * float(arg) * float(arg))
    return np.float32(div / np.float32(math.cos(float(np.float32(math.pi * float(arg))))))


def beam_jones(descs, images, beam_freq, ref_size_fact, fblank, pos, par_ang):
    """Jones table [ntype, ncomp, nchan, 2, 2] from beam images
    [ntype][8][nchan, ny, nx] (GetJonesKernel, ObitCUDABeamModel.cu:1227-1308):
    images in the order P1 re, P1 im, P2 re, P2 im, X1 re, X1 im, X2 re,
    X2 im; pos [ncomp, 2] RA, Dec offsets (deg), placed at (-RA, Dec)."""
    ntype, ncomp, nch = len(descs), len(pos), len(beam_freq)
    J = np.zeros((ntype, ncomp, nch, 2, 2), np.complex128)
    J[..., 0, 0] = J[..., 1, 1] = 1.0
    for t in range(ntype):
        d, img = descs[t], np.asarray(images[t], np.float32)
        hw = d.get("hwidth", 2)
        for k in range(ncomp):
            x, y = -float(np.float32(pos[k][0])), float(np.float32(pos[k][1]))
            pix = beam_pixel(d, x, y, par_ang)
            if pix is None:
                continue
            ang = np.float32(math.sqrt(np.float32(x * x + y * y)))
            for c in range(nch):
                ipb = inv_beam(ref_size_fact, beam_freq[c], ang)
                val = lambda i: np.float32(ipb * beam_interp(img[i, c], pix[0], pix[1], hw, fblank))
                p1r, p2r = val(0), val(2)
                if ipb < 10.0 and abs(p1r) <= 1.3:
                    J[t, k, c] = [[p1r + 1j * val(1), val(4) + 1j * val(5)],
                                  [val(6) + 1j * val(7), p2r + 1j * val(3)]]
    return J
