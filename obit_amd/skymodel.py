"""DFT sky model on the GPU, through Obit's C entry points.

``SkyModelDFT`` mirrors how Obit's ObitSkyModel(MF) drives its GPU class
(src/ObitSkyModel.c:850-1026 SubUV, src/ObitSkyModelMF.c:884-969 InitMod /
InitModel, src/ObitGPUSkyModel.c:207-579): DFTInit with the uv descriptor,
DFTSetMod with the component table, then DFTCalc per uv buffer - or, for a
field whose visibilities stay resident in HBM, one CalcResident launch over
the whole data set per major cycle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .synth import UVData, SkyModel

OPS = {"sub": (0, 0), "div": (1, 0), "repl": (0, 1)}


def host_uvdesc(uv: UVData):
    """HostUVDesc for uv; returns (struct, keep-alive arrays)."""
    fa = np.ascontiguousarray(uv.freqArr, dtype=np.float64)
    fs = np.ascontiguousarray(uv.fscale, dtype=np.float32)
    d = _lib.HostUVDesc()
    d.nvis, d.lrec, d.nrparm = uv.nrow, uv.lrec, uv.nrparm
    d.ilocu, d.ilocv, d.ilocw = uv.ilocu, uv.ilocv, uv.ilocw
    d.iloct, d.ilocb, d.ilocsu = uv.iloct, uv.ilocb, -1
    d.nchan, d.nif, d.nstok = uv.nchan, uv.nif, uv.nstok
    d.freqBeforeIF, d.hasIF = 1, 1
    d.incs, d.incf, d.incif = uv.incs, uv.incf, uv.incif
    d.ncorr = uv.nchan * uv.nif * uv.nstok
    d.crvalStokes = uv.crval_stokes
    d.freq = uv.freq
    d.freqArr = fa.ctypes.data_as(C.POINTER(C.c_double))
    d.fscale = fs.ctypes.data_as(C.POINTER(C.c_float))
    return d, (fa, fs)


def load_comps_point(cc, im_crval, im_crota, uv_crval, uv_crota=0.0, factor=1.0) -> np.ndarray:
    """Component table of one 3-D image field from its CC rows
    [flux, deltaX, deltaY (deg)], as ObitSkyModelLoadComps builds it
    (ObitHostLoadCompsPoint): float32 [n, 4] rows [flux*factor, x, y, z]."""
    cc = np.ascontiguousarray(cc, dtype=np.float32)
    g = _lib.HostFieldGeom()
    g.imCrval[0], g.imCrval[1] = float(im_crval[0]), float(im_crval[1])
    g.imCrota = float(im_crota)
    g.do3D = 1
    g.uvCrval[0], g.uvCrval[1] = float(uv_crval[0]), float(uv_crval[1])
    g.uvCrota = float(uv_crota)
    out = np.zeros((cc.shape[0], 4), np.float32)
    n = _lib.load().ObitHostLoadCompsPoint(C.byref(g), cc.ctypes.data_as(C.POINTER(C.c_float)),
                                           cc.shape[0], cc.shape[1], float(factor),
                                           out.ctypes.data_as(C.POINTER(C.c_float)))
    if n != cc.shape[0]:
        raise _lib.HipError("ObitHostLoadCompsPoint failed")
    return out


def cc_merge(rows, have_z: bool = False, no_parms: int = 0, status=None, start: int = 1, end: int | None = None):
    """ObitTableCCUtilMergeSel on CC rows [Flux, DeltaX, DeltaY, (DeltaZ),
    parms[no_parms]] (ObitHostCCMergeSel): merged rows [flux, X, Y, (Z),
    spectral terms] by decreasing |flux|, and the first selected row's
    parms[0..3]."""
    lib = _lib.load()
    rows = np.ascontiguousarray(rows, dtype=np.float32)
    nrow = rows.shape[0]
    if rows.ndim != 2 or rows.shape[1] != 3 + int(have_z) + no_parms:
        raise ValueError("rows must be [nrow, 3 + haveZ + no_parms]")
    width = lib.ObitHostCCMergeWidth(int(have_z), no_parms)
    out = np.zeros((max(nrow, 1), width), np.float32)
    parms = np.zeros(4, np.float32)
    st = None if status is None else np.ascontiguousarray(status, dtype=np.int32)
    n = lib.ObitHostCCMergeSel(_lib.ptr(rows), nrow, int(have_z), no_parms,
                               None if st is None else _lib.ptr(st), start,
                               nrow if end is None else end, _lib.ptr(out), _lib.ptr(parms))
    if n < 0:
        raise ValueError("ObitHostCCMergeSel: bad input")
    return out[:n].copy(), parms


def load_comps(geom: dict, merged, parms, have_z: bool = False, nterms: int = 0, mod_type: int = 0,
               n_spec_term: int = 0, factor: float = 1.0, min_flux: float = 0.0, min_dft: float = 0.0,
               no_neg: bool = False) -> np.ndarray:
    """ObitSkyModelLoadComps for one field (ObitHostLoadComps) from merged
    CC rows (cc_merge); geom holds the field's and the uv data's geometry
    (ObitHostFieldGeom members: im_crval, im_crota, do3D, xPxOff, yPxOff,
    uv_crval, uv_crota, im_crpix, im_cdelt, im_naxis, xshift, yshift)."""
    lib = _lib.load()
    g = _lib.HostFieldGeom()
    g.imCrval[0], g.imCrval[1] = geom["im_crval"]
    g.imCrota = geom["im_crota"]
    g.do3D = int(geom["do3D"])
    g.xPxOff, g.yPxOff = geom["xPxOff"], geom["yPxOff"]
    g.uvCrval[0], g.uvCrval[1] = geom["uv_crval"]
    g.uvCrota = geom["uv_crota"]
    g.imCrpix[0], g.imCrpix[1] = geom["im_crpix"]
    g.imCdelt[0], g.imCdelt[1] = geom["im_cdelt"]
    g.imNaxis[0], g.imNaxis[1] = geom["im_naxis"]
    g.xshift, g.yshift = geom["xshift"], geom["yshift"]
    o = _lib.HostLoadOpts()
    o.modType, o.nSpecTerm, o.factor = mod_type, n_spec_term, factor
    o.minFlux, o.minDFT, o.noNeg = min_flux, min_dft, int(no_neg)
    merged = np.ascontiguousarray(merged, dtype=np.float32)
    parms = np.ascontiguousarray(parms, dtype=np.float32)
    nm = merged.shape[0]
    width = 4 + (3 if mod_type in (1, 11) else 0) + (n_spec_term if mod_type >= 10 else 0)
    out = np.zeros((max(nm, 1), width), np.float32)
    n = lib.ObitHostLoadComps(C.byref(g), _lib.ptr(merged), nm, int(have_z), nterms, _lib.ptr(parms),
                              C.byref(o), _lib.ptr(out), nm)
    if n < 0:
        raise ValueError("ObitHostLoadComps: bad input or a combination of model types Obit refuses")
    return out[:n].copy()


class SkyModelDFT:
    """A component model applied to visibilities by the HIP DFT.

    op: "sub" (data -= model), "div" (data /= model, weight *= |model|) or
    "repl" (data = model).  Channel/IF/Stokes selections are 1-relative, as
    Obit's startChannel/numberChannel etc.  stokFactor scales each Stokes
    after the first (Obit's stokFactor); doFlip multiplies the model by i.
    """

    def __init__(self, uv: UVData, sky: SkyModel, op: str = "sub", factor: float = 1.0,
                 stokFactor: float = 1.0, doFlip: bool = False, startChannel: int = 1,
                 numberChannel: int | None = None, startIF: int = 1, numberIF: int | None = None,
                 startPoln: int = 1, numberPoln: int | None = None, device: int = 0,
                 useFreq64: bool = True, nVisPIO: int = 10000):
        self.lib = _lib.load()
        self.uv = uv
        self.sky = sky
        dv, self._keep_uv = host_uvdesc(uv)
        self._desc = dv
        sm = _lib.HostSkyModel()
        sm.modType = sky.modType
        sm.nSpecTerm = sky.nSpecTerm
        sm.doDivide, sm.doReplace = OPS[op]
        sm.doFlip = int(doFlip)
        sm.factor, sm.stokFactor = factor, stokFactor
        sm.startChannel = startChannel
        sm.numberChannel = uv.nchan - startChannel + 1 if numberChannel is None else numberChannel
        sm.startIF = startIF
        sm.numberIF = uv.nif - startIF + 1 if numberIF is None else numberIF
        sm.startPoln = startPoln
        sm.numberPoln = min(2, uv.nstok) if numberPoln is None else numberPoln
        self._keep_sm = []
        if sky.specFreq is not None:
            sf = np.ascontiguousarray(sky.specFreq, dtype=np.float64)
            si = np.ascontiguousarray(sky.specIndex, dtype=np.int32)
            self._keep_sm += [sf, si]
            sm.isMF = 1
            sm.nSpec = len(sf)
            sm.specFreq = sf.ctypes.data_as(C.POINTER(C.c_double))
            sm.specIndex = si.ctypes.data_as(C.POINTER(C.c_int))
        sm.imgRefFreq = sky.imgRefFreq
        sm.cuda_device = device
        sm.useFreq64 = int(useFreq64)
        self._sm = sm
        self.g = self.lib.ObitHostDFTCreate()
        _lib.check(self.lib.ObitHostDFTInit(self.g, C.byref(dv), C.byref(sm), nVisPIO), "DFTInit")
        self.set_model(sky.comps)
        self._resident = None

    # -- model --
    def set_model(self, comps: np.ndarray) -> None:
        comps = np.ascontiguousarray(comps, dtype=np.float32)
        self._comps = comps
        _lib.check(self.lib.ObitHostDFTSetMod(self.g, C.byref(self._sm), _lib.ptr(comps),
                                              comps.shape[1], comps.shape[0]), "DFTSetMod")

    # -- per-buffer path (Obit's DFTCalc) --
    def calc(self, buffer: np.ndarray, times: np.ndarray | None = None) -> np.ndarray:
        """Apply the model in place to a host buffer, nVisPIO rows per call.
        times (float64[3]), when given, accumulates the wall time, the time
        inside ObitCUDASkyModelDFTCalc and the host copies' own time."""
        if buffer.dtype != np.float32 or not buffer.flags["C_CONTIGUOUS"]:
            raise ValueError("buffer must be C-contiguous float32")
        if times is None:
            _lib.check(self.lib.ObitHostDFTCalc(self.g, _lib.ptr(buffer), buffer.shape[0]), "DFTCalc")
        else:
            if times.dtype != np.float64 or times.shape != (3,):
                raise ValueError("times must be float64[3]")
            _lib.check(self.lib.ObitHostDFTCalcTimed(self.g, _lib.ptr(buffer), buffer.shape[0],
                                                     times.ctypes.data_as(C.POINTER(C.c_double))), "DFTCalc")
        return buffer

    # -- resident path --
    def set_vis(self, buffer: np.ndarray) -> None:
        buffer = np.ascontiguousarray(buffer, dtype=np.float32)
        _lib.check(self.lib.ObitHIPSkyModelDFTSetVis(self.g, _lib.ptr(buffer), buffer.shape[0]), "SetVis")
        self._resident = buffer.shape[0]

    def attach_vis(self, dptr: int, nrow: int) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTAttachVis(self.g, C.c_void_p(dptr), nrow), "AttachVis")
        self._resident = nrow

    def get_vis(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty((self._resident, self.uv.lrec), dtype=np.float32)
        _lib.check(self.lib.ObitHIPSkyModelDFTGetVis(self.g, _lib.ptr(out), out.shape[0]), "GetVis")
        return out

    def calc_resident(self) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTCalcResident(self.g), "CalcResident")

    def calc_fp64(self) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTCalcFP64(self.g), "CalcFP64")

    def graph_launch(self) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTGraphLaunch(self.g), "GraphLaunch")

    def set_variant(self, cpt: int, tile: int, wg_per_cu: int) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTSetVariant(self.g, cpt, tile, wg_per_cu), "SetVariant")

    def variant(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self.lib.ObitHIPSkyModelDFTGetVariant(self.g, C.byref(a), C.byref(b), C.byref(c)),
                   "GetVariant")
        return a.value, b.value, c.value

    def path(self) -> int:
        """Kernel of the last launch: 0 split precision, 1 the FP64 kernel
        (taken when the channel frequencies are not uniform within an IF)."""
        p = C.c_int()
        _lib.check(self.lib.ObitHIPSkyModelDFTGetPath(self.g, C.byref(p)), "GetPath")
        return p.value

    def buffer_shape(self) -> tuple[int, int]:
        """(channels per thread, component parts) of the last per-buffer
        launch (calc)."""
        a, b = C.c_int(), C.c_int()
        _lib.check(self.lib.ObitHIPSkyModelDFTGetBufferShape(self.g, C.byref(a), C.byref(b)),
                   "GetBufferShape")
        return a.value, b.value

    def set_precision(self, prec: int) -> None:
        _lib.check(self.lib.ObitHIPSkyModelDFTSetPrecision(self.g, prec), "SetPrecision")

    def autotune(self, sample_rows: int = 16384, reps: int = 2):
        _lib.check(self.lib.ObitHIPSkyModelDFTAutotune(self.g, sample_rows, reps), "Autotune")
        return self.variant()

    def stream(self) -> int:
        return self.lib.ObitHIPSkyModelDFTStream(self.g)

    def close(self) -> None:
        if self.g:
            self.lib.ObitHostDFTShutdown(self.g)
            self.g = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
