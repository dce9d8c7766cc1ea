"""Jones-corrected (antenna-beam-corrected) DFT sky model for linear feeds
through ObitHIPBeamDFT* (include/ObitHIPBeam.h).

Drives the layer the way Obit's beam-model host class drives its CUDA
layer (reference src/ObitGPUBeamModel.c: Init, SetMod, Calc per
parallactic angle, Shutdown; src/ObitCUDABeamModel.cu:1368-1437).  The
Jones table is handed in, set to unit matrices, or formed on the device
from beam images (GetJonesKernel, ObitCUDABeamModel.cu:1227-1308).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

OPS = {"sub": 0, "div": 1, "repl": 2}


def tspec_amps(flux, si, spec_index, freq_scale, freq_rat):
    """Per-channel amplitudes of tabulated spectra as the reference's
    beamPointTSpec / beamGaussTSpec kernels form them
    (ObitCUDABeamModel.cu:559-587): flux_sb exp(SI_sb ln(nu/nu0 * nu0/nu_sb))
    for the subband sb = spec_index[channel], in FP64.
    flux, si [ncomp, nspec]; spec_index, freq_scale [nif*nchan];
    freq_rat [nspec] = nu0 / nu_sb."""
    flux, si = np.asarray(flux, np.float64), np.asarray(si, np.float64)
    sb = np.asarray(spec_index)
    lnr = np.log(np.asarray(freq_scale, np.float64) * np.asarray(freq_rat, np.float64)[sb])
    return (flux[:, sb] * np.exp(si[:, sb] * lnr[None, :])).astype(np.float32)


def pspec_amps(amp, coef, freq_scale, freq_rat):
    """Per-channel amplitudes of a parameterised spectrum as the reference's
    beamPointSpec / beamGaussSpec kernels form them
    (ObitCUDABeamModel.cu:367-398): a exp(-sum_t c_t L^(t+1)) with
    L = -ln(nu/nu0 * freq_rat), in FP64.  amp [ncomp], coef [ncomp, nterm]."""
    L = -np.log(np.asarray(freq_scale, np.float64) * float(freq_rat))
    coef = np.asarray(coef, np.float64)
    arg = sum(coef[:, t:t + 1] * L[None, :] ** (t + 1) for t in range(coef.shape[1]))
    return (np.asarray(amp, np.float64)[:, None] * np.exp(-arg)).astype(np.float32)


def freq_lines(uv):
    """nu/nu0 of channel 0 of each IF and the per-channel step, from the
    IF-major frequency table (channels are evenly spaced within an IF)."""
    f = np.asarray(uv.freqArr, np.float64).reshape(uv.nif, uv.nchan) / uv.freq
    step = (f[:, 1] - f[:, 0]) if uv.nchan > 1 else np.zeros(uv.nif)
    lin = f[:, :1] + np.arange(uv.nchan)[None, :] * step[:, None]
    if np.max(np.abs(lin - f)) > 1e-9 * np.max(np.abs(f)):
        raise ValueError("channel frequencies are not evenly spaced within an IF")
    return np.ascontiguousarray(f[:, 0]), np.ascontiguousarray(step)


class BeamDFT:
    """One data layout, component table and Jones table on the device."""

    def __init__(self, uv, comps, ant_type, vis_jchan, njchan, ntype=None, stok_type=1, op="sub",
                 stok_fact=(1.0, 1.0, 1.0, 1.0), ncorr=None, stream=None, iloca=None,
                 chan_sel=None, if_sel=None):
        self.lib = _lib.load()
        self.uv = uv
        comps = np.ascontiguousarray(comps, np.float32)
        ant_type = np.ascontiguousarray(ant_type, np.int32)
        vis_jchan = np.ascontiguousarray(vis_jchan, np.int32)
        d = _lib.BeamDesc()
        d.lrec, d.nrparm = uv.lrec, uv.nrparm
        d.ilocu, d.ilocv, d.ilocw = uv.ilocu, uv.ilocv, uv.ilocw
        # iloca: (ANTENNA1, ANTENNA2) parameter offsets for data without BASELINE
        d.ilocb, d.iloca1, d.iloca2 = (uv.ilocb, -1, -1) if iloca is None else (-1, iloca[0], iloca[1])
        d.nif, d.nchan, d.incf, d.incif = uv.nif, uv.nchan, uv.incf, uv.incif
        d.ncorr = int(ncorr if ncorr is not None else uv.nstok)
        d.ncomp, d.compLen = comps.shape
        d.ntype = int(ntype if ntype is not None else ant_type.max() + 1)
        d.nant = len(ant_type)
        d.njchan = int(njchan)
        d.stokType, d.opType = int(stok_type), OPS[op]
        for i in range(4):
            d.stokFact[i] = float(stok_fact[i])
        # channel / IF selection (0-rel, inclusive); all by default
        d.chanb, d.chane = chan_sel if chan_sel is not None else (0, uv.nchan - 1)
        d.IFb, d.IFe = if_sel if if_sel is not None else (0, uv.nif - 1)
        self.desc = d
        f0, df = freq_lines(uv)
        self.h = C.c_void_p()
        _lib.check(self.lib.ObitHIPBeamDFTCreate(C.byref(self.h), C.byref(d), f0.ctypes.data, df.ctypes.data,
                                                 ant_type.ctypes.data, vis_jchan.ctypes.data, stream),
                   "ObitHIPBeamDFTCreate")
        _lib.check(self.lib.ObitHIPBeamDFTSetMod(self.h, comps.ctypes.data), "ObitHIPBeamDFTSetMod")

    def set_jones(self, jones=None):
        """jones: complex [ntype, ncomp, njchan, 2, 2] (or None: unit)."""
        if jones is None:
            _lib.check(self.lib.ObitHIPBeamDFTSetJones(self.h, None), "ObitHIPBeamDFTSetJones")
            return
        j = np.asarray(jones, np.complex64)
        flat = np.ascontiguousarray(j.reshape(j.shape[:3] + (4,)).view(np.float32))
        _lib.check(self.lib.ObitHIPBeamDFTSetJones(self.h, flat.ctypes.data), "ObitHIPBeamDFTSetJones")

    def set_amps(self, amps=None):
        """Per-channel amplitudes [ncomp, nif*nchan] (spectral types), or
        None for the table's amplitude column."""
        if amps is None:
            _lib.check(self.lib.ObitHIPBeamDFTSetAmps(self.h, None), "ObitHIPBeamDFTSetAmps")
            return
        a = np.ascontiguousarray(amps, np.float32)
        _lib.check(self.lib.ObitHIPBeamDFTSetAmps(self.h, a.ctypes.data), "ObitHIPBeamDFTSetAmps")

    def get_jones(self):
        """The device's Jones table as complex [ntype, ncomp, njchan, 2, 2]."""
        d = self.desc
        out = np.empty((d.ntype, max(d.ncomp, 1), d.njchan, 8), np.float32)
        _lib.check(self.lib.ObitHIPBeamDFTGetJones(self.h, out.ctypes.data), "ObitHIPBeamDFTGetJones")
        return out.view(np.complex64).reshape(d.ntype, max(d.ncomp, 1), d.njchan, 2, 2)

    def set_beams(self, descs, images, beam_freq, ref_size_fact, fblank):
        """descs: per type dicts (nx, ny, proj, hwidth, crval, crpix, cdelt);
        images float32 [ntype, 8, njchan, ny, nx] (or a list of per-type
        arrays [8, njchan, ny, nx]); beam_freq [njchan] Hz."""
        imgs = [np.asarray(x, np.float32) for x in images]
        if len(imgs) != len(descs) or len(descs) != self.desc.ntype:
            raise ValueError("one image set and one descriptor per antenna type")
        arr = (_lib.BeamImageDesc * len(descs))()
        for a, dd, im in zip(arr, descs, imgs):
            if im.ndim != 4 or im.shape[0] != 8 or im.shape[2:] != (dd["ny"], dd["nx"]):
                raise ValueError(f"beam images must be [8, nchan, {dd['ny']}, {dd['nx']}], not {im.shape}")
            a.nx, a.ny, a.nchan = dd["nx"], dd["ny"], im.shape[1]
            a.proj, a.hwidth = dd.get("proj", 0), dd.get("hwidth", 2)
            for i in range(2):
                a.crval[i], a.crpix[i], a.cdelt[i] = dd["crval"][i], dd["crpix"][i], dd["cdelt"][i]
        flat = np.ascontiguousarray(np.concatenate([x.ravel() for x in imgs]))
        bf = np.ascontiguousarray(beam_freq, np.float32)
        if bf.size != self.desc.njchan:
            raise ValueError(f"{bf.size} beam frequencies for {self.desc.njchan} beam channels")
        self._beams = (arr, flat, bf)
        _lib.check(self.lib.ObitHIPBeamDFTSetBeams(self.h, arr, flat.ctypes.data, bf.ctypes.data,
                                                   float(ref_size_fact), float(fblank)), "ObitHIPBeamDFTSetBeams")

    def jones_from_beams(self, pos, par_ang=0.0):
        """pos: float32 [ncomp, 2] RA and Dec offsets (deg)."""
        p = np.ascontiguousarray(pos, np.float32)
        _lib.check(self.lib.ObitHIPBeamDFTJonesFromBeams(self.h, p.ctypes.data, float(par_ang)),
                   "ObitHIPBeamDFTJonesFromBeams")

    def calc_device(self, d_vis, nrow, par_ang=0.0):
        _lib.check(self.lib.ObitHIPBeamDFTCalc(self.h, C.c_void_p(d_vis), nrow, float(par_ang)),
                   "ObitHIPBeamDFTCalc")

    def calc(self, buffer, par_ang=0.0):
        """Apply the model to a copy of host rows; returns the new rows."""
        buf = np.ascontiguousarray(buffer, np.float32).copy()
        d = C.c_void_p()
        _lib.check(self.lib.ObitHIPMalloc(C.byref(d), max(buf.nbytes, 4)), "malloc")
        try:
            _lib.check(self.lib.ObitHIPMemcpyH2D(d, buf.ctypes.data, buf.nbytes), "H2D")
            self.calc_device(d.value, buf.shape[0], par_ang)
            _lib.check(self.lib.ObitHIPDeviceSynchronize(), "sync")
            _lib.check(self.lib.ObitHIPMemcpyD2H(buf.ctypes.data, d, buf.nbytes), "D2H")
        finally:
            self.lib.ObitHIPFree(d)
        return buf

    def close(self):
        if self.h:
            self.lib.ObitHIPBeamDFTDestroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
