"""uv gridding on the GPU, through Obit's C entry points.

``GPUGrid`` mirrors how Obit's imaging code drives ObitGPUGrid
(src/ObitImageUtil.c:1285-1322 create/set-up per facet,
src/ObitUVGrid.c:605-636 per-buffer 2GPU + Grid, then Flip and 2Host):
create for nfacet facets, set each facet up, grid buffers (or a resident
device buffer), flip, and copy every facet's planes back trimmed and
centred as the FFT expects.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .skymodel import host_uvdesc
from .synth import UVData, Facet


def plane_of_channel(nchanIF: int, bChan, eChan, cpu_parity: bool = True) -> np.ndarray:
    """Coarse-bin plane per channel (ObitHostPlanesFromBins)."""
    lib = _lib.load()
    b = np.ascontiguousarray(bChan, dtype=np.int32)
    e = np.ascontiguousarray(eChan, dtype=np.int32)
    out = np.empty(nchanIF, dtype=np.int32)
    lib.ObitHostPlanesFromBins(nchanIF, len(b), b.ctypes.data_as(C.POINTER(C.c_int)),
                               e.ctypes.data_as(C.POINTER(C.c_int)), int(cpu_parity),
                               out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def conv_func_sph() -> np.ndarray:
    """The 708-entry spheroidal GCF table (ConvFunc(in,5) convfn)."""
    out = np.zeros(708, dtype=np.float32)
    _lib.load().ObitHostConvFuncSph(_lib.ptr(out))
    return out


class GPUGrid:
    def __init__(self, uv: UVData, facets: list[Facet], nplane: int = 1, plane_of_chan=None,
                 sigma=None, beam: bool = False, devices=(0,), nVisPIO: int = 8192,
                 bmTaper: float = 0.0, blmax: float = 0.0, blmin: float = 0.0):
        self.lib = _lib.load()
        self.uv = uv
        self.facets = facets
        self.nplane = nplane
        dv, self._keep = host_uvdesc(uv)
        self._desc = dv
        gp = (C.c_int * len(devices))(*devices)
        self.c = self.lib.ObitHostGridCreate(len(facets), len(devices), gp, C.byref(dv), nplane, 1, nVisPIO)
        if not self.c:
            raise _lib.HipError("ObitHostGridCreate failed (bad device?)")
        nci = uv.nchan * uv.nif
        poc = None
        if plane_of_chan is not None:
            poc = np.ascontiguousarray(plane_of_chan, dtype=np.int32)
            assert poc.shape[0] == nci
        sig = [None, None, None]
        if sigma is not None:
            sig = [np.ascontiguousarray(s, dtype=np.float32) for s in sigma]
        self._keep2 = (poc, sig)
        for i, f in enumerate(facets):
            hf = _lib.HostFacet()
            hf.nx, hf.ny = f.nx, f.ny
            hf.cellx, hf.celly = f.cell_deg, f.cell_deg
            for k in range(3):
                hf.shift[k] = float(f.shift[k])
            for k in range(9):
                hf.rotUV[k] = float(f.rot[k])
            hf.blmax, hf.blmin = blmax, blmin
            hf.bmTaperUV = bmTaper
            hf.doBeam = int(beam)
            _lib.check(self.lib.ObitHostGridSetFacet(
                self.c, i, C.byref(hf), C.byref(dv), _lib.ptr(poc),
                _lib.ptr(sig[0]), _lib.ptr(sig[1]), _lib.ptr(sig[2])), f"SetFacet {i}")

    def grid(self, buffer: np.ndarray) -> None:
        """Grid a host buffer (nVisPIO rows per ObitCUDAGridGrid call)."""
        buffer = np.ascontiguousarray(buffer, dtype=np.float32)
        _lib.check(self.lib.ObitHostGridBuffer(self.c, _lib.ptr(buffer), buffer.shape[0]), "GridBuffer")

    def grid_device(self, dptr: int, nrow: int, igpu: int = 0) -> None:
        """Grid nrow rows resident in device memory in one pass."""
        _lib.check(self.lib.ObitHIPGridGridDevice(self.c, igpu, C.c_void_p(dptr), nrow), "GridDevice")

    def grid_device_reuse(self, dptr: int, nrow: int, igpu: int = 0) -> None:
        """Grid rows with the same u, v, w as the last pass, reusing its
        track lists (ObitHIPGridGridDeviceReuse)."""
        _lib.check(self.lib.ObitHIPGridGridDeviceReuse(self.c, igpu, C.c_void_p(dptr), nrow), "GridDeviceReuse")

    def flip(self) -> None:
        _lib.check(self.lib.ObitHostGridFlip(self.c), "Flip")

    def to_host(self, ifacet: int) -> np.ndarray:
        """Planes of a facet, trimmed and centred: complex64 [nplane, ny, nx/2+1]."""
        f = self.facets[ifacet]
        out = np.zeros((self.nplane, f.ny, f.nx // 2 + 1, 2), dtype=np.float32)
        _lib.check(self.lib.ObitHostGrid2Host(self.c, ifacet, _lib.ptr(out)), "Grid2Host")
        return out[..., 0] + 1j * out[..., 1]

    def to_host_device_path(self, ifacet: int) -> np.ndarray:
        """As to_host, with trim/centre done by the device kernel."""
        f = self.facets[ifacet]
        out = np.zeros((self.nplane, f.ny, f.nx // 2 + 1, 2), dtype=np.float32)
        _lib.check(self.lib.ObitHIPGrid2HostPlanes(self.c, ifacet, _lib.ptr(out)), "Grid2HostPlanes")
        return out[..., 0] + 1j * out[..., 1]

    def fft_image(self, ifacet: int, is_beam: bool, beam_norm=None):
        """Planes of a flipped facet -> corrected, normalised images on the
        device (ObitHIPGridFFT2Image), copied back: (float32 [nplane, ny, nx],
        per-plane beam normalisation)."""
        f = self.facets[ifacet]
        norm = np.zeros(self.nplane, np.float32)
        if not is_beam:
            norm[:] = np.asarray(beam_norm, np.float32)
        out = np.empty((self.nplane, f.ny, f.nx), np.float32)
        d = C.c_void_p()
        _lib.check(self.lib.ObitHIPMalloc(C.byref(d), out.nbytes), "malloc image")
        try:
            _lib.check(self.lib.ObitHIPGridFFT2Image(self.c, ifacet, int(is_beam),
                                                     norm.ctypes.data_as(C.POINTER(C.c_float)), d),
                       "FFT2Image")
            _lib.check(self.lib.ObitHIPMemcpyD2H(_lib.ptr(out), d, out.nbytes), "image to host")
        finally:
            self.lib.ObitHIPFree(d)
        return out, norm

    def prep_debug(self, ifacet: int, buffer: np.ndarray):
        """Per-sample (iu, iv, itu, itv, vis) as the device computes them."""
        buffer = np.ascontiguousarray(buffer, dtype=np.float32)
        n = buffer.shape[0] * self.uv.nchan * self.uv.nif
        iu, iv, itu, itv = (np.empty(n, np.int32) for _ in range(4))
        vis = np.empty(2 * n, np.float32)
        _lib.check(self.lib.ObitHIPGridPrepDebug(self.c, ifacet, _lib.ptr(buffer), buffer.shape[0],
                                                 _lib.ptr(iu), _lib.ptr(iv), _lib.ptr(itu),
                                                 _lib.ptr(itv), _lib.ptr(vis)), "PrepDebug")
        return iu, iv, itu, itv, vis

    def stream(self, igpu: int = 0) -> int:
        return self.lib.ObitHIPGridStream(self.c, igpu)

    def close(self) -> None:
        if getattr(self, "c", None):
            self.lib.ObitHostGridShutdown(self.c)
            self.c = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
