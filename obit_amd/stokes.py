"""Stokes I records on the device (ObitUVCalSelect with Stokes 'I',
src/ObitUVCalSelect.c:540-695, :697-760), the way the gridder consumes
them: one pass over a resident data set through ObitHIPStokesI."""
import ctypes as C

import numpy as np

from . import _lib
from .synth import UVData


def stokes_i_device(lib, d_in, d_out, uv, formal_i=False, stream=None):
    """Launch ObitHIPStokesI on device buffers laid out as uv describes."""
    _lib.check(lib.ObitHIPStokesI(C.c_void_p(d_in), C.c_void_p(d_out), uv.nrow, uv.lrec, uv.nrparm,
                                  uv.nif, uv.nchan, uv.nstok, uv.incs, uv.incf, uv.incif,
                                  int(round(uv.crval_stokes)), 1 if formal_i else 0, stream),
               "ObitHIPStokesI")


def stokes_i(uv, formal_i=False):
    """Stokes I records of uv (host buffer in, host UVData out), formed on
    the device."""
    lib = _lib.load()
    lout = uv.nrparm + 3 * uv.nif * uv.nchan
    out = np.empty((uv.nrow, lout), np.float32)
    buf = np.ascontiguousarray(uv.buffer, np.float32)
    d_in, d_out = C.c_void_p(), C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d_in), max(buf.nbytes, 4)), "malloc")
    try:
        _lib.check(lib.ObitHIPMalloc(C.byref(d_out), max(out.nbytes, 4)), "malloc")
        try:
            _lib.check(lib.ObitHIPMemcpyH2D(d_in, buf.ctypes.data, buf.nbytes), "H2D")
            stokes_i_device(lib, d_in.value, d_out.value, uv, formal_i)
            _lib.check(lib.ObitHIPDeviceSynchronize(), "sync")
            _lib.check(lib.ObitHIPMemcpyD2H(out.ctypes.data, d_out, out.nbytes), "D2H")
        finally:
            lib.ObitHIPFree(d_out)
    finally:
        lib.ObitHIPFree(d_in)
    return UVData(buffer=out, nchan=uv.nchan, nif=uv.nif, nstok=1, freq=uv.freq,
                  freqArr=uv.freqArr, crval_stokes=1.0, nrparm=uv.nrparm, ilocu=uv.ilocu,
                  ilocv=uv.ilocv, ilocw=uv.ilocw, ilocb=uv.ilocb, iloct=uv.iloct)
