"""ObitHIPSkyModelDFTCalcNoAlloc (the DFT call that allocates nothing)
against ObitHIPSkyModelDFTCalcResident on the same inputs: two calls in a
row on one set-up, the second with fewer rows and components, each
bit-identical to the resident path's result (same kernels, same frequency
lines); sizes beyond the scratch and spectral types are refused."""
import ctypes as C

import numpy as np
import pytest

from obit_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _to_device(lib, a):
    d = C.c_void_p()
    _lib.check(lib.ObitHIPMalloc(C.byref(d), a.nbytes), "malloc")
    _lib.check(lib.ObitHIPMemcpyH2D(d, _lib.ptr(a), a.nbytes), "upload")
    return d


def _resident(uv, sky, **kw):
    from obit_amd.skymodel import SkyModelDFT
    with SkyModelDFT(uv, sky, **kw) as sm:
        sm.set_vis(uv.buffer)
        sm.calc_resident()
        return sm.get_vis()


def test_no_alloc_dft_matches_resident():
    from obit_amd.skymodel import SkyModelDFT
    lib = _lib.load()
    uv, sky, _ = synth.workload(3, rows=400)      # points and Gaussians, 64 channels
    assert sky.comps.shape[1] == 7 and sky.comps.shape[0] > 20000
    uv2 = uv.subset(np.arange(150))
    sky2 = synth.SkyModel(comps=np.ascontiguousarray(sky.comps[:3000]), modType=sky.modType)
    cases = [(uv, sky), (uv2, sky2)]
    refs = [_resident(u, s, op="sub") for u, s in cases]
    with SkyModelDFT(uv, sky, op="sub") as sm:
        for (u, s), ref in zip(cases, refs):
            dv = _to_device(lib, u.buffer)
            dm = _to_device(lib, s.comps)
            try:
                _lib.check(lib.ObitHIPSkyModelDFTCalcNoAlloc(sm.g, dm, s.comps.shape[0], dv, u.nrow), "NoAlloc")
                _lib.check(lib.ObitHIPDeviceSynchronize(), "sync")
                out = np.empty_like(u.buffer)
                _lib.check(lib.ObitHIPMemcpyD2H(_lib.ptr(out), dv, out.nbytes), "download")
            finally:
                lib.ObitHIPFree(dv)
                lib.ObitHIPFree(dm)
            print(f"{u.nrow} rows x {s.comps.shape[0]} components: "
                  f"max |d| {np.abs(out - ref).max():.3e}", flush=True)
            assert np.array_equal(out, ref)
        # more components than the scratch holds: refused before any launch
        rc = lib.ObitHIPSkyModelDFTCalcNoAlloc(sm.g, C.c_void_p(16), 900000, C.c_void_p(16), 10)
        assert rc == 1 and b"scratch" in lib.ObitHIPLastError()


def test_no_alloc_dft_refuses_spectral_types():
    from obit_amd.skymodel import SkyModelDFT
    lib = _lib.load()
    uv, sky, _ = synth.workload(5, rows=8)
    with SkyModelDFT(uv, sky, op="sub") as sm:
        rc = lib.ObitHIPSkyModelDFTCalcNoAlloc(sm.g, C.c_void_p(16), 10, C.c_void_p(16), 8)
        assert rc == 1 and b"spectral" in lib.ObitHIPLastError()
