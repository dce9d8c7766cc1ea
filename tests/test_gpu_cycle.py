"""Batched multi-facet DFT and the wideband major cycle (one subtract launch
for every facet's components, then regridding of the residual into every
facet x coarse-frequency plane), with errors reported per facet against the
FP64 oracle."""
import numpy as np
import pytest

from obit_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5    # DFT: max |dV| / peak |V_model|
GTOL = 2e-6   # gridding: max |dG| / peak |G| per plane


def _stokes0(uv, buf):
    d = buf[:, uv.nrparm:].reshape(buf.shape[0], uv.nif, uv.nchan, uv.nstok, 3)
    return d[..., 0, 0].astype(np.float64) + 1j * d[..., 0, 1]


def _model(uv, sky):
    from obit_amd.skymodel import SkyModelDFT
    with SkyModelDFT(uv, sky, op="repl", nVisPIO=uv.nrow) as sm:
        sm.set_vis(uv.buffer)
        sm.calc_resident()
        return _stokes0(uv, sm.get_vis())


def _facet_sky(sky, f):
    sel = sky.facet_of == f
    return synth.SkyModel(comps=np.ascontiguousarray(sky.comps[sel]), modType=sky.modType,
                          nSpecTerm=sky.nSpecTerm, imgRefFreq=sky.imgRefFreq)


def test_batched_multifacet_per_facet(oracle):
    """Workload 3 (40 facets, 10..20,000 components each, 20 % Gaussian):
    the one batched launch against the oracle, every facet's model against
    the oracle, and the facets' models adding up to the batched one."""
    uv, sky, _ = synth.workload(3, rows=96)
    batched = _model(uv, sky)
    ref = oracle.dft_model_fp64(uv, sky, nthreads=8)
    peak = np.abs(ref).max()
    err = np.abs(batched - ref).max() / peak
    print(f"batched, {sky.comps.shape[0]} components in 40 facets: max {err:.3e}", flush=True)
    assert err <= TOL
    total = np.zeros_like(batched)
    errs = []
    for f in range(40):
        skf = _facet_sky(sky, f)
        mf = _model(uv, skf)
        rf = oracle.dft_model_fp64(uv, skf, nthreads=8)
        d = np.abs(mf - rf)
        errs.append((d.max() / np.abs(rf).max(), np.sqrt(np.mean(d ** 2)) / np.abs(rf).max()))
        total += mf
    errs = np.array(errs)
    w = int(np.argmax(errs[:, 0]))
    print(f"per facet: worst max {errs[w, 0]:.3e} (facet {w}, {int((sky.facet_of == w).sum())} comps), "
          f"worst rms {errs[:, 1].max():.3e}", flush=True)
    assert (errs[:, 0] <= TOL).all()
    assert np.abs(total - batched).max() <= TOL * peak


def test_wideband_cycle_per_facet(oracle):
    """Workload 5's cycle on a row sample: subtract the 40-facet spectral
    model (S(nu) = S0 exp(alpha ln(nu/nu0))) from 1,024-channel RR/LL data
    in one launch, form Stokes I, and grid it into 40 facets x 8 coarse
    bins.  The subtraction is checked per facet through the oracle of each
    facet's components, the grids per facet and plane against the oracle's
    FP64 gridder applied to the same residual."""
    from obit_amd.gridder import GPUGrid, plane_of_channel
    from obit_amd.skymodel import SkyModelDFT
    uv, sky, facets = synth.workload(5, rows=24)
    with SkyModelDFT(uv, sky, op="sub", nVisPIO=uv.nrow) as sm:
        sm.set_vis(uv.buffer)
        sm.calc_resident()
        res = sm.get_vis()
    got = _stokes0(uv, uv.buffer) - _stokes0(uv, res)
    ref = oracle.dft_model_fp64(uv, sky, nthreads=8)
    peak = np.abs(ref).max()
    # the residual is an FP32 difference: allow its rounding on top
    slack = 4e-7 * np.abs(uv.buffer[:, 5:]).max() / peak
    err = np.abs(got - ref).max() / peak
    print(f"subtract, {sky.comps.shape[0]} spectral components: max {err:.3e}", flush=True)
    assert err <= TOL + slack
    errs = []
    for f in range(40):
        skf = _facet_sky(sky, f)
        mf = _model(uv, skf)
        rf = oracle.dft_model_fp64(uv, skf, nthreads=8)
        errs.append(np.abs(mf - rf).max() / np.abs(rf).max())
    print(f"subtract per facet: worst max {max(errs):.3e} (facet {int(np.argmax(errs))})", flush=True)
    assert max(errs) <= TOL

    # Stokes I records from the residual (RR, LL -> I)
    d = res[:, 5:].reshape(uv.nrow, uv.nif * uv.nchan, 2, 3).astype(np.float64)
    ok = (d[..., 0, 2] > 0) & (d[..., 1, 2] > 0)
    rec = np.zeros((uv.nrow, 5 + 3 * uv.nif * uv.nchan), np.float32)
    rec[:, :5] = res[:, :5]
    ri = rec[:, 5:].reshape(uv.nrow, -1, 3)
    ri[..., 0] = np.where(ok, 0.5 * (d[..., 0, 0] + d[..., 1, 0]), 0)
    ri[..., 1] = np.where(ok, 0.5 * (d[..., 0, 1] + d[..., 1, 1]), 0)
    ri[..., 2] = np.where(ok, 0.5 * (d[..., 0, 2] + d[..., 1, 2]), 0)
    uvI = synth.UVData(buffer=rec, nchan=uv.nchan, nif=uv.nif, nstok=1, freq=uv.freq,
                       freqArr=uv.freqArr, crval_stokes=1.0)
    small = [synth.Facet(nx=256, ny=256, cell_deg=f.cell_deg, shift=f.shift, rot=f.rot)
             for f in facets]
    nci = uv.nchan * uv.nif
    b, e = synth.coarse_bins(nci, 8)
    poc = plane_of_channel(nci, b, e, cpu_parity=False)
    worst = (0.0, -1, -1)
    with GPUGrid(uvI, small, nplane=8, plane_of_chan=poc, nVisPIO=uv.nrow) as gg:
        gg.grid(uvI.buffer)
        gg.flip()
        for i, f in enumerate(small):
            g = gg.to_host(i)
            for p in range(8):
                refg = oracle.trim_centre(oracle.grid_fp64(uvI, f, bChan=int(b[p]), eChan=int(e[p]) + 1), f)
                pk = np.abs(refg).max()
                assert pk > 0
                eg = np.abs(g[p] - refg).max() / pk
                if eg > worst[0]:
                    worst = (eg, i, p)
    print(f"regrid per facet: worst max |dG|/peak {worst[0]:.3e} (facet {worst[1]}, plane {worst[2]})",
          flush=True)
    assert worst[0] < GTOL
