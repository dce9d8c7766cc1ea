"""Write tests/golden/uvsub_fixture.npz from Obit's UVSub test data.

Reads (as data, with a small FITS reader of its own) three files the
reference's own UVSub test uses (testScripts/testUVSub.py,
test/testUVSub.c):

  testData/UVSubTestIn.uvtab      VLA 8.4 GHz uv data, 2 IF x RR/LL/RL/LR
  testData/UVSubTestModIn.fits.gz model image; CC table version 2 = 13 points
  testData/UVSubTestMaster.uvtab  the input divided by the DFT of that model

and keeps a seeded subset of rows of the input and of the master, the 13
components, the frequencies, and component tables in Obit's convention
(ObitSkyModel.c:1965-1990: [flux, 2 pi x, 2 pi y, 2 pi z]).

The component positions relative to the uv phase centre need Obit's image
and uv geometry (ObitSkyModelLoadComps with ObitUVDescShiftPhase /
ObitUVDescShift3DMatrix, ObitSkyModel.c:1889-1990), which this build does
not restate.  The script uses the textbook geometry instead (AIPS rotated
SIN projection about the image's CRVAL, direction cosines about the uv
phase centre, all FP64): `comps_geom`, with nothing fitted.  For
reference it also refines a linear position map and one flux scale by
least squares against the master's implied model M = V_in / V_master
(`comps`, with the fit's residual); the tests use `comps_geom`.
Everything else the master holds is left to the tests: the DFT's form and
sign, the per-IF frequency scaling, the amplitudes, and the divide
operation (data / M, weight x |M|).

Run in a container that has the reference tree:
    python tests/golden/make_uvsub_fixture.py [/root/reference]
"""
import gzip
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TD = os.path.join(REF, "ObitSystem", "Obit", "testData")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uvsub_fixture.npz")
NKEEP = 2000
SEED = 20261202


# ---------------------------------------------------------------- FITS
def _hdus(data):
    """Yield (keywords, data bytes) for every HDU of a FITS byte string."""
    off = 0
    while off + 2880 <= len(data):
        kw, cards = {}, 0
        while True:
            block = data[off:off + 2880]
            off += 2880
            done = False
            for i in range(36):
                card = block[80 * i:80 * (i + 1)].decode("ascii", "replace")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    val = card[10:]
                    if val.lstrip().startswith("'"):
                        v = val.split("'")[1].rstrip()
                    else:
                        v = val.split("/")[0].strip()
                    kw[key] = v
            if done:
                break
        naxis = int(kw.get("NAXIS", 0))
        size = 0
        if naxis:
            size = abs(int(kw.get("BITPIX", 8))) // 8
            for i in range(1, naxis + 1):
                size *= int(kw["NAXIS%d" % i])
            if "XTENSION" in kw:
                size += int(kw.get("PCOUNT", 0))
        yield kw, data[off:off + size]
        off += ((size + 2879) // 2880) * 2880


def _table(kw, raw):
    """Columns of a binary table of E/D/J fields as numpy arrays."""
    n1, n2 = int(kw["NAXIS1"]), int(kw["NAXIS2"])
    fmt = {"E": (">f4", 4), "D": (">f8", 8), "J": (">i4", 4)}
    cols, off, dt = [], 0, []
    for i in range(1, int(kw["TFIELDS"]) + 1):
        tf = kw["TFORM%d" % i].strip()
        cnt = int(tf[:-1] or 1)
        code, width = fmt[tf[-1]]
        name = kw["TTYPE%d" % i].strip()
        dt.append((name, code, (cnt,)))
        cols.append(name)
        off += cnt * width
    assert off == n1, (off, n1)
    arr = np.frombuffer(raw[:n1 * n2], dtype=np.dtype(dt))
    return {c: np.array(arr[c]) for c in cols}


def read_uvtab(path):
    data = open(path, "rb").read()
    out = {}
    for kw, raw in _hdus(data):
        ext = kw.get("EXTNAME", "")
        if ext == "AIPS UV":
            t = _table(kw, raw)
            cols = list(t)
            rows = np.concatenate([t[c].astype(np.float32).reshape(len(t[c]), -1) for c in cols], axis=1)
            out["rows"] = rows
            out["uvkw"] = kw
        elif ext == "AIPS FQ":
            out["fq"] = _table(kw, raw)
    return out


def read_model(path):
    data = gzip.decompress(open(path, "rb").read())
    img, cc = None, None
    for kw, raw in _hdus(data):
        if "XTENSION" not in kw:
            img = kw
        elif kw.get("EXTNAME") == "AIPS CC" and int(kw.get("EXTVER", 0)) == 2:
            t = _table(kw, raw)
            cc = np.stack([t["FLUX"][:, 0], t["DELTAX"][:, 0], t["DELTAY"][:, 0]], axis=1)
    return img, cc


# ------------------------------------------------------------ geometry
def direction_cosines(img, cc, ra_c, dec_c, sign):
    """(l, m, n-1) about the uv phase centre of each component: the CC
    offsets (deg, image axes) rotated by CROTA2, deprojected by SIN about
    the image's CRVAL, then projected about (ra_c, dec_c).  FP64."""
    d2r = np.pi / 180.0
    rho = sign * float(img["CROTA2"]) * d2r
    dx, dy = cc[:, 1].astype(np.float64), cc[:, 2].astype(np.float64)
    L = (dx * np.cos(rho) - dy * np.sin(rho)) * d2r
    M = (dx * np.sin(rho) + dy * np.cos(rho)) * d2r
    a0, d0 = float(img["CRVAL1"]) * d2r, float(img["CRVAL2"]) * d2r
    n = np.sqrt(1.0 - L * L - M * M)
    dec = np.arcsin(M * np.cos(d0) + n * np.sin(d0))
    ra = a0 + np.arctan2(L, n * np.cos(d0) - M * np.sin(d0))
    ac, dc = ra_c * d2r, dec_c * d2r
    l = np.cos(dec) * np.sin(ra - ac)
    m = np.sin(dec) * np.cos(dc) - np.cos(dec) * np.sin(dc) * np.cos(ra - ac)
    nn = np.sin(dec) * np.sin(dc) + np.cos(dec) * np.cos(dc) * np.cos(ra - ac)
    return np.stack([l, m, nn - 1.0], axis=1)


def model(uvw, fsc, flux, lmn):
    """sum_k flux_k exp(+2 pi i (u l + v m + w n) f/f0): [nrow, nif]."""
    ph = 2.0 * np.pi * (uvw @ lmn.T)  # [nrow, ncomp] at the reference frequency
    return np.stack([(np.exp(1j * ph * f) * flux).sum(axis=1) for f in fsc], axis=1)


def main():
    uvin = read_uvtab(os.path.join(TD, "UVSubTestIn.uvtab"))
    master = read_uvtab(os.path.join(TD, "UVSubTestMaster.uvtab"))
    img, cc = read_model(os.path.join(TD, "UVSubTestModIn.fits.gz"))
    kw = uvin["uvkw"]
    rin, rms_ = uvin["rows"], master["rows"]
    assert rin.shape == rms_.shape and np.array_equal(rin[:, :5], rms_[:, :5])
    dims = [int(x) for x in kw["TDIM6"].strip("()").split(",")]
    ncx, nstok, nchan, nif = dims[:4]
    assert ncx == 3 and nchan == 1
    f0 = float(kw["3CRVL6"])
    fq = uvin["fq"]
    freqs = f0 + fq["IF FREQ"][0].astype(np.float64)
    ra_c, dec_c = float(kw["5CRVL6"]), float(kw["6CRVL6"])
    print(f"{rin.shape[0]} rows, {nif} IF x {nchan} ch x {nstok} Stokes, f0={f0:.6g}, IF freqs {freqs}")
    print(f"uv centre {ra_c:.7f} {dec_c:.7f}; image CRVAL {img['CRVAL1']} {img['CRVAL2']} "
          f"CROTA2 {img['CROTA2']}; {len(cc)} components")

    vin = rin[:, 5:].reshape(-1, nif, nstok, 3).astype(np.float64)
    vma = rms_[:, 5:].reshape(-1, nif, nstok, 3).astype(np.float64)
    good = (vin[..., :2, 2] > 0) & (vma[..., :2, 2] > 0)
    Min = (vin[..., :2, 0] + 1j * vin[..., :2, 1])
    Mma = (vma[..., :2, 0] + 1j * vma[..., :2, 1])
    Mobs = np.where(good, Min / np.where(good, Mma, 1.0), 0.0)  # [nrow, nif, 2]
    uvw = rin[:, :3].astype(np.float64)
    fsc = freqs / f0
    flux = cc[:, 0].astype(np.float64)

    best = None
    for sign in (+1, -1):
        lmn = direction_cosines(img, cc, ra_c, dec_c, sign)
        mod = model(uvw, fsc, flux, lmn)[..., None]
        corr = np.abs(np.sum(np.conj(mod) * Mobs * good)) / np.sqrt(
            np.sum(np.abs(mod) ** 2 * good) * np.sum(np.abs(Mobs) ** 2 * good))
        print(f"rotation sign {sign:+d}: normalised correlation with the master's model {corr:.6f}")
        if best is None or corr > best[0]:
            best = (corr, sign, lmn)
    corr, sign, lmn0 = best

    # refine: lmn_k = lmn0_k + c + A (dx_k, dy_k), flux scale s  (10 parameters)
    from scipy.optimize import least_squares
    dxy = (cc[:, 1:3].astype(np.float64) - cc[:, 1:3].mean(axis=0)) * np.pi / 180.0
    sel = np.nonzero(good.any(axis=(1, 2)))[0]

    def unpack(p):
        c = p[0:3] * 1e-7
        A = p[3:9].reshape(3, 2) * 1e-2
        return lmn0 + c + dxy @ A.T, 1.0 + p[9] * 1e-3

    def resid(p):
        lmn, s = unpack(p)
        mod = s * model(uvw[sel], fsc, flux, lmn)[..., None]
        r = (mod - Mobs[sel]) * good[sel]
        return np.concatenate([r.real.ravel(), r.imag.ravel()])

    p0 = np.zeros(10)
    r0 = resid(p0)
    fit = least_squares(resid, p0, x_scale="jac")
    lmn, scale = unpack(fit.x)
    rfin = fit.fun
    npts = 2 * good.sum()
    print(f"fit: rms |M - M_master| {np.sqrt((r0 ** 2).sum() / npts):.3e} -> "
          f"{np.sqrt((rfin ** 2).sum() / npts):.3e} Jy (peak |M| {np.abs(Mobs).max():.3f}); "
          f"flux scale {scale:.8f}; centre moved {np.abs(fit.x[:3]).max() * 1e-7:.2e} rad")

    def table(flux_scale, pos):
        t = np.zeros((len(cc), 4), np.float32)
        t[:, 0] = (cc[:, 0].astype(np.float64) * flux_scale).astype(np.float32)
        t[:, 1:4] = (2.0 * np.pi * pos).astype(np.float32)
        return t

    comps_geom = table(1.0, lmn0)  # textbook geometry, nothing fitted
    comps = np.zeros((len(cc), 4), np.float32)
    comps[:, 0] = (cc[:, 0].astype(np.float64) * scale).astype(np.float32)
    comps[:, 1:4] = (2.0 * np.pi * lmn).astype(np.float32)

    rng = np.random.default_rng(SEED)
    keep = np.sort(rng.choice(rin.shape[0], NKEEP, replace=False))
    np.savez_compressed(
        OUT,
        rows_in=rin[keep], rows_master=rms_[keep], comps=comps, comps_geom=comps_geom,
        cc=cc.astype(np.float32),
        freqs=freqs, f0=np.float64(f0), nif=nif, nchan=nchan, nstok=nstok,
        uv_centre=np.array([ra_c, dec_c]), rot_sign=sign, flux_scale=scale,
        img_crval=np.array([float(img["CRVAL1"]), float(img["CRVAL2"])]),
        img_crota=np.float32(float(img["CROTA2"])), uv_crota=np.float32(float(kw["6CROT6"])),
        factor=np.float32(0.5), nfield=2,
        fit_rms=np.sqrt((rfin ** 2).sum() / npts), start_corr=corr, rows_kept=keep)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
