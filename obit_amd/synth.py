"""Seeded synthetic visibility data sets and sky models.

The five workloads are those the benchmark and the parity tests use:

1. 27 antennas uniform-random on a 3-arm Y (<= 36 km), dec +30, 600 x 10 s,
   1 IF x 64 ch x 1 MHz from 1.4 GHz, RR/LL, 1,000 point components
   (|l|,|m| <= 0.5 deg, exponential fluxes, mean 10 mJy); 1024^2 grid.
2. 64 antennas (8 km core + 3 log-radial arms), 450 x 8 s, 4 IF x 64 ch at
   1.28 GHz, RR/LL, 10,000 point components, power-law fluxes (dN/dS ~
   S^-1.5 on [1 uJy, 1 Jy]), |l|,|m| <= 0.7 deg.
3. Workload-1 visibilities, 40 facets with log-uniform component counts
   10..20,000, 20 % Gaussians (FWHM 2..30 arcsec).
4. Workload-1 geometry, 256 ch, Stokes-I records with weights U(0.5, 2),
   8 coarse bins x 40 facet grids of 2048^2.
5. Workload-2 array, 1,800 integrations, 4 IF x 256 ch, 40 facets of
   power-law point components with spectral index N(-0.7, 0.3).

Data layout is Obit's uv record (ObitUVDesc): 5 random parameters
(UU, VV, WW in wavelengths at the reference frequency, BASELINE, DATE)
then [IF][chan][Stokes][re, im, weight] float32.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

C_LIGHT = 299792458.0
DG2RAD = math.pi / 180.0
SIDEREAL = 2.0 * math.pi / 86164.0905  # rad/s
SEED0 = 20261202


@dataclasses.dataclass
class UVData:
    """A uv buffer with the ObitUVDesc fields the hot path reads."""
    buffer: np.ndarray          # float32 [nrow, lrec]
    nchan: int
    nif: int
    nstok: int
    freq: float                 # reference frequency (Hz)
    freqArr: np.ndarray         # float64 [nif*nchan], IF-major (freq axis first)
    crval_stokes: float = -1.0  # -1: RR,LL,...  +1: I
    nrparm: int = 5
    ilocu: int = 0
    ilocv: int = 1
    ilocw: int = 2
    ilocb: int = 3
    iloct: int = 4

    @property
    def nrow(self) -> int:
        return self.buffer.shape[0]

    @property
    def lrec(self) -> int:
        return self.buffer.shape[1]

    @property
    def incs(self) -> int:
        return 3

    @property
    def incf(self) -> int:
        return 3 * self.nstok

    @property
    def incif(self) -> int:
        return 3 * self.nstok * self.nchan

    @property
    def fscale(self) -> np.ndarray:
        return (self.freqArr / self.freq).astype(np.float32)

    def data_view(self) -> np.ndarray:
        """[nrow, nif, nchan, nstok, 3] view of the correlations."""
        return self.buffer[:, self.nrparm:].reshape(self.nrow, self.nif, self.nchan, self.nstok, 3)

    def subset(self, rows) -> "UVData":
        return dataclasses.replace(self, buffer=np.ascontiguousarray(self.buffer[rows]))


@dataclasses.dataclass
class SkyModel:
    """An Obit component table (ObitSkyModel comps member) and its type."""
    comps: np.ndarray           # float32 [ncomp, lcomp]
    modType: int                # ObitSkyModelCompType
    nSpecTerm: int = 0
    facet_of: np.ndarray | None = None   # facet index per component
    specFreq: np.ndarray | None = None   # tabulated-spectrum bin frequencies
    specIndex: np.ndarray | None = None  # bin per channel
    imgRefFreq: float = 0.0


@dataclasses.dataclass
class Facet:
    nx: int
    ny: int
    cell_deg: float
    shift: np.ndarray          # float32[3]
    rot: np.ndarray            # float32[9]
    l0: float = 0.0            # facet centre offset (rad)
    m0: float = 0.0


# ---------------------------------------------------------------- arrays

def y_array(rng, nant=27, rmax_km=21.0, rmin_km=0.1, arms_deg=(5.0, 125.0, 245.0)):
    """Antennas uniform-random in radius on a 3-arm Y; ENU metres."""
    per = nant // len(arms_deg)
    pos = []
    for k, az in enumerate(arms_deg):
        n = per + (1 if k < nant - per * len(arms_deg) else 0)
        r = rng.uniform(rmin_km, rmax_km, n) * 1e3
        a = az * DG2RAD
        pos += [(ri * math.sin(a), ri * math.cos(a), 0.0) for ri in r]
    return np.array(pos)


def core_arm_array(rng, ncore=40, core_diam_km=8.0, narm=3, per_arm=8, rlo_km=4.5, rhi_km=25.0):
    """Uniform core disc plus log-radially spaced arm stations; ENU metres."""
    rr = core_diam_km / 2 * 1e3 * np.sqrt(rng.uniform(0, 1, ncore))
    th = rng.uniform(0, 2 * math.pi, ncore)
    pos = [(r * math.sin(t), r * math.cos(t), 0.0) for r, t in zip(rr, th)]
    radii = np.geomspace(rlo_km, rhi_km, per_arm) * 1e3
    for k in range(narm):
        a = (5.0 + 120.0 * k) * DG2RAD
        pos += [(r * math.sin(a), r * math.cos(a), 0.0) for r in radii]
    return np.array(pos)


def uvw_tracks(enu, lat_deg, dec_deg, ntime, dt_s, freq):
    """uvw (wavelengths at freq) for every baseline i<j and integration,
    hour angles centred on transit.  Rows are time-major then baseline."""
    lat, dec = lat_deg * DG2RAD, dec_deg * DG2RAD
    E, N, U = enu[:, 0], enu[:, 1], enu[:, 2]
    X = -math.sin(lat) * N + math.cos(lat) * U
    Y = E
    Z = math.cos(lat) * N + math.sin(lat) * U
    ia, ib = np.triu_indices(len(enu), 1)
    bx, by, bz = X[ib] - X[ia], Y[ib] - Y[ia], Z[ib] - Z[ia]
    t = (np.arange(ntime) - (ntime - 1) / 2.0) * dt_s
    H = (SIDEREAL * t)[:, None]
    sd, cd = math.sin(dec), math.cos(dec)
    u = np.sin(H) * bx + np.cos(H) * by
    v = -sd * np.cos(H) * bx + sd * np.sin(H) * by + cd * bz
    w = cd * np.cos(H) * bx - cd * np.sin(H) * by + sd * bz
    scale = freq / C_LIGHT
    nb = len(ia)
    bl = np.broadcast_to((256 * (ia + 1) + (ib + 1)).astype(np.float64), (ntime, nb))
    tt = np.broadcast_to((t / 86400.0)[:, None], (ntime, nb))
    hdr = np.stack([u * scale, v * scale, w * scale, bl, tt], axis=-1).reshape(-1, 5)
    return hdr


def freq_table(f0, nif, nchan, df, if_step=None):
    if if_step is None:
        if_step = nchan * df
    f = np.empty(nif * nchan)
    for i in range(nif):
        f[i * nchan:(i + 1) * nchan] = f0 + i * if_step + np.arange(nchan) * df
    return f


def make_uv(hdr, nif, nchan, nstok, freq, freqArr, rng, weights=None, crval_stokes=-1.0,
            data=True):
    nrow = hdr.shape[0]
    lrec = 5 + 3 * nif * nchan * nstok
    buf = np.empty((nrow, lrec), dtype=np.float32)
    buf[:, :5] = hdr
    d = buf[:, 5:].reshape(nrow, nif * nchan * nstok, 3)
    if data:
        d[:, :, 0:2] = rng.standard_normal((nrow, nif * nchan * nstok, 2), dtype=np.float32)
    else:
        d[:, :, 0:2] = 0.0
    if weights is None:
        d[:, :, 2] = 1.0
    else:
        d[:, :, 2] = weights(rng, (nrow, nif * nchan * nstok)).astype(np.float32)
    return UVData(buffer=buf, nchan=nchan, nif=nif, nstok=nstok, freq=freq,
                  freqArr=np.asarray(freqArr, dtype=np.float64), crval_stokes=crval_stokes)


# ---------------------------------------------------------------- sky models

def point_row(flux, l_rad, m_rad):
    """Obit table row of a point (ObitSkyModel.c:1965-1990): amplitude,
    2 pi l, 2 pi m, 2 pi (n - 1) (radians per wavelength)."""
    n1 = np.sqrt(np.maximum(0.0, 1.0 - l_rad ** 2 - m_rad ** 2)) - 1.0
    return np.stack([flux, 2 * math.pi * l_rad, 2 * math.pi * m_rad, 2 * math.pi * n1], axis=-1)


def gauss_coeffs(maj_deg, min_deg, pa_deg):
    """gp1, gp2, gp3 of an elliptical Gaussian from FWHM (deg) and position
    angle, as ObitSkyModelLoadComps forms them (ObitSkyModel.c:1722,
    1939-1947): konst2 = DG2RAD sqrt(2) pi / 2.35482044."""
    konst2 = DG2RAD * math.sqrt(2.0) * math.pi / 2.35482044
    cpa, spa = np.cos(DG2RAD * pa_deg), np.sin(DG2RAD * pa_deg)
    xmaj, xmin = maj_deg * konst2, min_deg * konst2
    gp1 = -((cpa * xmaj) ** 2 + (spa * xmin) ** 2)
    gp2 = -((spa * xmaj) ** 2 + (cpa * xmin) ** 2)
    gp3 = -2.0 * cpa * spa * (xmaj * xmaj - xmin * xmin)
    return gp1, gp2, gp3


def exp_fluxes(rng, n, mean=0.01):
    return rng.exponential(mean, n)


def powerlaw_fluxes(rng, n, smin=1e-6, smax=1.0):
    """dN/dS ~ S^-1.5 on [smin, smax] by inversion of N(>S) ~ S^-0.5."""
    a, b = smin ** -0.5, smax ** -0.5
    return (a - rng.uniform(0, 1, n) * (a - b)) ** -2.0


def hex_centres(n, spacing_rad):
    """n facet centres on a hexagonal tiling around the phase centre."""
    pts = [(0.0, 0.0)]
    ring = 1
    while len(pts) < n:
        for side in range(6):
            a0 = math.pi / 3 * side
            a1 = math.pi / 3 * (side + 1)
            for k in range(ring):
                f = k / ring
                x = ring * ((1 - f) * math.cos(a0) + f * math.cos(a1))
                y = ring * ((1 - f) * math.sin(a0) + f * math.sin(a1))
                pts.append((x * spacing_rad, y * spacing_rad))
        ring += 1
    return np.array(pts[:n])


def facet_rotation(l0, m0):
    """uvw rotation to a facet tangent point (Rodrigues rotation taking the
    phase-centre direction to the facet direction)."""
    n0 = math.sqrt(max(0.0, 1.0 - l0 * l0 - m0 * m0))
    a = np.array([0.0, 0.0, 1.0])
    b = np.array([l0, m0, n0])
    v = np.cross(a, b)
    s = np.linalg.norm(v)
    c = float(a @ b)
    if s < 1e-15:
        return np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * ((1 - c) / s ** 2)


def make_facets(nfacet, nx, cell_deg, spacing_rad):
    out = []
    for (l0, m0) in hex_centres(nfacet, spacing_rad):
        n0 = math.sqrt(1 - l0 * l0 - m0 * m0)
        shift = np.array([-2 * math.pi * l0, -2 * math.pi * m0, -2 * math.pi * (n0 - 1)], np.float32)
        rot = facet_rotation(l0, m0).astype(np.float32).reshape(9)
        out.append(Facet(nx=nx, ny=nx, cell_deg=cell_deg, shift=shift, rot=rot, l0=l0, m0=m0))
    return out


# ---------------------------------------------------------------- workloads

def _rows_subset(hdr, rows, rng):
    if rows is None or rows >= hdr.shape[0]:
        return hdr
    idx = np.sort(rng.choice(hdr.shape[0], rows, replace=False))
    return hdr[idx]


def workload(cfg: int, rows: int | None = None, seed: int | None = None, data: bool = True,
             ncomp: int | None = None):
    """Build workload cfg (1..5).  rows selects a seeded subset of rows (the
    geometry is that of the full workload); ncomp overrides the component
    count of single-facet models.  The sky model is drawn from its own
    seeded stream and does not depend on rows.  Returns (uv, sky, facets)."""
    rng = np.random.default_rng(SEED0 + cfg if seed is None else seed)
    if cfg in (1, 3, 4):
        enu = y_array(rng)
        nchan = 256 if cfg == 4 else 64
        f0 = 1.4e9
        freqArr = freq_table(f0, 1, nchan, 1e6)
        hdr = uvw_tracks(enu, 34.08, 30.0, 600, 10.0, f0)
        hdr = _rows_subset(hdr, rows, rng)
        if cfg == 4:
            uv = make_uv(hdr, 1, nchan, 1, f0, freqArr, rng,
                         weights=lambda r, s: r.uniform(0.5, 2.0, s), crval_stokes=1.0, data=data)
        else:
            uv = make_uv(hdr, 1, nchan, 2, f0, freqArr, rng, data=data)
    else:
        enu = core_arm_array(rng)
        f0 = 1.28e9
        if cfg == 2:
            freqArr = freq_table(f0, 4, 64, 1e6)
            hdr = uvw_tracks(enu, 34.08, 30.0, 450, 8.0, f0)
            hdr = _rows_subset(hdr, rows, rng)
            uv = make_uv(hdr, 4, 64, 2, f0, freqArr, rng, data=data)
        else:
            freqArr = freq_table(f0, 4, 256, 1e6)
            hdr = uvw_tracks(enu, 34.08, 30.0, 1800, 8.0, f0)
            hdr = _rows_subset(hdr, rows, rng)
            uv = make_uv(hdr, 4, 256, 2, f0, freqArr, rng, data=data)

    # the sky has a stream of its own, so the model of a workload is the
    # same whatever row subset the visibilities take
    rng = np.random.default_rng([SEED0 + cfg if seed is None else seed, 1])
    facets = []
    if cfg == 1:
        n = ncomp or 1000
        l = rng.uniform(-0.5, 0.5, n) * DG2RAD
        m = rng.uniform(-0.5, 0.5, n) * DG2RAD
        sky = SkyModel(comps=point_row(exp_fluxes(rng, n), l, m).astype(np.float32), modType=0)
        facets = make_facets(1, 1024, 0.4 / 3600.0, 0.0)
    elif cfg == 2:
        n = ncomp or 10000
        l = rng.uniform(-0.7, 0.7, n) * DG2RAD
        m = rng.uniform(-0.7, 0.7, n) * DG2RAD
        sky = SkyModel(comps=point_row(powerlaw_fluxes(rng, n), l, m).astype(np.float32), modType=0)
    elif cfg in (3, 5):
        nf = 40
        spacing = 0.3 * DG2RAD
        centres = hex_centres(nf, spacing)
        counts = np.round(np.exp(rng.uniform(math.log(10), math.log(20000), nf))).astype(int)
        rows_ = []
        fac = []
        for fi, (lc, mc) in enumerate(centres):
            k = counts[fi]
            l = lc + rng.uniform(-0.15, 0.15, k) * DG2RAD
            m = mc + rng.uniform(-0.15, 0.15, k) * DG2RAD
            if cfg == 3:
                flux = exp_fluxes(rng, k)
                base = point_row(flux, l, m)
                g = np.zeros((k, 3))
                isg = rng.uniform(0, 1, k) < 0.2
                maj = rng.uniform(2.0, 30.0, k) / 3600.0
                mnr = maj * rng.uniform(0.3, 1.0, k)
                pa = rng.uniform(0, 180.0, k)
                gp = np.stack(gauss_coeffs(maj, mnr, pa), axis=-1)
                g[isg] = gp[isg]
                rows_.append(np.concatenate([base, g], axis=1))
            else:
                flux = powerlaw_fluxes(rng, k)
                alpha = rng.normal(-0.7, 0.3, k)
                rows_.append(np.concatenate([point_row(flux, l, m), alpha[:, None]], axis=1))
            fac.append(np.full(k, fi))
        comps = np.concatenate(rows_).astype(np.float32)
        if cfg == 3:
            sky = SkyModel(comps=comps, modType=1, facet_of=np.concatenate(fac))
        else:
            sky = SkyModel(comps=comps, modType=10, nSpecTerm=1, facet_of=np.concatenate(fac),
                           imgRefFreq=f0)
        facets = make_facets(nf, 2048, 0.4 / 3600.0, spacing)
    else:  # cfg == 4: gridding only
        sky = None
        facets = make_facets(40, 2048, 0.4 / 3600.0, 0.3 * DG2RAD)
    return uv, sky, facets


def coarse_bins(nchanIF: int, nbin: int):
    """Equal coarse-frequency bins: 0-relative inclusive (bChan, eChan)."""
    per = nchanIF // nbin
    b = np.arange(nbin) * per
    e = b + per - 1
    e[-1] = nchanIF - 1
    return b.astype(np.int32), e.astype(np.int32)
