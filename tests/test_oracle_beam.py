"""The Jones-corrected DFT oracle (oracle.beam_*), pinned without running
the reference: against the plain-DFT oracle (itself pinned by Obit's
golden vectors) for unit Jones matrices, against the reference GPU's
term-by-term expansion of the products (include/CUDALinCorSum.h), against
closed forms for diagonal Jones matrices, a brute-force loop on a tiny
case, and the symmetries the coherency matrices must have."""
import math

import numpy as np
import pytest

from obit_amd import synth
from oracle import oracle as orc


def _jones(rng, ntype, ncomp, njchan, scale=0.2):
    j = np.zeros((ntype, ncomp, njchan, 2, 2), np.complex128)
    j[..., 0, 0] = 1.0
    j[..., 1, 1] = 1.0
    j += scale * (rng.normal(size=j.shape) + 1j * rng.normal(size=j.shape))
    return j


def _small(rng, rows=48, ncomp=12):
    uv, sky, _ = synth.workload(1, rows=rows)
    comps = sky.comps[:ncomp].copy()
    return uv, comps


def test_unit_jones_stokes_i_is_the_plain_dft():
    """J = 1, Stokes I: XX = YY = the plain DFT model, XY = YX = 0."""
    rng = np.random.default_rng(1)
    uv, comps = _small(rng)
    sky = synth.SkyModel(comps=comps, modType=0)
    plain = orc.dft_model_fp64(uv, sky).reshape(uv.nrow, -1)
    nant = 27
    jones = _jones(rng, 1, len(comps), 1, scale=0.0)
    m, live = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, np.zeros(nant, int),
                                  np.zeros(uv.nif * uv.nchan, int), jones)
    assert live.all()
    peak = np.abs(plain).max()
    assert np.abs(m[:, :, 0, 0] - plain).max() < 1e-9 * peak
    assert np.abs(m[:, :, 1, 1] - plain).max() < 1e-9 * peak
    assert np.abs(m[:, :, 0, 1]).max() == 0 and np.abs(m[:, :, 1, 0]).max() == 0


def _gpu_expansion(stok, g, h, c, s):
    """The reference kernels' products, term by term (include/CUDALinCorSum.h:
    I :119-126, Q :226-240, U :345-359, V :444-452); g, h the Jones matrices
    as 4 complex (XX, XY, YX, YY)."""
    mc = lambda a, b: a * np.conj(b)
    if stok == 1:
        return [mc(g[0], h[0]) + mc(g[1], h[1]), mc(g[1], h[3]) + mc(g[0], h[2]),
                mc(g[2], h[0]) + mc(g[3], h[1]), mc(g[2], h[2]) + mc(g[3], h[3])]
    if stok == 2:
        f = lambda a, b, cc, d: c * (mc(g[a], h[b]) - mc(g[cc], h[d]))
        t = lambda a, b, cc, d: s * (mc(g[a], h[b]) + mc(g[cc], h[d]))
        return [f(0, 0, 1, 1) - t(1, 0, 0, 1), f(0, 2, 1, 3) - t(1, 2, 0, 3),
                f(2, 0, 3, 1) - t(3, 0, 2, 1), f(2, 2, 3, 3) - t(3, 2, 2, 3)]
    if stok == 3:
        f = lambda a, b, cc, d: c * (mc(g[a], h[b]) + mc(g[cc], h[d]))
        t = lambda a, b, cc, d: s * (mc(g[a], h[b]) - mc(g[cc], h[d]))
        return [f(1, 0, 0, 1) + t(0, 0, 1, 1), f(1, 2, 0, 3) + t(0, 2, 1, 3),
                f(3, 0, 2, 1) + t(2, 0, 3, 1), f(3, 2, 2, 3) + t(2, 2, 3, 3)]
    J = 1j
    return [J * (mc(g[0], h[1]) - mc(g[1], h[3])), J * (mc(g[0], h[3]) - mc(g[1], h[2])),
            J * (mc(g[2], h[1]) - mc(g[3], h[0])), J * (mc(g[2], h[3]) - mc(g[3], h[2]))]


@pytest.mark.parametrize("stok", [1, 2, 3, 4])
def test_products_match_the_reference_expansion(stok):
    """J1 P J2^H (matrix products) against the GPU's expanded terms; for
    Stokes V the GPU's XX term uses J2's YY element where the product has
    J2's XX element, so XX is compared with that element exchanged."""
    rng = np.random.default_rng(10 + stok)
    jones = _jones(rng, 2, 5, 3, scale=0.4)
    chi = 0.37
    c, s = math.cos(2 * chi), math.sin(2 * chi)
    prods = orc.beam_products(jones, [(0, 1), (1, 0), (1, 1)], stok, chi)
    for (t1, t2), M in prods.items():
        g = jones[t1].reshape(5, 3, 4).transpose(2, 0, 1)
        h = jones[t2].reshape(5, 3, 4).transpose(2, 0, 1)
        ref = _gpu_expansion(stok, g, h, c, s)
        got = M.reshape(5, 3, 4).transpose(2, 0, 1)
        for e in range(4):
            if stok == 4 and e == 0:
                h_fix = h.copy()
                h_fix[3] = h[0]          # the product's term g_xy conj(h_xx)
                ref0 = 1j * (g[0] * np.conj(h_fix[1]) - g[1] * np.conj(h_fix[3]))
                assert np.abs(got[0] - ref0).max() < 1e-12
                assert np.abs(got[0] - ref[0]).max() > 1e-6   # the GPU term differs
            else:
                assert np.abs(got[e] - ref[e]).max() < 1e-12


def test_diagonal_jones_closed_form():
    """J = diag(gx, gy): XX = g1x conj(g2x) P_xx V, XY = g1x conj(g2y) P_xy V, ..."""
    rng = np.random.default_rng(3)
    jones = np.zeros((2, 4, 2, 2, 2), np.complex128)
    jones[..., 0, 0] = rng.normal(size=(2, 4, 2)) + 1j * rng.normal(size=(2, 4, 2))
    jones[..., 1, 1] = rng.normal(size=(2, 4, 2)) + 1j * rng.normal(size=(2, 4, 2))
    for stok in (1, 2, 3, 4):
        P = orc.beam_coherency(stok, 0.2)
        M = orc.beam_products(jones, [(0, 1)], stok, 0.2)[(0, 1)]
        g, h = jones[0], jones[1]
        for a in range(2):
            for b in range(2):
                want = g[..., a, a] * np.conj(h[..., b, b]) * P[a, b]
                assert np.abs(M[..., a, b] - want).max() < 1e-12


def test_q_rotates_into_u():
    """Q at chi - pi/4 is U at chi; I and V do not depend on chi."""
    for chi in (0.0, 0.4, -1.1):
        assert np.allclose(orc.beam_coherency(2, chi - math.pi / 4), orc.beam_coherency(3, chi))
        assert np.allclose(orc.beam_coherency(1, chi), np.eye(2))
        assert np.allclose(orc.beam_coherency(4, chi), orc.beam_coherency(4, 0.0))
        for s in (1, 2, 3, 4):
            P = orc.beam_coherency(s, chi)
            assert np.allclose(P, P.conj().T)       # Hermitian


def test_brute_force_tiny_case():
    """Explicit loops over rows, channels, components and matrix entries."""
    rng = np.random.default_rng(5)
    uv, comps = _small(rng, rows=3, ncomp=2)
    comps[:, 0] = [1.0, 0.5]
    gauss = np.zeros((2, 3), np.float32)
    gauss[1] = synth.gauss_coeffs(20 / 3600.0, 10 / 3600.0, 30.0)
    comps = np.concatenate([comps, gauss], axis=1).astype(np.float32)
    nant = 27
    at = rng.integers(0, 2, nant)
    nf = uv.nif * uv.nchan
    jmap = (np.arange(nf) * 3) // nf
    jones = _jones(rng, 2, 2, 3, scale=0.3)
    chi, stok = 0.6, 2
    m, live = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, at, jmap, jones, stok, chi)
    P = orc.beam_coherency(stok, chi)
    for r in range(3):
        cb = int(uv.buffer[r, 3])
        a1, a2 = cb // 256 - 1, cb % 256 - 1
        for f in range(0, nf, 7):
            fs = uv.freqArr[f] / uv.freq
            S = np.zeros((2, 2), complex)
            for k in range(2):
                u, v, w = (float(x) for x in uv.buffer[r, :3])
                x, y, z = (float(t) for t in comps[k, 1:4])
                g0, g1, g2 = (float(t) for t in comps[k, 4:7])
                amp = float(comps[k, 0]) * math.exp(fs * fs * (g0 * u * u + g1 * v * v + g2 * u * v))
                V = amp * complex(math.cos(fs * (u * x + v * y + w * z)), math.sin(fs * (u * x + v * y + w * z)))
                J1, J2 = jones[at[a1], k, jmap[f]], jones[at[a2], k, jmap[f]]
                for a in range(2):
                    for b in range(2):
                        for p in range(2):
                            for q in range(2):
                                S[a, b] += V * J1[a, p] * P[p, q] * np.conj(J2[b, q])
            assert np.abs(m[r, f] - S).max() < 1e-12 * max(1.0, np.abs(S).max())


def test_swapping_antennas_conjugates():
    """(a2, a1) with -u, -v, -w gives the conjugate transpose of (a1, a2)."""
    rng = np.random.default_rng(7)
    uv, comps = _small(rng, rows=16, ncomp=6)
    at = rng.integers(0, 2, 27)
    jones = _jones(rng, 2, 6, 1, scale=0.3)
    nf = uv.nif * uv.nchan
    jm = np.zeros(nf, int)
    b = uv.buffer.copy()
    cb = b[:, 3].astype(np.int64)
    a1, a2 = cb // 256, cb % 256
    b2 = b.copy()
    b2[:, :3] *= -1
    b2[:, 3] = (a2 * 256 + a1).astype(np.float32)
    for stok in (1, 2, 3, 4):
        m1, _ = orc.beam_model_fp64(b, comps, uv.freqArr / uv.freq, at, jm, jones, stok, 0.3)
        m2, _ = orc.beam_model_fp64(b2, comps, uv.freqArr / uv.freq, at, jm, jones, stok, 0.3)
        assert np.abs(m2 - np.conj(np.swapaxes(m1, -1, -2))).max() < 1e-9 * np.abs(m1).max()


def test_apply_rules():
    """Flags, replace, divide and the out-of-table rows (CPU rules)."""
    nrparm, nchan = 4, 2
    ncorr = 4
    incf = 12
    buf = np.zeros((3, nrparm + nchan * incf), np.float32)
    buf[:, 3] = [258.0, 259.0, 9 * 256 + 1]          # (1,2), (1,3), antenna 9 not in the table
    for c in range(nchan):
        o = nrparm + c * incf
        for k in range(4):
            buf[:, o + 3 * k] = 1.0 + k
            buf[:, o + 3 * k + 1] = 0.5
            buf[:, o + 3 * k + 2] = 1.0
    o1 = nrparm + incf
    buf[0, nrparm + 2] = -1.0                           # row 0, channel 0: XX flagged
    buf[1, o1 + 2] = buf[1, o1 + 5] = buf[1, o1 + 8] = buf[1, o1 + 11] = 0.0   # row 1 ch 1 all flagged
    model = np.full((3, nchan, 2, 2), 0.25 + 0.5j)
    live = np.array([True, True, False])
    sub = orc.beam_apply(buf, model, live, nrparm, 1, nchan, incf, 0, ncorr, "sub")
    assert sub[0, nrparm] == 1.0 and sub[0, nrparm + 3] == 2.0 - 0.25   # XX flagged kept, YY subtracted
    assert np.array_equal(sub[1, o1:o1 + 12], buf[1, o1:o1 + 12])       # all flagged: skipped
    assert np.array_equal(sub[2], buf[2])                                # out of the table
    assert sub[0, nrparm + 6] == 3.0 - 0.25 and sub[0, nrparm + 7] == 0.0  # XY subtracted
    rep = orc.beam_apply(buf, model, live, nrparm, 1, nchan, incf, 0, ncorr, "repl")
    assert rep[1, o1] == 0.25 and rep[1, o1 + 1] == 0.5                  # replaced though flagged
    div = orc.beam_apply(buf, model, live, nrparm, 1, nchan, incf, 0, ncorr, "div")
    assert np.array_equal(div[0, nrparm + 6:nrparm + 12], buf[0, nrparm + 6:nrparm + 12])  # cross hands kept
    mr, mi = np.float32(0.25), np.float32(0.5)
    wm = mr * mr + mi * mi
    re, im = np.float32(2.0), np.float32(0.5)
    assert div[0, nrparm + 3] == (mr / wm) * re + (mi / wm) * im
    assert div[0, nrparm + 5] == np.float32(1.0) * np.sqrt(wm)


def test_spectral_amplitude_helpers():
    """Tabulated spectra: the subband flux at its own frequency, a power law
    in between, SI = 0 flat; parameterised: one term is a power law."""
    from obit_amd.beam import pspec_amps, tspec_amps
    fs = np.array([0.9, 1.0, 1.1, 1.25])
    spec_index = np.array([0, 0, 1, 1])
    freq_rat = np.array([1.0 / 0.95, 1.0 / 1.2])        # nu0 / nu_sb
    flux = np.array([[2.0, 1.0], [0.5, 0.25]])
    si = np.array([[-0.7, 0.3], [0.0, 0.0]])
    a = tspec_amps(flux, si, spec_index, fs, freq_rat)
    for f in range(4):
        sb = spec_index[f]
        assert abs(a[0, f] - flux[0, sb] * (fs[f] / (1.0 / freq_rat[sb])) ** si[0, sb]) < 1e-6 * flux[0, sb]
        assert a[1, f] == np.float32(flux[1, sb])
    p = pspec_amps(np.array([3.0]), np.array([[0.8]]), fs, 1.0)
    assert np.allclose(p[0], 3.0 * fs ** 0.8, rtol=1e-6)
    p2 = pspec_amps(np.array([1.0]), np.array([[0.0, 0.5]]), fs, 1.0)
    assert np.allclose(p2[0], np.exp(-0.5 * np.log(fs) ** 2), rtol=1e-6)


def test_amplitude_table_replaces_the_column():
    """A table that repeats the column gives the same model; a table
    scaled per channel scales the model per channel."""
    rng = np.random.default_rng(8)
    uv, comps = _small(rng, rows=12, ncomp=5)
    nf = uv.nif * uv.nchan
    at = np.zeros(27, int)
    jm = np.zeros(nf, int)
    jones = _jones(rng, 1, 5, 1, scale=0.2)
    m0, _ = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, at, jm, jones, 1, 0.0)
    tab = np.repeat(comps[:, :1], nf, axis=1)
    m1, _ = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, at, jm, jones, 1, 0.0, amps=tab)
    assert np.abs(m1 - m0).max() < 1e-12 * np.abs(m0).max()
    scale = np.linspace(0.5, 2.0, nf)
    m2, _ = orc.beam_model_fp64(uv.buffer, comps, uv.freqArr / uv.freq, at, jm, jones, 1, 0.0,
                                amps=(tab * scale[None, :]).astype(np.float32))
    assert np.abs(m2 - m0 * scale[None, :, None, None]).max() < 1e-6 * np.abs(m0).max()
