"""The oracle's Jones-from-beam-images restatement (oracle.beam_lagrange,
beam_interp, beam_pixel, inv_beam, beam_jones), pinned by what the
mathematics fixes: Lagrange weights reproduce polynomials, interpolation
returns pixel values at pixel centres, the projections against an
independent direction-cosine derivation, the rotation, InvBeam's closed
form, and the Jones matrix of constant beams."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

FBLANK = -1.0e30


@pytest.mark.parametrize("hw", [1, 2, 3, 4])
def test_lagrange_reproduces_polynomials(hw):
    n = 40
    for p in (2.3, 7.5, 17.91, 1.2, 39.7):
        cen, k = orc.beam_lagrange(p, n, hw)
        nodes = cen + 1 + np.arange(2 * hw + 1, dtype=np.float64)
        assert abs(k.astype(np.float64).sum() - 1.0) < 1e-5
        for deg in range(2 * hw + 1):
            f = lambda x: (x / 10.0) ** deg
            assert abs(np.dot(k.astype(np.float64), f(nodes)) - f(p)) < 2e-5 * max(1.0, f(p))


def test_interp_polynomial_image_and_pixel_centres():
    ny, nx = 30, 34
    y, x = np.mgrid[1:ny + 1, 1:nx + 1].astype(np.float64)
    f = lambda x, y: 0.3 + 0.01 * x - 0.02 * y + 1e-3 * x * y - 2e-4 * x * x + 1e-5 * y ** 3
    img = f(x, y).astype(np.float32)
    for px, py in ((5.3, 7.8), (17.5, 12.25), (2.1, 28.9), (33.6, 1.4)):
        assert abs(orc.beam_interp(img, px, py, 2, FBLANK) - f(px, py)) < 2e-5
    assert orc.beam_interp(img, 11.0, 13.0, 2, FBLANK) == img[12, 10]
    assert orc.beam_interp(img, 0.99, 5.0, 2, FBLANK) == np.float32(FBLANK)
    assert orc.beam_interp(img, 5.0, ny + 0.01, 2, FBLANK) == np.float32(FBLANK)


def test_interp_blanks():
    img = np.full((20, 20), 0.7, np.float32)
    img[9, 9] = FBLANK                                   # pixel (10, 10)
    assert orc.beam_interp(img, 10.0, 10.0, 2, FBLANK) == np.float32(FBLANK)   # its own weight is 1
    assert orc.beam_interp(img, 10.45, 10.0, 2, FBLANK) == np.float32(FBLANK)   # blank weight ~0.6
    assert abs(orc.beam_interp(img, 11.9, 10.0, 2, FBLANK) - 0.7) < 1e-6         # small blank weight
    assert abs(orc.beam_interp(img, 15.2, 15.3, 2, FBLANK) - 0.7) < 1e-6        # blank outside the kernel


def _dircos(ra, dec, ra0, dec0):
    """(l, m, n) of (ra, dec) about (ra0, dec0) from 3-vectors (radians)."""
    p = np.array([math.cos(dec) * math.cos(ra), math.cos(dec) * math.sin(ra), math.sin(dec)])
    el = np.array([-math.sin(ra0), math.cos(ra0), 0.0])
    em = np.array([-math.sin(dec0) * math.cos(ra0), -math.sin(dec0) * math.sin(ra0), math.cos(dec0)])
    en = np.array([math.cos(dec0) * math.cos(ra0), math.cos(dec0) * math.sin(ra0), math.sin(dec0)])
    return p @ el, p @ em, p @ en


@pytest.mark.parametrize("proj", [1, 2, 3])
def test_projections_against_direction_cosines(proj):
    d = dict(crval=(30.0, -45.0), crpix=(65.0, 70.0), cdelt=(-0.05, 0.05), proj=proj)
    for (x, y) in ((30.8, -44.1), (29.0, -46.5), (31.7, -43.2)):
        px, py = orc.beam_pixel(d, x, y, 0.0)
        l, m, n = _dircos(math.radians(x), math.radians(y), math.radians(30.0), math.radians(-45.0))
        if proj == 2:
            l, m = l / n, m / n
        elif proj == 3:
            t = math.acos(n)
            l, m = l * t / math.sin(t), m * t / math.sin(t)
        want = (math.degrees(l) / -0.05 + 65.0, math.degrees(m) / 0.05 + 70.0)
        assert abs(px - want[0]) < 2e-3 and abs(py - want[1]) < 2e-3


def test_linear_pixel_and_rotation():
    d = dict(crval=(0.0, 0.0), crpix=(33.0, 33.0), cdelt=(0.1, 0.1), proj=0)
    px, py = orc.beam_pixel(d, 0.3, -0.2, 0.0)
    assert abs(px - 36.0) < 1e-5 and abs(py - 31.0) < 1e-5
    px, py = orc.beam_pixel(d, 0.3, -0.2, math.pi / 2)          # (x, y) -> (y, -x)
    assert abs(px - 31.0) < 1e-5 and abs(py - 30.0) < 1e-5
    assert orc.beam_pixel(dict(d, cdelt=(0.0, 0.1)), 0.3, 0.2, 0.0) is None
    # RA wraps by 360 degrees
    d2 = dict(d, crval=(359.9, 0.0))
    px, _ = orc.beam_pixel(d2, 0.1, 0.0, 0.0)
    assert abs(px - 35.0) < 1e-4


def test_inv_beam_closed_form():
    assert orc.inv_beam(1.0, 1.0, 0.0) == 1.0
    assert orc.inv_beam(2.0, 1.0, 0.52) == 100.0
    assert abs(orc.inv_beam(0.5, 1.0, 0.5) - 0.75 * math.sqrt(2.0)) < 1e-6


def test_jones_of_constant_beams():
    """Constant images: each element is ipb x its constant; off the image
    (or with P1 beyond 1.3 x the gain) the unit matrix."""
    nch, ny, nx = 2, 16, 16
    consts = [0.9, 0.05, 0.8, -0.04, 0.1, 0.02, -0.1, 0.03]
    imgs = np.stack([np.full((nch, ny, nx), c, np.float32) for c in consts])
    d = dict(crval=(0.0, 0.0), crpix=(8.5, 8.5), cdelt=(0.5, 0.5), proj=0, hwidth=2)
    bf = np.array([1.0e9, 1.5e9])
    fact = 1.0e-10
    pos = np.array([[0.5, 1.0], [40.0, 0.0], [-1.0, -2.0]], np.float32)
    J = orc.beam_jones([d], [imgs], bf, fact, FBLANK, pos, 0.2)
    for k in (0, 2):
        ang = math.hypot(pos[k, 0], pos[k, 1])
        for c in range(nch):
            ipb = orc.inv_beam(fact, bf[c], ang)
            want = ipb * np.array([[consts[0] + 1j * consts[1], consts[4] + 1j * consts[5]],
                                   [consts[6] + 1j * consts[7], consts[2] + 1j * consts[3]]])
            assert np.abs(J[0, k, c] - want).max() < 1e-5
    assert np.array_equal(J[0, 1], np.broadcast_to(np.eye(2), (nch, 2, 2)))
