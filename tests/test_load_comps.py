"""The component-table builder: ObitHostCCMergeSel (ObitTableCCUtilMergeSel)
and ObitHostLoadComps (ObitSkyModelLoadComps), host code, checked against
results fixed independently of the restatement: merged duplicates summing
their fluxes in the order the reference documents, the textbook Fourier
transform of an elliptical Gaussian (by numerical integration of the image),
the 3-D path on a 2-D field that needs no re-projection, and the 2-D w term
as the first-order expansion of 2 pi (n - 1) about the field centre."""
import numpy as np
import pytest

from obit_amd.skymodel import cc_merge, load_comps

DEG = np.pi / 180.0


def _rows(entries, nparms=4, z=None):
    """CC rows [Flux, DeltaX, DeltaY, (DeltaZ), parms] from (flux, x, y, parms)."""
    out = []
    for i, (f, x, y, p) in enumerate(entries):
        r = [f, x, y] + ([z[i]] if z is not None else []) + list(p) + [0.0] * (nparms - len(p))
        out.append(r)
    return np.array(out, np.float32)


def test_merge_sums_duplicates_and_orders_by_flux():
    p = (0.0, 0.0, 0.0, 0.0)
    rows = _rows([(0.5, 1.0, 1.0, p), (0.2, 2.0, 2.0, p), (0.3, 1.0, 1.0, p), (-0.9, 3.0, 3.0, p),
                  (0.1, 2.0, 2.0, p), (0.05, 4.0, 4.0, p)])
    status = np.array([1, 1, 1, 1, -1, 1], np.int32)
    m, parms = cc_merge(rows, no_parms=4, status=status)
    assert np.allclose(m, [[-0.9, 3.0, 3.0], [0.8, 1.0, 1.0], [0.2, 2.0, 2.0], [0.05, 4.0, 4.0]], rtol=1e-6, atol=0)
    # rows 2..5 only
    m, _ = cc_merge(rows, no_parms=4, status=status, start=2, end=5)
    assert np.allclose(m, [[-0.9, 3.0, 3.0], [0.3, 1.0, 1.0], [0.2, 2.0, 2.0]], rtol=1e-6, atol=0)
    # every row deselected
    m, _ = cc_merge(rows, no_parms=4, status=-np.ones(6, np.int32))
    assert m.shape == (0, 3)


def test_merge_ties_keep_position_order_and_parms_of_first_row():
    rows = _rows([(0.5, 2.0, 0.0, (3.0, 2.0, 45.0, 1.0)), (-0.5, 1.0, 0.0, (9.0, 9.0, 9.0, 1.0)),
                  (0.5, 1.0, -1.0, (0.0, 0.0, 0.0, 1.0))])
    m, parms = cc_merge(rows, no_parms=4)
    # |flux| ties: position order, x then y
    assert m[:, 1:].tolist() == [[1.0, -1.0], [1.0, 0.0], [2.0, 0.0]]
    assert parms.tolist() == [3.0, 2.0, 45.0, 1.0]


def test_merge_spectra():
    # tabulated spectra (type 20): the terms of merged rows add up too
    t = [(0.0, 0.0, 0.0, 20.0, 1.0, 2.0, 3.0), (0.0, 0.0, 0.0, 20.0, 0.5, 0.5, 0.5)]
    rows = _rows([(1.0, 1.0, 1.0, t[0]), (2.0, 1.0, 1.0, t[1]), (0.5, 0.0, 1.0, t[1])], nparms=7)
    m, parms = cc_merge(rows, no_parms=7)
    assert np.allclose(m, [[3.0, 1.0, 1.0, 1.5, 2.5, 3.5], [0.5, 0.0, 1.0, 0.5, 0.5, 0.5]], rtol=1e-6, atol=0)
    assert parms[3] == 20.0
    # parameterised spectra (type 10) are not merged
    s = [(0.0, 0.0, 0.0, 10.0, -0.7), (0.0, 0.0, 0.0, 10.0, -0.2)]
    rows = _rows([(1.0, 1.0, 1.0, s[0]), (2.0, 1.0, 1.0, s[1])], nparms=5)
    m, _ = cc_merge(rows, no_parms=5)
    assert np.allclose(m, [[2.0, 1.0, 1.0, -0.2], [1.0, 1.0, 1.0, -0.7]], rtol=1e-6, atol=0)
    # a DeltaZ column is carried through
    rows = _rows([(1.0, 1.0, 1.0, (0, 0, 0, 0)), (2.0, 1.0, 1.0, (0, 0, 0, 0))], z=[0.25, 0.5])
    m, _ = cc_merge(rows, have_z=True, no_parms=4)
    assert np.allclose(m, [[3.0, 1.0, 1.0, 0.25]], rtol=1e-6, atol=0)


def _geom3d(crval=(180.0, 35.0), uv=(180.0, 35.0)):
    return dict(im_crval=crval, im_crota=0.0, do3D=1, xPxOff=0.0, yPxOff=0.0, uv_crval=uv, uv_crota=0.0,
                im_crpix=(0.0, 0.0), im_cdelt=(0.0, 0.0), im_naxis=(0, 0), xshift=0.0, yshift=0.0)


@pytest.mark.parametrize("pa", [0.0, 30.0, 117.0])
def test_gaussian_terms_are_the_textbook_transform(pa):
    """One elliptical Gaussian (FWHM 40" x 12") at the phase centre: the
    row's exp(g1 u^2 + g2 v^2 + g3 uv) equals the numerical Fourier
    transform of the image exp(-4 ln2 ((x'/M)^2 + (y'/m)^2)), x' along
    (cos PA, sin PA) in (l, m) -- the orientation of the reference's
    coefficients (src/ObitSkyModel.c:1939-1947) -- normalised by its flux."""
    maj, mnr = 40.0 / 3600.0, 12.0 / 3600.0
    rows = _rows([(1.0, 0.0, 0.0, (maj, mnr, pa, 1.0))])
    m, parms = cc_merge(rows, no_parms=4)
    c = load_comps(_geom3d(), m, parms, mod_type=1)
    assert c.shape == (1, 7)
    g1, g2, g3 = (float(v) for v in c[0, 4:7])
    # the image on a grid fine enough for |u| <= 3e4 wavelengths
    n, h = 401, 4.0 * maj * DEG
    ax = np.linspace(-h, h, n)
    L, M = np.meshgrid(ax, ax, indexing="ij")
    th = pa * DEG
    xp = L * np.cos(th) + M * np.sin(th)
    yp = -L * np.sin(th) + M * np.cos(th)
    img = np.exp(-4.0 * np.log(2.0) * ((xp / (maj * DEG)) ** 2 + (yp / (mnr * DEG)) ** 2))
    tot = img.sum()
    for u, v in [(1.5e4, 0.0), (0.0, 1.5e4), (1.0e4, 1.0e4), (-2.0e4, 7.0e3), (3.0e3, -2.5e4)]:
        num = (img * np.exp(-2j * np.pi * (u * L + v * M))).sum() / tot
        want = np.exp(g1 * u * u + g2 * v * v + g3 * u * v)
        assert abs(num.real - want) < 2e-5 and abs(num.imag) < 1e-9, (u, v, num, want)


def test_mixed_types_and_filters():
    gparm = (20.0 / 3600.0, 10.0 / 3600.0, 30.0, 1.0)
    rows = _rows([(1.0, 0.01, 0.02, gparm), (0.4, -0.01, 0.0, gparm), (-0.3, 0.0, 0.01, gparm),
                  (0.2, 0.02, 0.0, gparm)])
    m, parms = cc_merge(rows, no_parms=4)
    # a point field in a Gaussian table: zero Gaussian terms
    pt = np.array([0.0, 0.0, 0.0, 0.0], np.float32)
    c = load_comps(_geom3d(), m, pt, mod_type=1)
    assert c.shape == (4, 7) and np.all(c[:, 4:7] == 0.0)
    g = load_comps(_geom3d(), m, parms, mod_type=1)
    assert np.all(g[:, 4:7] == g[0, 4:7]) and g[0, 4] < 0.0 and g[0, 5] < 0.0 and g[0, 6] != 0.0
    assert np.array_equal(g[:, :4], c[:, :4])
    # noNeg stops at the first non-positive flux (order: 1.0, 0.4, -0.3, 0.2)
    c = load_comps(_geom3d(), m, pt, mod_type=0, no_neg=True)
    assert c[:, 0].tolist() == pytest.approx([1.0, 0.4])
    # minFlux drops |flux| at or below it; the factor scales the rest
    c = load_comps(_geom3d(), m, pt, mod_type=0, min_flux=0.3, factor=2.0)
    assert c[:, 0].tolist() == pytest.approx([2.0, 0.8])
    # spectral terms: copied for a spectral field, zero for a point field
    srows = _rows([(1.0, 0.01, 0.0, (0, 0, 0, 10.0, -0.7, 0.1)), (0.5, 0.0, 0.01, (0, 0, 0, 10.0, -0.2, 0.3))],
                  nparms=6)
    sm, sp = cc_merge(srows, no_parms=6)
    s = load_comps(_geom3d(), sm, sp, nterms=2, mod_type=10, n_spec_term=2)
    assert np.allclose(s[:, 4:], [[-0.7, 0.1], [-0.2, 0.3]], rtol=1e-6, atol=0)
    s = load_comps(_geom3d(), sm, pt, nterms=2, mod_type=10, n_spec_term=2)
    assert np.all(s[:, 4:] == 0.0)
    # a Gaussian field in a point-spectrum table: refused as the reference does
    with pytest.raises(ValueError):
        load_comps(_geom3d(), m, parms, mod_type=10, n_spec_term=2)


def test_2d_field_without_offsets_is_the_3d_table():
    rows = _rows([(1.0, 0.05, -0.03, (0, 0, 0, 0)), (0.5, -0.2, 0.1, (0, 0, 0, 0))])
    m, parms = cc_merge(rows, no_parms=4)
    c3 = load_comps(_geom3d(), m, parms, mod_type=0)
    g2 = dict(_geom3d(), do3D=0, im_crpix=(513.0, 513.0), im_cdelt=(-1.0 / 3600.0, 1.0 / 3600.0),
              im_naxis=(1024, 1024))
    c2 = load_comps(g2, m, parms, mod_type=0)
    assert np.array_equal(c2, c3)
    assert np.all(c3[:, 3] == 0.0)


def test_2d_field_w_term_is_the_expansion_about_the_centre():
    """A 2-D facet on the phase centre's tangent plane whose centre pixel
    lies (300, -200) pixels from its reference pixel, with reference-pixel
    offsets (12.5, -7.25): x, y are 2 pi (Delta + offset) and z the
    first-order term of 2 pi (n - 1) about the facet centre's direction
    cosines (l_c, m_c) = ((1 + n/2 - crpix) cdelt): -(l_c x + m_c y) / n_c."""
    cdelt = (-2.0 / 3600.0, 2.0 / 3600.0)
    crpix = (213.0, 713.0)
    naxis = (1024, 1024)
    g2 = dict(_geom3d(), do3D=0, xPxOff=12.5, yPxOff=-7.25, im_crpix=crpix, im_cdelt=cdelt, im_naxis=naxis)
    ent = [(1.0, 0.01, -0.02, (0, 0, 0, 0)), (0.7, -0.05, 0.03, (0, 0, 0, 0)), (0.3, 0.2, 0.15, (0, 0, 0, 0))]
    m, parms = cc_merge(_rows(ent), no_parms=4)
    c = load_comps(g2, m, parms, mod_type=0).astype(np.float64)
    lc = (1.0 + naxis[0] / 2 - crpix[0]) * cdelt[0] * DEG
    mc = (1.0 + naxis[1] / 2 - crpix[1]) * cdelt[1] * DEG
    nc = np.sqrt(1.0 - lc * lc - mc * mc)
    x = 2.0 * np.pi * DEG * (m[:, 1].astype(np.float64) + 12.5 * cdelt[0])
    y = 2.0 * np.pi * DEG * (m[:, 2].astype(np.float64) - 7.25 * cdelt[1])
    assert np.allclose(c[:, 1], x, rtol=1e-6, atol=0)
    assert np.allclose(c[:, 2], y, rtol=1e-6, atol=0)
    z = -(lc * x + mc * y) / nc
    assert np.allclose(c[:, 3], z, rtol=2e-5, atol=1e-12)
    assert np.abs(z).max() > 1e-5                      # the term is there


def test_general_builder_reproduces_the_golden_point_table():
    """On the golden UVSub field (a 3-D image off the phase centre, rotated,
    tests/golden/uvsub_fixture.npz) the general builder's point table is
    the point builder's, bit for bit (that table is pinned against the
    textbook geometry and Obit's master in test_golden_uvsub)."""
    import os
    from obit_amd.skymodel import load_comps_point
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uvsub_fixture.npz"))
    cc = np.ascontiguousarray(fx["cc"][:, :3], np.float32)
    want = load_comps_point(cc, fx["img_crval"], fx["img_crota"], fx["uv_centre"], fx["uv_crota"], 0.5)
    geom = dict(_geom3d(crval=tuple(fx["img_crval"]), uv=tuple(fx["uv_centre"])), im_crota=float(fx["img_crota"]),
                uv_crota=float(fx["uv_crota"]))
    got = load_comps(geom, cc, np.zeros(4, np.float32), mod_type=0, factor=0.5)
    assert np.array_equal(got, want)
