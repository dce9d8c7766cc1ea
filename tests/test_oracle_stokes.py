"""The oracle's Stokes I formation (ObitUVCalSelect with Stokes 'I',
src/ObitUVCalSelect.c:540-760) pinned by closed forms: the mean of the two
hands with half the weight sum; a flagged hand replaced by the other at
half its weight; nothing when both are flagged or when formal I lacks a
hand; a single correlation or true Stokes I copied; the cases Obit refuses.
"""
import numpy as np
import pytest

from obit_amd import synth

NR = 5


def _rows(corrs, nstok):
    """One row per entry of corrs: [(re, im, wt) per Stokes]."""
    out = np.zeros((len(corrs), NR + 3 * nstok), np.float32)
    out[:, :NR] = np.arange(NR, dtype=np.float32) + 10
    for i, c in enumerate(corrs):
        for s, v in enumerate(c):
            out[i, NR + 3 * s:NR + 3 * s + 3] = v
    return out


def _si(oracle, rows, nstok, kstoke0, formal=False):
    return oracle.stokes_i(rows, NR, 1, 1, nstok, 3, 3 * nstok, 3 * nstok, kstoke0, formal)


def test_closed_forms(oracle):
    rr, ll = (1.5, -2.0, 4.0), (0.5, 6.0, 2.0)
    rows = _rows([[rr, ll], [rr, (9.0, 9.0, 0.0)], [(9.0, 9.0, -1.0), ll],
                  [(9.0, 9.0, 0.0), (9.0, 9.0, -3.0)]], 2)
    out = _si(oracle, rows, 2, -1)
    assert np.array_equal(out[:, :NR], rows[:, :NR])
    np.testing.assert_array_equal(out[0, NR:], [1.0, 2.0, 3.0])   # mean, (4 + 2) / 2
    np.testing.assert_array_equal(out[1, NR:], [1.5, -2.0, 2.0])  # RR alone, half its weight
    np.testing.assert_array_equal(out[2, NR:], [0.5, 6.0, 1.0])   # LL alone
    np.testing.assert_array_equal(out[3, NR:], [0.0, 0.0, 0.0])   # both flagged
    # formal I needs both hands
    f = _si(oracle, rows, 2, -1, formal=True)
    np.testing.assert_array_equal(f[0, NR:], [1.0, 2.0, 3.0])
    assert not f[1:, NR:].any()
    # XX, YY form I the same way
    np.testing.assert_array_equal(_si(oracle, rows, 2, -5), out)
    # one correlation, and true Stokes I, are copied with their weight
    one = _rows([[rr], [(9.0, 9.0, 0.0)]], 1)
    np.testing.assert_array_equal(_si(oracle, one, 1, -1)[:, NR:], [[1.5, -2.0, 4.0], [0, 0, 0]])
    iq = _rows([[rr, ll]], 2)
    np.testing.assert_array_equal(_si(oracle, iq, 2, 1)[0, NR:], [1.5, -2.0, 4.0])


def test_unavailable(oracle):
    rows = _rows([[(1, 1, 1)] * 4], 4)
    for k0, nst, formal in ((2, 4, False), (-3, 4, False), (-7, 4, False), (-1, 1, True),
                            (-2, 2, True), (-6, 2, True)):
        with pytest.raises(ValueError):
            _si(oracle, rows[:, :NR + 3 * nst], nst, k0, formal)
    # LL-first data still give (non-formal) I from the first two correlations
    _si(oracle, rows, 4, -2)


def test_symmetric_and_linear(oracle):
    rng = np.random.default_rng(5)
    n, nif, nch, nst = 300, 2, 8, 4
    buf = rng.normal(size=(n, NR + 3 * nst * nch * nif)).astype(np.float32)
    d = buf[:, NR:].reshape(n, nif, nch, nst, 3)
    d[..., 2] = rng.choice([-1.0, 0.0, 0.5, 2.0], size=d[..., 2].shape)
    a = oracle.stokes_i(buf, NR, nif, nch, nst, 3, 3 * nst, 3 * nst * nch, -1)
    sw = buf.copy()
    s = sw[:, NR:].reshape(n, nif, nch, nst, 3)
    s[..., 0, :], s[..., 1, :] = d[..., 1, :].copy(), d[..., 0, :].copy()
    np.testing.assert_array_equal(oracle.stokes_i(sw, NR, nif, nch, nst, 3, 3 * nst, 3 * nst * nch, -1), a)
    twice = buf.copy()
    twice[:, NR:] *= 2
    b = oracle.stokes_i(twice, NR, nif, nch, nst, 3, 3 * nst, 3 * nst * nch, -1)
    np.testing.assert_array_equal(b[:, NR:], 2 * a[:, NR:])


def test_fixture_rows(oracle):
    """Obit's UVSub input (2 IF x RR, LL, RL, LR): every sample falls in
    one of the closed-form cases."""
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "uvsub_fixture.npz"))
    rows = fx["rows_in"]
    nif, nst = int(fx["nif"]), int(fx["nstok"])
    out = oracle.stokes_i(rows, NR, nif, 1, nst, 3, 3 * nst, 3 * nst, -1)
    d = rows[:, NR:].reshape(-1, nif, nst, 3).astype(np.float64)
    o = out[:, NR:].reshape(-1, nif, 3)
    both = (d[..., 0, 2] > 0) & (d[..., 1, 2] > 0)
    assert both.sum() > 1000
    np.testing.assert_allclose(o[both][:, 0], 0.5 * (d[..., 0, 0] + d[..., 1, 0])[both], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(o[both][:, 2], 0.5 * (d[..., 0, 2] + d[..., 1, 2])[both], rtol=1e-6)
    none = (d[..., 0, 2] <= 0) & (d[..., 1, 2] <= 0)
    assert not o[none].any()
