"""ObitHIPStokesI against the oracle's restatement of ObitUVCalSelect
(Stokes 'I'): bit-identical on flagged random data in several layouts and
on Obit's UVSub rows."""
import os

import numpy as np
import pytest

from obit_amd import synth

pytestmark = pytest.mark.gpu
NR = 5


def _uv(buf, nif, nch, nst, k0):
    return synth.UVData(buffer=np.ascontiguousarray(buf, np.float32), nchan=nch, nif=nif, nstok=nst,
                        freq=1.4e9, freqArr=np.full(nif * nch, 1.4e9), crval_stokes=float(k0))


@pytest.mark.parametrize("nif,nch,nst,k0,formal", [
    (1, 64, 2, -1, False), (4, 64, 2, -1, True), (2, 37, 4, -1, False), (2, 37, 4, -5, True),
    (1, 200, 1, -1, False), (3, 16, 4, 1, False), (2, 8, 4, -2, False)])
def test_matches_oracle(oracle, nif, nch, nst, k0, formal):
    from obit_amd.stokes import stokes_i
    rng = np.random.default_rng(nif * 100 + nch + nst)
    n = 777
    buf = rng.normal(size=(n, NR + 3 * nst * nch * nif)).astype(np.float32)
    d = buf[:, NR:].reshape(n, nif, nch, nst, 3)
    d[..., 2] = rng.choice(np.array([-1.0, 0.0, 0.25, 1.0, 3.0, np.nan], np.float32), size=d[..., 2].shape,
                           p=[0.1, 0.1, 0.2, 0.3, 0.29, 0.01])
    uv = _uv(buf, nif, nch, nst, k0)
    got = stokes_i(uv, formal_i=formal).buffer
    ref = oracle.stokes_i(buf, NR, nif, nch, nst, 3, 3 * nst, 3 * nst * nch, k0, formal)
    print(f"layout nif={nif} nchan={nch} nstok={nst} kstoke0={k0} formal={formal}", flush=True)
    np.testing.assert_array_equal(got, ref)


def test_uvsub_rows(oracle):
    from obit_amd.stokes import stokes_i
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "uvsub_fixture.npz"))
    rows, nif, nst = fx["rows_in"], int(fx["nif"]), int(fx["nstok"])
    got = stokes_i(_uv(rows, nif, 1, nst, -1)).buffer
    ref = oracle.stokes_i(rows, NR, nif, 1, nst, 3, 3 * nst, 3 * nst, -1)
    assert np.array_equal(got, ref)


def test_refuses_missing_stokes():
    from obit_amd import _lib
    from obit_amd.stokes import stokes_i
    buf = np.ones((4, NR + 3 * 4), np.float32)
    with pytest.raises(_lib.HipError, match="not available"):
        stokes_i(_uv(buf, 1, 1, 4, 2))
    with pytest.raises(_lib.HipError, match="not available"):
        stokes_i(_uv(buf[:, :NR + 3], 1, 1, 1, -1), formal_i=True)
