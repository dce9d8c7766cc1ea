"""Obit's uv FITS tables through obit_amd.uvtab: a synthetic data set
written and read back unchanged; Obit's own UVSub input (when the
reference tree is present) read into the records the fixture holds, with
its frequencies, and rewritten so that an independent reader (the fixture
script's) sees the same rows."""
import os
import sys

import numpy as np
import pytest

from obit_amd import synth, uvtab

REF = os.path.join("/root/reference", "ObitSystem", "Obit", "testData", "UVSubTestIn.uvtab")
HERE = os.path.dirname(os.path.abspath(__file__))


def test_synthetic_round_trip(tmp_path):
    uv, _, _ = synth.workload(1, rows=300)
    p = str(tmp_path / "w1.uvtab")
    uvtab.write(p, uv, uvtab.table_for(uv, jd_obs=2460000.5, ra=150.0, dec=30.0))
    back, t = uvtab.read(p)
    assert np.array_equal(back.buffer, uv.buffer)
    assert (back.nchan, back.nif, back.nstok, back.nrparm) == (uv.nchan, uv.nif, uv.nstok, uv.nrparm)
    assert back.crval_stokes == uv.crval_stokes and back.freq == uv.freq
    assert (back.ilocu, back.ilocv, back.ilocw, back.ilocb, back.iloct) == (0, 1, 2, 3, 4)
    np.testing.assert_allclose(back.freqArr, uv.freqArr, rtol=1e-12)
    assert t.jd_obs == 2460000.5
    # the file is FITS: 2880-byte blocks, string values from column 11
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0
    assert b"XTENSION= 'BINTABLE'" in raw


@pytest.mark.skipif(not os.path.exists(REF), reason="reference tree not present")
def test_uvsub_input(tmp_path):
    fx = np.load(os.path.join(HERE, "golden", "uvsub_fixture.npz"))
    uv, t = uvtab.read(REF)
    assert (uv.nif, uv.nchan, uv.nstok, uv.nrparm) == (int(fx["nif"]), 1, int(fx["nstok"]), 5)
    assert uv.crval_stokes == -1.0 and uv.freq == float(fx["f0"])
    np.testing.assert_array_equal(uv.freqArr, fx["freqs"])
    rows = {r.tobytes() for r in uv.buffer}
    assert all(r.tobytes() in rows for r in fx["rows_in"])
    p = str(tmp_path / "copy.uvtab")
    uvtab.write(p, uv, t)
    back, _ = uvtab.read(p)
    assert np.array_equal(back.buffer, uv.buffer)
    # every HDU but the UV table is copied byte for byte; an independent
    # reader sees the same rows
    a, b = open(REF, "rb").read(), open(p, "rb").read()
    assert len(a) == len(b)
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_uvsub_fixture as fxs
    assert np.array_equal(fxs.read_uvtab(p)["rows"], uv.buffer)
