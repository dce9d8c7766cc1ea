"""The oracle's restatement of Obit's FP32 vector math (SSE Cephes sincos_ps
and exp_ps, include/sse_mathfun.h) checked against libm in double, known
values and the properties Cephes guarantees."""
import math

import numpy as np


def test_sincos_known_values(oracle):
    for x, s_ref, c_ref in ((0.0, 0.0, 1.0), (math.pi / 6, 0.5, math.sqrt(3) / 2),
                            (math.pi / 4, math.sqrt(0.5), math.sqrt(0.5)),
                            (math.pi / 2, 1.0, 0.0), (math.pi, 0.0, -1.0), (-math.pi / 3, -math.sqrt(3) / 2, 0.5)):
        s, c = oracle.sincos(np.float32(x))
        assert abs(s - s_ref) < 2e-7 and abs(c - c_ref) < 2e-7, (x, s, c)


def test_sincos_error_bound(oracle):
    """Cephes sinf/cosf: |err| < ~1.5 ulp of 1 for |x| < 8192 after the
    three-part Cody-Waite reduction (breaks down beyond that range)."""
    rng = np.random.default_rng(1)
    x = rng.uniform(-8192, 8192, 200000).astype(np.float32)
    s, c = oracle.sincos_vec(x)
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(xd)).max() < 1.5e-7
    assert np.abs(c - np.cos(xd)).max() < 1.5e-7
    # odd/even symmetry and Pythagoras
    s2, c2 = oracle.sincos_vec(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 5e-7


def test_sincos_octant_boundaries(oracle):
    """Values either side of every octant boundary up to 64 pi: a wrong
    octant selection or sign bit shows up as an O(1) error."""
    k = np.arange(1, 256)
    x = (k * math.pi / 4)
    for eps in (-1e-4, 1e-4):
        xs = (x + eps).astype(np.float32)
        s, c = oracle.sincos_vec(xs)
        assert np.abs(s - np.sin(xs.astype(np.float64))).max() < 2e-7
        assert np.abs(c - np.cos(xs.astype(np.float64))).max() < 2e-7


def test_exp(oracle):
    x = np.linspace(-87, 88, 20001).astype(np.float32)
    y = oracle.exp_vec(x)
    rel = np.abs(y / np.exp(x.astype(np.float64)) - 1)
    assert rel.max() < 2e-7
    assert oracle.exp_vec(np.array([0.0], np.float32))[0] == 1.0
    # clamping at +-88.376 (exp_hi / exp_lo)
    big = oracle.exp_vec(np.array([200.0, -200.0], np.float32))
    assert np.isfinite(big[0]) and big[0] > 1e38 and 0 <= big[1] < 1e-37


def test_single_element_uses_libm(oracle):
    """ObitSinCosVec sends n == 1 to sinf/cosf (include/ObitSinCos.h:59)."""
    x = np.array([1234.5678], np.float32)
    s, c = oracle.sincos_vec(x)
    assert abs(s[0] - math.sin(float(x[0]))) < 1e-6
    s4, _ = oracle.sincos_vec(np.repeat(x, 4))
    assert abs(s4[0] - s[0]) < 3e-7  # Cephes and libm agree to ~1 ulp
