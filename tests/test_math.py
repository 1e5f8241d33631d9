"""Accuracy of the shared reproducible math (hellovulkan_amd/csrc/hvx_math.h)
against numpy's float64 libm, through the oracle library (CPU only)."""
import numpy as np
import pytest

from oracle import hvx_oracle as O


def ulp_err(got, ref):
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return np.abs(np.asarray(got, np.float64) - ref) / sp


RNG = np.random.default_rng(1234)


def test_log2():
    x = np.exp(RNG.uniform(-85, 85, 100000)).astype(np.float32)
    assert ulp_err(O.math_f32("log2", x), np.log2(x.astype(np.float64))).max() < 3.0
    x = (1 + RNG.uniform(-0.01, 0.01, 100000)).astype(np.float32)
    assert ulp_err(O.math_f32("log2", x), np.log2(x.astype(np.float64))).max() < 3.0
    sp = O.math_f32("log2", np.array([0, -1, np.inf, 1, 2, 0.5, 1e-40, np.nan], np.float32))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and sp[3] == 0 and sp[4] == 1 and sp[5] == -1
    assert abs(sp[6] - np.log2(1e-40)) < 1e-4 and np.isnan(sp[7])


def test_exp2_and_pow():
    x = RNG.uniform(-140, 127.9, 100000).astype(np.float32)
    assert ulp_err(O.math_f32("exp2", x), np.exp2(x.astype(np.float64))).max() < 1.1
    sp = O.math_f32("exp2", np.array([0, 1, -1, 128, -151, 10], np.float32))
    assert list(sp) == [1, 2, 0.5, np.inf, 0, 1024]
    d = RNG.uniform(0.01, 40, 100000).astype(np.float32)
    assert ulp_err(O.math_f32("pow", d, np.full_like(d, 0.1)), d.astype(np.float64) ** 0.1).max() < 2.0
    y = (np.arange(49, dtype=np.float32) / np.float32(48))
    assert ulp_err(O.math_f32("pow", np.full(49, 1000, np.float32), y), 1000.0 ** y.astype(np.float64)).max() < 8.0
    sp = O.math_f32("pow", np.array([0, 0, 2, 1, -2], np.float32), np.array([2, 0, 0, 5, 2], np.float32))
    assert sp[0] == 0 and sp[1] == 1 and sp[2] == 1 and sp[3] == 1 and np.isnan(sp[4])
    x = RNG.uniform(0, 1, 100000).astype(np.float32)
    assert ulp_err(O.math_f32("pow5", x), x.astype(np.float64) ** 5).max() < 3.0


def test_trig_correctly_rounded():
    # the double-evaluated functions round once: within half an ulp (+ tiny)
    x = RNG.uniform(-10, 10, 100000).astype(np.float32)
    assert ulp_err(O.math_f32("sin", x), np.sin(x.astype(np.float64))).max() <= 0.5001
    assert ulp_err(O.math_f32("cos", x), np.cos(x.astype(np.float64))).max() <= 0.5001
    y, xx = RNG.normal(size=(2, 100000)).astype(np.float32)
    assert ulp_err(O.math_f32("atan2", y, xx), np.arctan2(y.astype(np.float64), xx.astype(np.float64))).max() <= 0.5001
    c = RNG.uniform(-1, 1, 100000).astype(np.float32)
    assert ulp_err(O.math_f32("acos", c), np.arccos(c.astype(np.float64))).max() <= 0.5001


def test_double_kernels():
    x = RNG.uniform(-50, 50, 100000)
    assert np.max(np.abs(O.math_f64("sin", x) - np.sin(x))) < 4e-15
    assert np.max(np.abs(O.math_f64("cos", x) - np.cos(x))) < 4e-15
    t = np.concatenate([RNG.normal(size=50000) * 10, RNG.uniform(-1, 1, 50000)])
    assert np.max(np.abs(O.math_f64("atan", t) - np.arctan(t))) < 4e-16 * 8
    # C99 atan2 special cases
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0])
    x = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -0.0])
    got = O.math_f64("atan2", y, x)
    np.testing.assert_array_equal(got, np.arctan2(y, x))
    assert np.signbit(got[1]) and np.signbit(got[3])


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_trig_identities_exact_points(fn):
    pts = np.array([0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi], np.float32)
    got = O.math_f32(fn, pts)
    ref = getattr(np, fn)(pts.astype(np.float64))
    assert np.all(np.abs(got - ref) <= np.spacing(np.float32(1.0)))
