"""Accuracy of the portable FP32 transcendentals (ap_fmath.h) that the core
and the oracle share, against float64 numpy/scipy. Being bit-identical on
host and device makes the render path reproducible; these tests pin that
they are also accurate, so a wrong coefficient, split constant or quadrant
fails here rather than as a silently biased image."""
import numpy as np
import pytest
from scipy import special

import oracle_lib as ol


def ulp_err(got, ref):
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / sp


RNG = np.random.default_rng(0)
CASES = {
    "sin": (np.concatenate([np.linspace(-7, 7, 400001), RNG.uniform(-500, 500, 200000)]), np.sin, 1.6),
    "cos": (np.concatenate([np.linspace(-7, 7, 400001), RNG.uniform(-500, 500, 200000)]), np.cos, 1.6),
    "exp": (np.concatenate([np.linspace(-87, 88, 400001), np.linspace(-103, -87, 20001)]), np.exp, 1.1),
    "log": (np.concatenate([np.exp(np.linspace(-87, 88, 400001)), RNG.uniform(0.5, 2, 200000), [1e-40, 3e-39]]),
            np.log, 1.1),
    "log1p": (np.concatenate([-RNG.uniform(0, 0.9999999, 200000), RNG.uniform(0, 10, 20000)]), np.log1p, 2.5),
    "pow075": (RNG.uniform(0, 1, 200000), lambda x: x ** 0.75, 2.5),
    "atan": (np.concatenate([np.linspace(-50, 50, 200001), RNG.standard_cauchy(100000)]), np.arctan, 3.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_ulp_accuracy(name):
    xs, f, bound = CASES[name]
    xs = np.asarray(xs, np.float32)
    got = ol.fmath(name, xs)
    err = ulp_err(got, f(xs.astype(np.float64)))
    assert err.max() <= bound, (name, float(err.max()), float(xs[np.argmax(err)]))


def test_erfc_relative_accuracy():
    xs = np.linspace(-4, 6, 200001).astype(np.float32)
    got = ol.fmath("erfc", xs)
    ref = special.erfc(xs.astype(np.float64))
    assert (np.abs(got - ref) / ref).max() < 3e-6


def test_special_values():
    nan = np.float32("nan")
    assert np.isnan(ol.fmath("sin", [np.inf, nan])).all() and np.isnan(ol.fmath("cos", [-np.inf])).all()
    assert ol.fmath("sin", [0.0])[0] == 0.0 and ol.fmath("cos", [0.0])[0] == 1.0
    e = ol.fmath("exp", [0.0, 89.0, -104.0, -np.inf, np.inf])
    assert e[0] == 1.0 and np.isinf(e[1]) and e[2] == 0.0 and e[3] == 0.0 and np.isinf(e[4])
    assert 0.0 < ol.fmath("exp", [-100.0])[0] < 1e-38  # subnormal results are kept
    lg = ol.fmath("log", [1.0, 0.0, -1.0, np.inf])
    assert lg[0] == 0.0 and lg[1] == -np.inf and np.isnan(lg[2]) and lg[3] == np.inf
    assert ol.fmath("log1p", [1e-30])[0] == np.float32(1e-30)
    assert ol.fmath("erfc", [0.0])[0] == 1.0
    assert ol.fmath("pow075", [0.0, 1.0]).tolist() == [0.0, 1.0]


def test_sin_cos_quadrant_identities():
    xs = RNG.uniform(-50, 50, 100000).astype(np.float32)
    s, c = ol.fmath("sin", xs), ol.fmath("cos", xs)
    assert np.abs(s * s + c * c - 1.0).max() < 4e-7
    assert np.array_equal(ol.fmath("sin", -xs), -s)  # odd
    assert np.array_equal(ol.fmath("cos", -xs), c)   # even


def test_atan2_quadrants_and_accuracy():
    y = RNG.uniform(-3, 3, 200000).astype(np.float32)
    x = RNG.uniform(-3, 3, 200000).astype(np.float32)
    got = ol.fmath_atan2(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert ulp_err(got, ref).max() <= 3.0
    sp = ol.fmath_atan2([0.0, 1.0, -1.0, 0.0, 0.0, 1.0, -1.0], [0.0, 0.0, 0.0, 1.0, -1.0, -1.0, -1.0])
    assert np.allclose(sp, [0.0, np.pi / 2, -np.pi / 2, 0.0, np.pi, np.arctan2(1, -1), np.arctan2(-1, -1)], atol=1e-7)
