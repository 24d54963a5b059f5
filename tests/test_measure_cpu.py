"""Pins of the oracle's FP64 image statistics (oracle/or_measure.cpp), CPU only.

  * DeviceMSE and DeviceMSELuminance (core/helper/measure.cuh:18-73): equal
    to the exactly rounded (math.fsum) sum of the same FP64 terms, with the
    luminance taken in FP32 by numpy in the reference's operation order;
    identical images give 0, a constant luminance offset c gives c^2, n = 0
    gives NaN;
  * relMSE: the oracle equals an FP64 numpy form with the matrix applied as a
    matrix product; an image that is (1 + k) times a gray reference whose
    channels are r gives k^2 r^2 / (r^2 + eps) per channel; per-pixel FP64
    sums with scale 1/spp equal the mean image.
"""
import math

import numpy as np

import oracle_lib as ol

M = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]],
             np.float32).astype(np.float64)


def lum32(x):
    x = np.asarray(x, np.float32).reshape(-1, 3)
    return np.float32(0.2126) * x[:, 0] + np.float32(0.7152) * x[:, 1] + np.float32(0.0722) * x[:, 2]


def test_mse_and_luminance_mse_equal_exact_sums():
    g = np.random.default_rng(1)
    for n in (1, 7, 1000, 100_003):
        a = g.gamma(0.5, 2.0, (n, 3)).astype(np.float32)
        b = (a * g.uniform(0.8, 1.2, (n, 3))).astype(np.float32)
        d = (lum32(a).astype(np.float64) - lum32(b).astype(np.float64))
        assert abs(ol.mse_luminance(a, b) - math.fsum(d * d) / n) <= 1e-11 * (math.fsum(d * d) / n)
        e = (a.reshape(-1) - b.reshape(-1)).astype(np.float64)
        assert abs(ol.mse(a, b) - math.fsum(e * e) / e.size) <= 1e-11 * (math.fsum(e * e) / e.size)
    a = g.random((64, 3)).astype(np.float32)
    assert ol.mse_luminance(a, a) == 0.0 and ol.mse(a, a) == 0.0
    # A constant luminance offset: x + c on every channel moves L by c (weights sum to 1).
    c = np.float32(0.25)
    b = np.full((64, 3), 0.5, np.float32)
    a = b + c
    assert abs(ol.mse_luminance(a, b) - 0.0625) < 1e-7
    assert math.isnan(ol.mse_luminance(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)))


def relmse_np(img, ref, eps):
    a = np.maximum(np.asarray(img, np.float64).reshape(-1, 3) @ M.T, 0.0)
    b = np.maximum(np.asarray(ref, np.float64).reshape(-1, 3) @ M.T, 0.0)
    return float(np.mean((a - b) ** 2 / (b * b + eps)))


def test_relmse_matches_numpy_and_closed_forms():
    g = np.random.default_rng(2)
    ref = g.gamma(1.0, 0.5, (5000, 3)).astype(np.float32)
    img = (ref * g.uniform(0.5, 1.5, ref.shape)).astype(np.float32)
    for eps in (1e-2, 1e-3):
        o = ol.relmse(img, ref, eps)
        assert abs(o - relmse_np(img, ref, eps)) <= 1e-12 * o
    # (1 + k) x a reference whose linear-sRGB channels are r (an XYZ chosen by the inverse matrix).
    r = np.array([0.3, 0.6, 0.9])
    xyz = np.linalg.solve(M, r)
    ref = np.tile(xyz, (100, 1))
    for k in (0.1, -0.5):
        o = ol.relmse(ref * (1 + k), ref, 1e-2)
        want = float(np.mean(k * k * r * r / (r * r + 1e-2)))
        assert abs(o - want) < 1e-12
    # FP64 sums with scale 1/spp equal their mean image.
    sums = g.gamma(1.0, 4.0, (300, 3))
    mean = (sums / 16.0).astype(np.float32)
    ref = g.gamma(1.0, 0.25, (300, 3)).astype(np.float32)
    assert abs(ol.relmse(sums, ref, img_scale=1 / 16) - relmse_np(sums / 16, ref, 1e-2)) < 1e-12
    assert abs(ol.relmse(mean, ref) - ol.relmse(sums, ref, img_scale=1 / 16)) < 1e-6
