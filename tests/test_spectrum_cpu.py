"""CPU tests pinning the oracle's spectral restatement (oracle/or_spectrum.cpp).

Pins, independent of the restated formulas:
  * the reference's own known-answer vectors inputTest/gtTest
    (core/spectrum/spectrum_basis.cuh:2472-2489; tests/golden/spectrum_golden.json).
    They are eval-mode PyTorch outputs on an unstated device, so they agree to
    FP32-network noise, not bitwise: measured max |err| 3.6e-4, median 7.6e-5,
    MSE 1.5e-8. A transposed or mis-ordered weight block gives errors ~1e-1;
  * ndtrif(0.5) = 0 and u = Phi(ndtrif(u)) (the reference's own check,
    sample_ciexyz.cuh:335-549);
  * the wavelength pdf integrates to one and matches a histogram of samples;
  * D65Norm normalises the D65 luminance to Y = 1 (ciexyz.cuh:85 comment);
  * D65 white converts to sRGB (1, 1, 1) (tonemap.cuh:116-126 constants);
  * the Monte Carlo round trip RGB -> spectrum -> XYZ -> sRGB returns the input
    colour (the reference's ValidationSpectrum, spectrum_lut.cuh:238-321).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    return json.load(open(os.path.join(HERE, "golden", "spectrum_golden.json")))


def test_golden_vectors():
    g = golden()
    rk = ol.rgb_encode(g["rgb"])
    lk = ol.lambda_encode(g["lambda_nm"])
    pred = np.array([ol.spectral_interact(rk[i], lk[i]) for i in range(100)])
    err = pred - np.array(g["gt"])
    assert np.abs(err).max() < 5e-4
    assert (err ** 2).mean() < 5e-8


def test_golden_vectors_catch_a_transposed_layer():
    """The same inputs through a deliberately wrong weight order must fail the
    tolerance above (checks that the golden test has teeth)."""
    g = golden()
    rk = ol.rgb_encode(g["rgb"])
    lk = ol.lambda_encode(g["lambda_nm"])
    pred = np.array([ol.spectral_interact(rk[i][::-1].copy(), lk[i]) for i in range(100)])
    assert np.abs(pred - np.array(g["gt"])).max() > 1e-2


def phi(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def test_ndtrif_round_trip():
    assert ol.ndtrif(0.5) == 0.0
    for u in np.concatenate([np.linspace(1e-6, 0.024, 50), np.linspace(0.03, 0.97, 200), np.linspace(0.976, 0.99999, 50)]):
        assert abs(phi(ol.ndtrif(u)) - u) < 2e-6 * max(1.0, u / 0.5)


def test_lambda_pdf_normalised_and_matches_samples():
    xs = np.linspace(360.0, 780.0, 84001)
    p = np.array([ol.pdf_cie1931(x) for x in xs])
    integral = np.trapezoid(p, xs)
    assert abs(integral - 1.0) < 2e-3
    g = np.random.default_rng(3)
    lam, pdf = ol.sample_cie1931(g.random((200000, 3), dtype=np.float32))
    assert lam.min() >= 360.0 and lam.max() <= 780.0
    for i in range(0, 200000, 997):
        assert abs(pdf[i] - max(ol.pdf_cie1931(lam[i]), 1e-16)) <= 1e-6 * pdf[i]
    hist, edges = np.histogram(lam, bins=42, range=(360.0, 780.0))
    expected = np.array([np.trapezoid(p[(xs >= a) & (xs <= b)], xs[(xs >= a) & (xs <= b)])
                         for a, b in zip(edges[:-1], edges[1:])]) * len(lam)
    mask = expected > 200
    chi2 = (((hist - expected) ** 2) / expected)[mask].sum()
    assert chi2 < 3.0 * mask.sum()


def test_d65_luminance_normalisation():
    xs = np.arange(360, 781, dtype=np.float32)
    y = np.array([ol.d65_norm(x) * ol.lambda_to_xyz(x)[1] for x in xs])
    assert abs(np.trapezoid(y, xs) - 1.0) < 2e-3


def test_cmf_table_and_srgb_white():
    a, b = ol.lambda_to_xyz(500.0), ol.lambda_to_xyz(501.0)
    mid = ol.lambda_to_xyz(500.5)
    assert np.allclose(mid, 0.5 * (a + b), rtol=1e-6, atol=1e-9)
    assert ol.lambda_to_xyz(555.0)[1] > 0.99  # photopic peak near 555 nm
    xs = np.arange(360, 781, dtype=np.float32)
    white = np.array([[ol.d65_norm(x) * v for v in ol.lambda_to_xyz(x)] for x in xs])
    xyz = np.trapezoid(white, xs, axis=0)
    assert np.allclose(ol.xyz_to_srgb(xyz), [1.0, 1.0, 1.0], atol=5e-3)


def test_query_lambda_at_cell_centres_equals_encoder():
    step = (780.0 - 360.0) / 840.0
    lam = (360.0 + (np.arange(5, 835, 37) + 0.5) * step).astype(np.float32)
    assert np.allclose(ol.query_lambda(lam), ol.lambda_encode(lam), rtol=0, atol=1e-6)


@pytest.mark.parametrize("rgb", [(1.0, 1.0, 1.0), (0.9, 0.2, 0.1), (0.1, 0.8, 0.2), (0.2, 0.3, 0.9), (0.5, 0.5, 0.5)])
def test_rgb_spectrum_round_trip(rgb):
    """ValidationSpectrum: integrate basis(rgb) . basis(lambda) * CMF * D65 / pdf."""
    g = np.random.default_rng(11)
    n = 20000
    lam, pdf = ol.sample_cie1931(g.random((n, 3), dtype=np.float32))
    lb = ol.query_lambda(lam)
    rb = ol.query_rgb_basis(np.array([rgb], np.float32))[0]
    refl = np.clip(lb @ rb, 0.0, 1.0)
    w = np.array([ol.d65_norm(l) for l in lam]) * refl / pdf
    xyz = (np.array([ol.lambda_to_xyz(l) for l in lam]) * w[:, None]).mean(0)
    out = ol.xyz_to_srgb(xyz)
    assert np.allclose(out, rgb, atol=0.04), (out, rgb)
