"""GPU parity of the spectral basis kernels (ap_spectrum_*) against the
oracle's restatement of core/spectrum (oracle/or_spectrum.cpp).

Both sides evaluate the same FP32 network in the same operation order with
contraction off, and take sin/cos/exp/log/erfc from the same portable
implementation (ap_fmath.h), so the bar is bit-exact equality; the reference's
own golden vectors go through the GPU encoders too
(tests/golden/spectrum_golden.json, same bar as the oracle's test).
"""
import json
import os

import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def spectrum(gpu):
    print("[spectrum] building device LUTs", flush=True)
    ap.spectrum_setup()
    return ap


def test_encoders_match_oracle(spectrum):
    g = np.random.default_rng(5)
    rgb = np.concatenate([g.random((4096, 3), dtype=np.float32),
                          np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
                                   np.float32)])
    lam = np.concatenate([np.float32(360.0) + np.float32(420.0) * g.random(4096, dtype=np.float32),
                          np.array([360.0, 780.0, 555.0], np.float32)])
    e_rgb = np.abs(ap.spectrum_encode_rgb(rgb) - ol.rgb_encode(rgb)).max()
    e_lam = np.abs(ap.spectrum_encode_lambda(lam) - ol.lambda_encode(lam)).max()
    print(f"[spectrum] encoder max|err| rgb {e_rgb:.3g} lambda {e_lam:.3g}", flush=True)
    assert e_rgb == 0.0 and e_lam == 0.0


def test_golden_vectors_on_gpu(spectrum):
    gold = json.load(open(os.path.join(HERE, "golden", "spectrum_golden.json")))
    rk = ap.spectrum_encode_rgb(np.array(gold["rgb"], np.float32))
    lk = ap.spectrum_encode_lambda(np.array(gold["lambda_nm"], np.float32))
    pred = np.array([ol.spectral_interact(rk[i], lk[i]) for i in range(len(rk))])
    err = pred - np.array(gold["gt"])
    assert np.abs(err).max() < 5e-4 and (err ** 2).mean() < 5e-8


def test_lut_queries_match_oracle(spectrum):
    g = np.random.default_rng(6)
    # Includes cell centres, cell faces, the clamped border and out-of-range inputs.
    rgb = np.concatenate([g.random((3000, 3), dtype=np.float32),
                          ((np.arange(0, 512, 37)[:, None] + 0.5) / 512.0).astype(np.float32).repeat(3, 1),
                          np.array([[0, 0, 0], [1, 1, 1], [1e-7, 0.999999, 0.5], [-0.1, 1.2, 0.3]], np.float32)])
    got = ap.spectrum_query_rgb(rgb)
    want = ol.query_rgb_basis(rgb)
    e_rgb = np.abs(got - want).max()
    lam = np.concatenate([np.linspace(350.0, 790.0, 4001, dtype=np.float32),
                          np.float32(360.0) + np.float32(420.0) * g.random(2000, dtype=np.float32)])
    e_lam = np.abs(ap.spectrum_query_lambda(lam) - ol.query_lambda(lam)).max()
    print(f"[spectrum] LUT max|err| rgb {e_rgb:.3g} lambda {e_lam:.3g}", flush=True)
    assert e_rgb == 0.0 and e_lam == 0.0


def test_wavelength_sampling_matches_oracle(spectrum):
    g = np.random.default_rng(7)
    u3 = g.random((200000, 3), dtype=np.float32)
    u3[:8] = [[0, 0, 0], [0.05, 0.5, 0.5], [0.1, 0.5, 0.5], [0.3, 1e-7, 0.2], [0.6, 0.999999, 0.7],
              [0.9, 0.5, 1e-6], [0.999, 0.5, 0.999999], [0.5, 0.5, 0.5]]
    lam, pdf = ap.spectrum_sample_lambda(u3)
    rl, rp = ol.sample_cie1931(u3)
    assert lam.min() >= 360.0 and lam.max() <= 780.0
    el = np.abs(lam / rl - 1.0).max()
    ep = np.abs(pdf / rp - 1.0).max()
    print(f"[spectrum] sampling max rel err lambda {el:.3g} pdf {ep:.3g}", flush=True)
    assert np.array_equal(lam, rl) and np.array_equal(pdf, rp)


def test_interact_matches_oracle(spectrum):
    g = np.random.default_rng(8)
    n = 20000
    lam = np.float32(360.0) + np.float32(420.0) * g.random(n, dtype=np.float32)
    rgb = g.random((n, 3), dtype=np.float32)
    rgb[: n // 4] *= np.float32(20.0)  # emissive values above one take the scaled branch
    got = ap.spectrum_interact(lam, rgb)
    want = ol.interact(lam, rgb)
    scale = np.maximum(1.0, rgb.max(1))
    err = (np.abs(got - want) / scale).max()
    print(f"[spectrum] interact max scaled err {err:.3g}", flush=True)
    assert np.array_equal(got, want)
    assert got.min() >= 0.0


def test_xyz_to_srgb_matches_oracle(spectrum):
    g = np.random.default_rng(9)
    xyz = (g.random((1000, 3), dtype=np.float32) * np.float32(3.0)).astype(np.float32)
    got = ap.xyz_to_srgb(xyz)
    want = np.array([ol.xyz_to_srgb(v) for v in xyz])
    assert np.array_equal(got, want)
