"""GPU FP64 image statistics (ap_mse_f64, ap_mse_luminance_f64,
ap_relmse_f64) against the oracle's sequential FP64 restatement.

Bar: relative difference <= 1e-11 (the same FP64 terms, bit-identical per
element; only the order of the FP64 additions differs: a fixed tree on the
device, a running sum in the oracle). The device result is the same bits on
every call. Sizes cover an empty input (NaN), one element, block edges, more
elements than one pass of the 1024-block grid, and a 1080p frame; images as
float3 means and as FP64 sums with a 1/spp scale; and the odd/even
framebuffers of a render, as Camera::Render reduces them (core/camera.cu:94-97).
"""
import math

import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import test_render_gpu as rg

pytestmark = pytest.mark.gpu

SIZES = [1, 255, 256, 257, 1000, 262_147, 1920 * 1080]


def close(g, o):
    return abs(g - o) <= 1e-11 * abs(o) + 1e-300


def test_luminance_mse_and_mse_match_the_oracle(gpu):
    g = np.random.default_rng(3)
    for n in SIZES:
        a = g.gamma(0.5, 2.0, (n, 3)).astype(np.float32)
        b = (a * g.uniform(0.8, 1.2, (n, 3))).astype(np.float32)
        da, db = ap.DeviceBuffer.from_numpy(a), ap.DeviceBuffer.from_numpy(b)
        gl, ol_ = ap.mse_luminance_f64(da, db, n), ol.mse_luminance(a, b)
        gm, om = ap.mse_f64(da, db, 3 * n), ol.mse(a, b)
        print(f"[measure] n {n}: luminance MSE {gl:.17g} (oracle {ol_:.17g}), MSE {gm:.17g} (oracle {om:.17g})",
              flush=True)
        assert close(gl, ol_) and close(gm, om)
        assert ap.mse_luminance_f64(da, db, n) == gl   # same bits every call
    empty = ap.DeviceBuffer(16)
    assert math.isnan(ap.mse_luminance_f64(empty, empty, 0))
    assert math.isnan(ap.mse_f64(empty, empty, 0))


def test_relmse_matches_the_oracle(gpu):
    g = np.random.default_rng(4)
    for n in SIZES:
        ref = g.gamma(1.0, 0.5, (n, 3)).astype(np.float32)
        img = (ref * g.uniform(0.5, 1.5, ref.shape)).astype(np.float32)
        sums = img.astype(np.float64) * 32.0
        dr, di, ds = (ap.DeviceBuffer.from_numpy(x) for x in (ref, img, sums))
        got = ap.relmse_f64(di, dr, n)
        want = ol.relmse(img, ref)
        got_s = ap.relmse_f64(ap.ImageView.of(ds, ap.IMAGE_XYZ_SUMS_F64, 1.0 / 32.0), dr, n, eps=1e-3)
        want_s = ol.relmse(sums, ref, 1e-3, img_scale=1.0 / 32.0)
        print(f"[measure] n {n}: relMSE {got:.17g} (oracle {want:.17g}); from sums {got_s:.17g} ({want_s:.17g})",
              flush=True)
        assert close(got, want) and close(got_s, want_s)
        assert ap.relmse_f64(di, dr, n) == got
    with pytest.raises(ap.ApertureError):
        ap.relmse_f64(di, ap.ImageView.of(dr, 7), 4)
    with pytest.raises(ap.ApertureError):
        ap.relmse_f64(di, dr, 4, eps=0.0)


def test_odd_even_variance_of_a_render(gpu):
    """The reference's per-frame variance estimate, 1e6 x DeviceMSELuminance
    of the odd and even framebuffers, written into a caller's buffer in stream
    order."""
    objs = ap.synth_box_scene(1, 1, 15.0)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 64, 48)
    scene = ap.SceneRenderer(objs)
    out = rg.render_gpu(scene, cam, rg.box_params(spp=16), odd_even=True)
    n = cam.width * cam.height
    d_odd, d_even = ap.DeviceBuffer.from_numpy(out["odd"]), ap.DeviceBuffer.from_numpy(out["even"])
    res = ap.DeviceBuffer(8)
    stream = ap.Stream()
    ap.mse_luminance_f64(d_odd, d_even, n, out=res, stream=stream)
    stream.synchronize()
    var = res.to_numpy(np.float64, 1)[0] * 1e6
    want = ol.mse_luminance(out["odd"], out["even"]) * 1e6
    print(f"[measure] odd/even variance x 1e6: {var:.12g} (oracle {want:.12g})", flush=True)
    assert var > 0 and close(var, want)
    scene.close()
