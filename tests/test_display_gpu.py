"""GPU checks of the display and statistics framebuffers against the oracle
(oracle/or_display.cpp, or_render_screen in oracle/or_render.cpp).

Bars: every field of the screen G-buffer and statistics records after a
render (frames recorded in frame order per pixel, several frames per batch,
across two calls, one and four wavelengths, both triangle material modes)
is bit-identical to the oracle's; ap_tonemap's ARGB8 output is
bit-identical to the oracle's Tonemap for every tone curve and display mode
(same FP32 operations and the shared portable pow); bad arguments are
refused.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GB_FIELDS = ("albedo", "normal", "metallic", "roughness", "depth", "valid_frame_index", "frame_index", "shading_model")
ST_FIELDS = ("num", "mean", "m2_div_num", "m3_div_num")


def same_bits(a, b, fields):
    for f in fields:
        x, y = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        if x.tobytes() != y.tobytes():
            return f
    return None


def render_records(scene, cam, prm, calls=1):
    n = cam.width * cam.height
    hdr = ap.DeviceBuffer(n * 12)
    hdr.zero()
    gb = ap.DeviceBuffer.from_numpy(np.zeros(n, ap.SCREEN_GBUFFER_DTYPE))
    st = ap.DeviceBuffer.from_numpy(np.zeros(n, ap.SCREEN_STATISTICS_DTYPE))
    spp = prm.spp
    per = spp // calls
    for k in range(calls):
        p = ap.RenderParams.from_buffer_copy(prm)
        p.frame_index = prm.frame_index + k * per
        p.spp = per if k < calls - 1 else spp - per * (calls - 1)
        scene.render(cam, p, hdr, gbuffer=gb, statistics=st)
    ap.synchronize()
    return (hdr.to_numpy(np.float32, n * 3), gb.to_numpy(ap.SCREEN_GBUFFER_DTYPE, n),
            st.to_numpy(ap.SCREEN_STATISTICS_DTYPE, n))


@pytest.mark.parametrize("tri_mode,wavelengths,calls", [(0, 1, 1), (1, 1, 2), (0, 4, 2)])
def test_records_match_the_oracle(gpu, tri_mode, wavelengths, calls):
    print(f"[display] records: triangle mode {tri_mode}, {wavelengths} wavelengths, {calls} calls", flush=True)
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 60.0, 24, 18)
    prm = ap.RenderParams.defaults(spp=6, max_depth=8, seed=13, wavelengths=wavelengths, triangle_material=tri_mode)
    scene = ap.SceneRenderer(objs)
    margin = np.float32(scene.info().line_margin)
    hdr, gb, st = render_records(scene, cam, prm, calls)
    o_hdr, _, o_gb, o_st = ol.OracleBVH(objs).render_screen(cam, prm, cull_margin=margin)
    assert np.array_equal(hdr, o_hdr.reshape(-1))
    bad_g, bad_s = same_bits(gb, o_gb, GB_FIELDS), same_bits(st, o_st, ST_FIELDS)
    print(f"[display]   hits {int((gb['valid_frame_index'] > 0).sum())} of {len(gb)} pixels; "
          f"mismatch: gbuffer {bad_g}, statistics {bad_s}", flush=True)
    assert bad_g is None and bad_s is None
    assert np.all(gb["frame_index"] == 6) and np.all(st["num"] == 6)
    scene.close()


def test_tonemap_matches_the_oracle(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 60.0, 40, 30)
    scene = ap.SceneRenderer(objs)
    _, gb, st = render_records(scene, cam, ap.RenderParams.defaults(spp=4, max_depth=8, seed=3))
    n = cam.width * cam.height
    g = np.random.default_rng(7)
    hdr = np.concatenate([g.uniform(0, 1.5, (n - 200, 3)), g.uniform(0, 200, (197, 3)),
                          [[0, 0, 0], [np.nan, 1, 0], [np.inf, 2, 1]]]).astype(np.float32)
    d_hdr, d_gb, d_st = ap.DeviceBuffer.from_numpy(hdr), ap.DeviceBuffer.from_numpy(gb), ap.DeviceBuffer.from_numpy(st)
    for display in range(10):
        tones = (ap.TONE_LINEAR, ap.TONE_GAMMA, ap.TONE_ACES) if display == ap.DISPLAY_IMAGE else (ap.TONE_LINEAR,)
        for tone in tones:
            for exposure in ((1.0, 0.6) if display == ap.DISPLAY_IMAGE else (1.0,)):
                out = ap.tonemap(d_hdr, cam.width, cam.height, tone=tone, display=display, exposure=exposure,
                                 gbuffer=d_gb, statistics=d_st)
                ap.synchronize()
                got = out.to_numpy(np.uint32, n)
                want = ol.tonemap(hdr, gb, st, n=n, tone=tone, display=display, exposure=exposure)
                diff = int((got != want).sum())
                print(f"[display] tonemap display {display} tone {tone} exposure {exposure}: "
                      f"{diff} pixels differ", flush=True)
                assert diff == 0
    scene.close()


def test_tonemap_refuses_bad_arguments(gpu):
    hdr = ap.DeviceBuffer(16 * 12)
    hdr.zero()
    with pytest.raises(ap.ApertureError):
        ap.tonemap(hdr, 4, 4, display=ap.DISPLAY_ALBEDO)      # no G-buffer
    with pytest.raises(ap.ApertureError):
        ap.tonemap(hdr, 4, 4, display=ap.DISPLAY_STATISTICS_MEAN)
    with pytest.raises(ap.ApertureError):
        ap.tonemap(hdr, 4, 4, tone=3)
    with pytest.raises(ap.ApertureError):
        ap.tonemap(hdr, 4, 4, display=10)
    with pytest.raises(ap.ApertureError):
        ap.tonemap(None, 4, 4)
    out = ap.tonemap(hdr, 4, 4, tone=ap.TONE_ACES)
    ap.synchronize()
    assert np.all(out.to_numpy(np.uint32, 16) == 0xff000000)
