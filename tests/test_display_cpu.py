"""CPU tests pinning the oracle's restatement of the display and statistics
framebuffers (oracle/or_display.cpp: RecordScreenGBuffer, RecordShadingModel,
RecordScreenStatisticBuffer, the Tonemap kernel with Gamma, ACES and the
viridis heat map) with properties fixed independently of it:

  * binary16 rounding equals numpy's float16 (normals, subnormals, ties,
    overflow to infinity);
  * the statistics record's mean, M2/n and M3/n equal numpy's two-pass mean,
    variance and third central moment of the Box-Cox transformed samples
    (lambda 0.5: 2 (sqrt(max(x, 0)) - 1));
  * the G-buffer record is the running mean of its inputs over the frames
    that hit (within binary16 precision), its shading-model histogram the
    fraction of frames per model (misses in slot 4), its counters the frame
    and hit counts, and a first hit after misses starts from zero;
  * tonemap: the linear curve is trunc(255 * saturate(XYZ2SRGB(hdr) *
    exposure)) exactly; the gamma curve is numpy's x^(1/2.2) within 1 LSB;
    ACES keeps grays gray, is monotone, maps 0 to 0 and saturates; the heat
    map hits its five knots at t = 0, 1/4, 1/2, 3/4 and keeps the
    reference's t = 1 -> third knot; normals map n * 0.5 + 0.5;
  * a render with the records on: per pixel the statistics equal numpy's
    moments of the pixel's own per-frame samples (read back frame by frame
    from the FP64 sums), and the G-buffer of a camera looking straight down
    at a floor holds the floor's albedo, normal, roughness and metallic.
"""
import numpy as np

import aperture as ap
import oracle_lib as ol
import test_nrc_render_cpu as nrc_cpu

G = np.random.default_rng(31)


def test_half_rounding_matches_numpy():
    x = np.concatenate([G.normal(0, 1, 20000), G.normal(0, 1e-5, 5000), G.normal(0, 3e4, 5000),
                        [65504.0, 65519.0, 65520.0, 1e6, -1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0, -0.0,
                         np.inf, -np.inf], np.float32(1.0) + np.arange(-4, 5) * 2.0 ** -11]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ol.half_round(x), want)


def box_cox(x):
    return 2.0 * (np.sqrt(np.maximum(x.astype(np.float64), 0.0)) - 1.0)


def test_statistics_are_the_box_cox_moments():
    for n, scale in ((1, 1.0), (7, 0.3), (1000, 2.0), (20000, 50.0)):
        x = (G.gamma(0.7, scale, (n, 3))).astype(np.float32)
        x[::17, 1] = 0.0
        st = ol.record_statistics(x)
        b = box_cox(x)
        mean = b.mean(0)
        m2 = ((b - mean) ** 2).mean(0)
        m3 = ((b - mean) ** 3).mean(0)
        assert st["num"] == n
        assert np.allclose(st["mean"], mean, rtol=2e-5, atol=2e-6 * max(1.0, np.abs(b).max()))
        assert np.allclose(st["m2_div_num"], m2, rtol=2e-4, atol=1e-6)
        assert np.allclose(st["m3_div_num"], m3, rtol=5e-3, atol=1e-5 * max(1.0, np.abs(m3).max()))
    # Split runs chain: recording a then b equals recording a + b at once.
    x = G.uniform(0, 3, (64, 3)).astype(np.float32)
    st = ol.record_statistics(x[40:], ol.record_statistics(x[:40]))
    whole = ol.record_statistics(x)
    for f in ("num", "mean", "m2_div_num", "m3_div_num"):
        assert np.array_equal(np.asarray(st[f]).view(np.uint32), np.asarray(whole[f]).view(np.uint32))


def gsamples(n, hit_frac=0.7, seed=0):
    g = np.random.default_rng(seed)
    s = np.zeros(n, ol.GSAMPLE_DTYPE)
    s["hit"] = g.uniform(size=n) < hit_frac
    s["albedo"] = g.uniform(0, 1, (n, 3))
    nn = g.normal(size=(n, 3))
    s["normal"] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    s["metallic"] = g.uniform(size=n)
    s["roughness"] = g.uniform(size=n)
    s["depth"] = g.uniform(0.5, 20, n)
    s["model"] = np.where(s["hit"], g.choice([0, 1, 3], n), 4)
    return s


def test_gbuffer_is_the_running_mean_of_its_hits():
    for n, frac in ((1, 1.0), (40, 0.7), (400, 0.3)):
        s = gsamples(n, frac, seed=n)
        r = ol.record_gbuffer(s)
        hits = s[s["hit"] != 0]
        assert r["frame_index"] == n and r["valid_frame_index"] == len(hits)
        if len(hits):
            # binary16 running means drift by a few half ulps per frame at most.
            assert np.allclose(r["albedo"].astype(np.float64), hits["albedo"].mean(0), atol=2e-3 + 5e-5 * n)
            assert np.allclose(r["normal"].astype(np.float64), hits["normal"].mean(0), atol=2e-3 + 5e-5 * n)
            assert abs(float(r["metallic"]) - hits["metallic"].mean()) < 2e-3 + 5e-5 * n
            assert abs(float(r["roughness"]) - hits["roughness"].mean()) < 2e-3 + 5e-5 * n
            assert abs(float(r["depth"]) - hits["depth"].mean()) < 1e-5 * hits["depth"].mean() * max(1, n / 10)
        frac_model = np.array([(s["model"] == k).mean() for k in range(5)])
        assert np.allclose(r["shading_model"].astype(np.float64), frac_model, atol=2e-3 + 5e-5 * n)
    # A first hit after misses starts the means from zero (not from stale bytes).
    miss = np.zeros(3, ol.GSAMPLE_DTYPE)
    miss["model"] = 4
    stale = np.zeros(1, ap.SCREEN_GBUFFER_DTYPE)
    stale["albedo"] = 9.0
    stale["depth"] = 123.0
    hit = gsamples(1, 1.0, seed=5)
    hit["albedo"] = 0.5
    hit["depth"] = 2.0
    r = ol.record_gbuffer(np.concatenate([miss, hit]), stale)
    assert np.all(r["albedo"] == np.float16(0.5)) and r["depth"] == 2.0
    assert r["valid_frame_index"] == 1 and r["frame_index"] == 4
    want = np.zeros(5)
    want[4], want[int(hit["model"][0])] = 0.75, 0.25   # three misses, then one hit
    assert np.allclose(r["shading_model"].astype(np.float64), want, atol=1e-3)


def srgb(xyz):
    m = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]], np.float32)
    x = np.asarray(xyz, np.float32).reshape(-1, 3)
    out = np.stack([(x[:, 0] * m[k, 0] + x[:, 1] * m[k, 1]) + x[:, 2] * m[k, 2] for k in range(3)], 1)
    return np.maximum(out, np.float32(0.0)).astype(np.float32)


def unpack(argb):
    a = np.asarray(argb, np.uint32)
    return np.stack([(a >> 16) & 255, (a >> 8) & 255, a & 255], 1).astype(np.int64), (a >> 24)


def test_tonemap_linear_and_gamma():
    hdr = np.concatenate([G.uniform(0, 1.2, (5000, 3)), G.uniform(0, 50, (500, 3)),
                          [[0, 0, 0], [np.nan, 1, 1], [0.9505, 1.0, 1.089]]]).astype(np.float32)
    for exposure in (1.0, 0.7):
        rgb, alpha = unpack(ol.tonemap(hdr, tone=ap.TONE_LINEAR, exposure=exposure))
        v = srgb(hdr) * np.float32(exposure)
        with np.errstate(invalid="ignore"):
            want = (np.float32(255.0) * np.clip(np.nan_to_num(v, nan=0.0), 0, 1).astype(np.float32)).astype(np.int64)
        assert np.all(alpha == 255) and np.array_equal(rgb, want)
        g, _ = unpack(ol.tonemap(hdr, tone=ap.TONE_GAMMA, exposure=exposure))
        with np.errstate(invalid="ignore"):
            gw = np.floor(255.0 * np.clip(np.nan_to_num(v.astype(np.float64) ** (1 / 2.2), nan=0.0), 0, 1))
        d = np.abs(g - gw)
        assert d.max() <= 1 and (d == 0).mean() > 0.99
    # D65 white at Y = 1 is (255, 255, 255) under every curve but ACES.
    for tone in (ap.TONE_LINEAR, ap.TONE_GAMMA):
        assert np.all(unpack(ol.tonemap(np.array([0.9505, 1.0, 1.089], np.float32), tone=tone))[0] >= 254)


def test_aces_properties():
    y = np.concatenate([[0.0], np.geomspace(1e-4, 1e3, 400)]).astype(np.float32)
    gray_rgb = np.repeat(y[:, None], 3, 1)
    # XYZ of sRGB gray: the inverse of the tonemap's matrix applied to (y, y, y).
    m = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]])
    xyz = (gray_rgb.astype(np.float64) @ np.linalg.inv(m).T).astype(np.float32)
    rgb, _ = unpack(ol.tonemap(xyz, tone=ap.TONE_ACES))
    assert np.abs(rgb - rgb[:, 1:2]).max() <= 2          # grays stay gray
    assert np.all(np.diff(rgb[:, 1]) >= 0)               # monotone
    assert rgb[0].tolist() == [0, 0, 0] and rgb[-1, 1] >= 245  # the fit tends to a/c = 0.949 before its gamma
    lin, _ = unpack(ol.tonemap(xyz, tone=ap.TONE_LINEAR))
    mid = (y > 0.02) & (y < 0.5)
    assert np.all(rgb[mid, 1] > lin[mid, 1])             # the curve lifts mid tones (its final gamma)


def viridis_knots():
    return np.array([[0.267004, 0.004874, 0.329415], [0.229739, 0.322361, 0.545706], [0.127568, 0.566949, 0.550556],
                     [0.369214, 0.788888, 0.382914], [0.993248, 0.906157, 0.143936]], np.float32)


def test_gbuffer_and_statistics_views():
    gb = np.zeros(6, ap.SCREEN_GBUFFER_DTYPE)
    t = np.array([0.0, 0.25, 0.5, 0.75, 1.0, 0.125], np.float32)
    gb["depth"] = np.float32(1.0) / np.where(t > 0, t, np.float32(1e-30)) - np.float32(1.0)
    gb["depth"][0] = 1e30
    gb["normal"] = [0.0, 1.0, 0.0]
    gb["albedo"] = [[0.5, 0.25, 1.0]] * 6
    gb["metallic"] = t
    rgb, _ = unpack(ol.tonemap(gbuffer=gb, display=ap.DISPLAY_DEPTH))
    k = viridis_knots()
    want = (np.float32(255.0) * k).astype(np.int64)
    assert np.array_equal(rgb[:4], want[:4])
    assert np.array_equal(rgb[4], want[3])                 # the reference's t = 1 lands on the third knot
    mid = (np.float32(255.0) * (k[0] + (k[1] - k[0]) * np.float32(0.5))).astype(np.int64)
    assert np.abs(rgb[5] - mid).max() <= 1
    assert np.array_equal(unpack(ol.tonemap(gbuffer=gb, display=ap.DISPLAY_METALLIC))[0], rgb)
    assert unpack(ol.tonemap(gbuffer=gb, display=ap.DISPLAY_NORMAL))[0][0].tolist() == [127, 255, 127]
    assert unpack(ol.tonemap(gbuffer=gb, display=ap.DISPLAY_ALBEDO))[0][0].tolist() == [127, 63, 255]
    st = np.zeros(3, ap.SCREEN_STATISTICS_DTYPE)
    st["num"] = [0, 1000, 4000000000]
    st["mean"] = [[-1.0, 2.0, 8.0]] * 3
    st["m2_div_num"] = [[0.4, 0.0, 3.0]] * 3
    rgb, _ = unpack(ol.tonemap(statistics=st, display=ap.DISPLAY_STATISTICS_NUM))
    want_r = np.floor(255.0 * np.minimum(np.log(st["num"].astype(np.float64) + 1.0) * 0.01, 1.0))
    assert np.array_equal(rgb[:, 0], want_r.astype(np.int64)) and np.all(rgb[:, 1:] == 0)
    assert unpack(ol.tonemap(statistics=st, display=ap.DISPLAY_STATISTICS_MEAN))[0][0].tolist() == [63, 127, 255]
    assert unpack(ol.tonemap(statistics=st, display=ap.DISPLAY_STATISTICS_M2))[0][0].tolist() == [25, 0, 191]


def test_render_records_the_pixels_own_samples():
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 12, 10)
    bvh = ol.OracleBVH(objs)
    spp = 6
    prm = ap.RenderParams.defaults(spp=spp, max_depth=8, seed=9)
    hdr, sums, gb, st = bvh.render_screen(cam, prm, cull_margin=2 ** -12)
    ref, ref_sums, _ = bvh.render(cam, prm, cull_margin=2 ** -12)
    assert np.array_equal(hdr, ref) and np.array_equal(sums, ref_sums)
    per_frame = []
    for f in range(spp):
        p1 = ap.RenderParams.defaults(spp=1, frame_index=f, max_depth=8, seed=9)
        per_frame.append(bvh.render(cam, p1, cull_margin=2 ** -12)[1].reshape(-1, 3))
    x = np.stack(per_frame, 1)  # (pixels, frames, 3): each frame's clamped XYZ sample
    b = box_cox(x)
    mean = b.mean(1)
    assert np.all(st["num"] == spp)
    assert np.allclose(st["mean"], mean, rtol=1e-5, atol=1e-5)
    assert np.allclose(st["m2_div_num"], ((b - mean[:, None]) ** 2).mean(1), rtol=1e-4, atol=1e-5)
    assert np.all(gb["frame_index"] == spp)


def test_render_gbuffer_of_a_floor():
    objs = nrc_cpu.floor_and_ceiling(rho_f=0.5, rho_c=0.8)
    objs["metallic"] = 0.25
    cam = ap.Camera.make([0.0, 0.5, 0.0], [0.0, -1.0, 0.0001], 20.0, 8, 8)
    _, _, gb, st = ol.OracleBVH(objs).render_screen(cam, ap.RenderParams.defaults(spp=5, max_depth=3, sky=0),
                                                    cull_margin=2 ** -12)
    assert np.all(gb["valid_frame_index"] == 5) and np.all(gb["frame_index"] == 5)
    assert np.allclose(gb["albedo"].astype(np.float64), 0.5, atol=2e-3)
    assert np.allclose(gb["normal"].astype(np.float64), [0.0, 1.0, 0.0], atol=1e-3)
    assert np.allclose(gb["roughness"].astype(np.float64), 0.25, atol=1e-3)
    assert np.allclose(gb["metallic"].astype(np.float64), 0.25, atol=1e-3)
    # Depth: the camera is 0.5 above the floor, looking straight down through a 20 degree field.
    assert np.all(gb["depth"] >= 0.5) and np.all(gb["depth"] < 0.5 / np.cos(np.radians(15.0)))
    model = int(objs["shading_model"][0])
    assert np.allclose(gb["shading_model"][:, model].astype(np.float64), 1.0, atol=1e-3)
    assert np.all(st["num"] == 5)
