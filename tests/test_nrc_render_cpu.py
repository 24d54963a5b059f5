"""CPU tests pinning the oracle's restatement of rendering with the learned
radiance cache (CalculateRadianceNRC in oracle/or_render.cpp; the contract is
include/aperture/ap_nrc.h) with properties fixed independently of it:

  * a cache that is never reached (spread threshold out of reach, or every
    path an unbiased training path) leaves the reference integrator exactly:
    same pixels, bit for bit, as or_render;
  * between a gray floor and a large gray ceiling with no light, a cache that
    predicts the constant c makes every path worth rho_floor * c;
  * a training target is (L_end - L_k) / M_k in sRGB: at the first vertex
    exactly XYZ2SRGB of the pixel's sample, and one vertex later larger by
    the floor's spectral albedo, the same factor in every channel;
  * with the camera straight above the floor and the ceiling one unit above
    it, a path ends at the ceiling iff pi / cos^4(theta) > spread * a0, so a
    threshold of 2 pi ends the fraction P(cos theta < 2^-1/4) = 1/sqrt(2) of
    cosine-sampled paths there.
"""
import math

import numpy as np

import aperture as ap
import oracle_lib as ol
import scenes

PIX_KEYS = 8


def quad(y, half, albedo, roughness=0.5):
    a, b, c, d = (-half, y, -half), (half, y, -half), (half, y, half), (-half, y, half)
    t = scenes.triangles([a, a], [c, d], [b, c])
    t["v0_normal"] = t["v1_normal"] = t["v2_normal"] = (0.0, 1.0, 0.0)
    t["albedo"] = albedo
    t["roughness"] = roughness
    return t


def floor_and_ceiling(rho_f=0.5, rho_c=0.8):
    return scenes.concat([quad(0.0, 200.0, rho_f, 0.25), quad(1.0, 200.0, rho_c, 0.75)])


def constant_cache(c):
    """Weights whose network outputs (c, c, c) for every input: the constant
    feature 46 feeds hidden unit 0 of every layer with weight 1."""
    p = np.zeros(ol.NRC_PARAMS, np.float32)
    p[0 * 4096 + 0 * 64 + 46] = 1.0
    for l in range(1, 5):
        p[l * 4096 + 0] = 1.0
    for o in range(3):
        p[5 * 4096 + o * 64 + 0] = c
    return p


BOUNDS = (np.array([-201, -1, -201], np.float32), np.array([201, 2, 201], np.float32))


def test_unreached_cache_is_the_reference_integrator():
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 16, 16)
    bvh = ol.OracleBVH(objs)
    prm = ap.RenderParams.defaults(spp=2, max_depth=8, seed=3)
    ref, ref_sums, _ = bvh.render(cam, prm, cull_margin=2 ** -12)
    w = constant_cache(0.7)
    hdr, sums, samples, nq = bvh.render_nrc(cam, prm, w, *BOUNDS, spread=1e30, train_paths=256,
                                            cull_margin=2 ** -12)
    assert nq == 0 and np.array_equal(hdr, ref) and np.array_equal(sums, ref_sums)
    assert len(samples) > 0
    hdr, sums, _, nq = bvh.render_nrc(cam, prm, w, *BOUNDS, spread=0.01, train_paths=256, unbiased_every=1,
                                      cull_margin=2 ** -12)
    assert nq == 0 and np.array_equal(hdr, ref) and np.array_equal(sums, ref_sums)


def test_constant_cache_closed_form():
    objs = floor_and_ceiling()
    cam = ap.Camera.make([0.0, 0.05, 0.0], [0.0, -1.0, 0.001], 20.0, 16, 16)
    prm = ap.RenderParams.defaults(spp=32, max_depth=8, sky=0)
    hdr, _, _, nq = ol.OracleBVH(objs).render_nrc(cam, prm, constant_cache(0.5), *BOUNDS, cull_margin=2 ** -12)
    assert nq >= 0.99 * 256
    y = float(hdr[..., 1].mean())
    assert abs(y / 0.25 - 1.0) < 0.02, y


def test_training_targets():
    """Targets are (L_end - L_k) / M_k in sRGB: at the first vertex (L = 0,
    M = 1) exactly XYZ2SRGB of the pixel's own sample, and with no light
    between vertices 0 and 1, target_0 = M_1 target_1 with M_1 the floor's
    spectral albedo at the path's wavelength, one factor for all channels."""
    objs = floor_and_ceiling()
    cam = ap.Camera.make([0.0, 0.05, 0.0], [0.0, -1.0, 0.001], 20.0, 32, 32)
    prm = ap.RenderParams.defaults(spp=1, max_depth=8, sky=0, seed=5)
    _, sums, s, nq = ol.OracleBVH(objs).render_nrc(cam, prm, constant_cache(0.5), *BOUNDS, train_paths=1024,
                                                    cull_margin=2 ** -12)
    assert nq >= 0.99 * 1024  # every path is a training path and ends at its second crossing
    key = s[:, 19].view(np.uint32)
    pix, vert = key // PIX_KEYS, key % PIX_KEYS
    assert len(np.unique(pix)) >= 0.99 * 1024
    v0, v1 = s[vert == 0], s[vert == 1]
    # Vertex 0 on the floor, seen from above; vertex 1 on the ceiling, seen from below.
    assert np.all(np.abs(v0[:, 1]) < 1e-5) and np.all(v0[:, 4] > 0.0)
    assert np.all(v0[:, 6:9] == (0.0, 1.0, 0.0)) and np.all(v0[:, 9] == 0.25) and np.all(v0[:, 10:13] == 0.5)
    assert np.all(np.abs(v1[:, 1] - 1.0) < 1e-5) and np.all(v1[:, 4] < 0.0)
    assert np.all(v1[:, 6:9] == (0.0, -1.0, 0.0)) and np.all(v1[:, 9] == 0.75) and np.all(v1[:, 10:13] == np.float32(0.8))
    assert np.allclose(np.linalg.norm(v0[:, 3:6], axis=1), 1.0, atol=1e-5)
    assert (s[:, 13:16] == 0.0).all()
    flat = sums.reshape(-1, 3).astype(np.float32)
    for row, p in zip(v0, pix[vert == 0]):
        assert np.array_equal(row[16:19], ol.xyz_to_srgb(flat[p]))
    t0 = {int(p): r[16:19] for p, r in zip(pix[vert == 0], v0)}
    ratios = []
    for p, r in zip(pix[vert == 1], v1):
        a, b = t0[int(p)].astype(np.float64), r[16:19].astype(np.float64)
        k = int(np.argmax(np.abs(b)))
        m = a[k] / b[k]
        assert np.allclose(a, m * b, rtol=1e-5, atol=1e-6 * abs(b[k]))
        ratios.append(m)
    assert 0.25 < min(ratios) and max(ratios) < 0.75 and abs(np.median(ratios) - 0.5) < 0.02


def test_spread_threshold_fraction():
    objs = floor_and_ceiling()
    h0 = 0.5
    cam = ap.Camera.make([0.0, h0, 0.0], [0.0, -1.0, 0.0001], 1.0, 128, 128)
    prm = ap.RenderParams.defaults(spp=1, max_depth=2, sky=0, use_nee=0)
    a0 = h0 * h0 / (4.0 * math.pi)
    _, _, _, nq = ol.OracleBVH(objs).render_nrc(cam, prm, constant_cache(0.5), *BOUNDS,
                                                 spread=2.0 * math.pi / a0, cull_margin=2 ** -12)
    frac = nq / (128 * 128)
    assert abs(frac - 2 ** -0.5) < 0.015, frac  # 4 standard errors
