"""CPU tests pinning the oracle integrator (oracle/or_render.cpp) with
closed forms of the reference's estimator (core/render/render.cuh:460-999):

  * a camera facing the emissive quad sees its radiance: Y = Le (D65Norm
    normalises Y to one), X/Y and Z/Y the D65 white point;
  * a gray Lambertian floor under a square light, sky off and depth 2, gives
    the MIS-combined direct illumination L = rho * Le * 4 F(a/h, a/h) with F
    the point-to-rectangle form factor; the NEE-only and BSDF-only limits of
    the same estimator converge to the same value;
  * the FP64 sums agree with the FP32 running mean; seeds change the noise
    only;
  * hero-wavelength sampling (four wavelengths per path, the core's
    extension) keeps the closed forms and cuts the colour noise of the
    light seen directly by a large factor.
"""
import math

import numpy as np

import aperture as ap
import oracle_lib as ol
import scenes

D65_XZ = (0.9505, 1.089)


def corner_form_factor(X, Y):
    """Differential area to a parallel rectangle, receiver under one corner."""
    a = X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
    b = Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))
    return (a + b) / (2 * math.pi)


def floor_and_light(rho, le, a=0.25, h=1.0):
    n = 8
    xs = np.linspace(-2.0, 2.0, n + 1, dtype=np.float32)
    p0, p1, p2 = [], [], []
    for j in range(n):
        for i in range(n):
            A = (xs[i], 0.0, xs[j]); B = (xs[i + 1], 0.0, xs[j]); C = (xs[i + 1], 0.0, xs[j + 1]); D = (xs[i], 0.0, xs[j + 1])
            p0 += [A, A]; p1 += [C, D]; p2 += [B, C]
    floor = scenes.triangles(p0, p1, p2)
    floor["v0_normal"] = floor["v1_normal"] = floor["v2_normal"] = (0.0, 1.0, 0.0)
    floor["albedo"] = rho
    light = scenes.cubes([(0.0, h + 0.001, 0.0)], (a, 0.001, a), emissive=le)
    return scenes.concat([floor, light])


def test_direct_view_of_light():
    objs = floor_and_light(0.5, 15.0, a=2.0, h=1.0)
    cam = ap.Camera.make([0.0, 0.2, 0.0], [0.0, 1.0, 0.001], 10.0, 8, 8)
    prm = ap.RenderParams.defaults(spp=4096, max_depth=1, sky=0)
    hdr, sums, prim = ol.OracleBVH(objs).render(cam, prm, cull_margin=2 ** -12)
    assert (prim == len(objs) - 1).all()
    xyz = hdr.reshape(-1, 3).mean(0)
    assert abs(xyz[1] / 15.0 - 1.0) < 0.02, xyz
    assert abs(xyz[0] / xyz[1] - D65_XZ[0]) < 0.02 and abs(xyz[2] / xyz[1] - D65_XZ[1]) < 0.02


def render_floor(use_nee, nee_always, spp, seed=0, wavelengths=1):
    objs = floor_and_light(0.5, 10.0, a=0.25, h=1.0)
    cam = ap.Camera.make([0.0, 0.5, 0.0], [0.0, -1.0, 0.001], 1.0, 4, 4)
    prm = ap.RenderParams.defaults(spp=spp, max_depth=2, sky=0, use_nee=use_nee, nee_always=nee_always, seed=seed,
                                   wavelengths=wavelengths)
    hdr, sums, prim = ol.OracleBVH(objs).render(cam, prm, cull_margin=2 ** -12)
    assert (prim >= 0).all() and (prim < len(objs) - 1).all()  # every camera ray lands on the floor
    return hdr, sums


def test_floor_irradiance_closed_form():
    expected = 0.5 * 10.0 * 4 * corner_form_factor(0.25, 0.25)
    hdr, sums = render_floor(1, 1, 8192)
    y = hdr.reshape(-1, 3)[:, 1].mean()
    assert abs(y / expected - 1.0) < 0.03, (y, expected)
    assert np.allclose(sums / 8192, hdr, rtol=1e-3, atol=1e-6)


def test_nee_and_bsdf_limits_agree():
    nee, _ = render_floor(1, 1, 4096, seed=1)
    bsdf, _ = render_floor(0, 0, 65536, seed=2)
    a, b = nee.reshape(-1, 3)[:, 1].mean(), bsdf.reshape(-1, 3)[:, 1].mean()
    assert abs(a / b - 1.0) < 0.05, (a, b)


def test_seed_changes_noise_not_mean():
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 32, 32)
    bvh = ol.OracleBVH(objs)
    m = []
    for seed in (0, 1):
        prm = ap.RenderParams.defaults(spp=32, max_depth=8, nee_always=1, seed=seed)
        hdr, _, _ = bvh.render(cam, prm, cull_margin=2 ** -12)
        m.append(hdr)
    assert not np.array_equal(m[0], m[1])
    # Pixels are independent estimates: the mean of the per-pixel difference
    # must sit within a few standard errors of zero.
    d = (m[0] - m[1]).reshape(-1, 3).sum(1)
    assert abs(d.mean()) < 4.0 * d.std() / math.sqrt(len(d)), (d.mean(), d.std())


def test_hero_wavelengths_keep_the_closed_form():
    expected = 0.5 * 10.0 * 4 * corner_form_factor(0.25, 0.25)
    hdr, sums = render_floor(1, 1, 2048, seed=3, wavelengths=4)
    y = hdr.reshape(-1, 3)[:, 1].mean()
    assert abs(y / expected - 1.0) < 0.03, (y, expected)
    bsdf, _ = render_floor(0, 0, 8192, seed=4, wavelengths=4)
    b = bsdf.reshape(-1, 3)[:, 1].mean()
    assert abs(b / expected - 1.0) < 0.08, (b, expected)


def test_hero_wavelengths_cut_colour_noise():
    objs = floor_and_light(0.5, 15.0, a=2.0, h=1.0)
    cam = ap.Camera.make([0.0, 0.2, 0.0], [0.0, 1.0, 0.001], 10.0, 16, 16)
    bvh = ol.OracleBVH(objs)
    out = {}
    for nl in (1, 4):
        prm = ap.RenderParams.defaults(spp=4, max_depth=1, sky=0, wavelengths=nl, seed=nl)
        hdr, _, _ = bvh.render(cam, prm, cull_margin=2 ** -12)
        px = hdr.reshape(-1, 3).astype(np.float64)
        out[nl] = (px.mean(0), px.var(0))
        assert abs(px[:, 1].mean() / 15.0 - 1.0) < 0.03, (nl, px.mean(0))
    # Every pixel sees the same emitter: the spread across pixels is the
    # spectral-sampling noise alone.
    ratio = out[1][1] / out[4][1]
    assert (ratio > 8.0).all(), ratio
