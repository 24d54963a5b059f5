"""CPU tests pinning the oracle's restatement of glass and media
(oracle/or_render.cpp: IORFromAbbeNM, SigmaAFromTiltNM, BeerLambert_Spectrum,
FresnelDielectricExact, refract, SampleBSDF / EvaluateBSDFWithCos of
MAT_GLASS, the MediaContext of CalculateRadiance) with properties fixed
independently of it:

  * the dispersion law returns nd at the d line and recovers the Abbe number
    (nd - 1) / (nF - nC); a non-positive Abbe number means no dispersion;
    the absorption tilt changes sigma_a by sigma_a * tilt between the C and F
    lines;
  * the Fresnel term equals the textbook unpolarised reflectance (Rs + Rp) / 2
    of Snell's law in FP64, ((n - 1) / (n + 1))^2 at normal incidence and
    Rs / 2 at Brewster's angle; Beer-Lambert equals resp^(d sigma) in FP64;
  * sampled refractions obey Snell's law in the plane of incidence, sampled
    reflections are mirror images, reflections are drawn with frequency Fr,
    total internal reflection always reflects with weight 0.99 T, and the
    refraction weight over its pdf is eta^2 cosT / cosI times the
    transmittance;
  * in the integrator, a glass slab seen at normal incidence in front of an
    emitter passes (1 - F) / (1 + F) of its radiance (all internal
    reflection orders), and a glass sphere of IOR 1 leaves a closed box's
    image unchanged in expectation.
"""
import numpy as np

import aperture as ap
import oracle_lib as ol
import scenes

LF, LD, LC = 486.1327, 587.5618, 656.2725


def test_dispersion_recovers_the_abbe_number():
    for nd, vd in ((1.5, 20.0), (1.7, 35.0), (2.4, 55.0), (1.33, 5.0)):
        assert ol.glass_ior(np.float32(LD), nd, vd) == np.float32(nd)
        nF, nC = ol.glass_ior(LF, nd, vd), ol.glass_ior(LC, nd, vd)
        assert nF > nd > nC
        assert abs((nd - 1.0) / (nF - nC) - vd) < 2e-3 * vd
        lams = np.linspace(380, 780, 41)
        ns = np.array([ol.glass_ior(l, nd, vd) for l in lams])
        assert np.all(np.diff(ns) < 0)
        for v in (0.0, -3.0):
            assert all(ol.glass_ior(l, nd, v) == np.float32(nd) for l in lams)


def test_absorption_tilt():
    for sa, tilt in ((10.0, 0.5), (3.0, -0.8), (1.0, 0.0)):
        assert abs(ol.glass_sigma(sa, LD, tilt) - sa) < 1e-6 * sa
        assert abs((ol.glass_sigma(sa, LF, tilt) - ol.glass_sigma(sa, LC, tilt)) - sa * tilt) < 2e-5 * sa
    assert ol.glass_sigma(5.0, 380.0, -4.0) == 0.0   # clamped at zero


def test_fresnel_is_the_textbook_reflectance():
    g = np.random.default_rng(3)
    for _ in range(400):
        n1, n2 = g.uniform(1.0, 2.5, 2)
        eta = n1 / n2
        sin_i = g.uniform(0.0, 1.0)
        sin_t = eta * sin_i
        if sin_t >= 1.0:
            continue
        ci, ct = np.sqrt(1 - sin_i ** 2), np.sqrt(1 - sin_t ** 2)
        rs = ((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2
        rp = ((n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)) ** 2
        assert abs(ol.glass_fresnel(ci, ct, eta) - 0.5 * (rs + rp)) < 2e-6
    for n in (1.33, 1.5, 2.4):
        assert abs(ol.glass_fresnel(1.0, 1.0, 1.0 / n) - ((n - 1) / (n + 1)) ** 2) < 1e-7
        tb = np.arctan(n)   # Brewster: Rp = 0
        ci, ct = np.cos(tb), np.cos(np.arcsin(np.sin(tb) / n))
        rs = ((ci - n * ct) / (ci + n * ct)) ** 2
        assert abs(ol.glass_fresnel(ci, ct, 1.0 / n) - 0.5 * rs) < 2e-6


def test_beer_lambert():
    for d, resp, sa in ((0.3, 0.5, 2.0), (1.7, 0.9, 0.4), (0.0, 0.2, 5.0), (2.0, 1.0, 9.0), (1.0, 1e-20, 1.0)):
        want = max(resp, 1e-12) ** (d * sa)
        assert abs(ol.glass_beer(d, resp, sa) - want) < 2e-6 * max(want, 1e-30) + 1e-30


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_sampled_directions_and_weights():
    g = np.random.default_rng(5)
    refl = 0
    trials = 0
    fr_sum = 0.0
    for k in range(3000):
        n = unit(g.normal(size=3))
        d = unit(-n + g.normal(size=3) * 0.6)
        if np.dot(d, n) >= -0.05:
            continue
        inside = bool(k % 2)
        ior, lam = g.uniform(1.2, 2.0), g.uniform(400, 700)
        nd = np.float32(n)
        nxt, pdf, brdf, inside_after, delta = ol.glass_bsdf(d, n, lam, ior=ior, inside=inside, seed=k)
        assert delta
        eta = ior if inside else 1.0 / ior
        ci = -np.dot(d, n)
        s2t = eta * eta * (1 - ci * ci)
        if s2t >= 1.0:                        # total internal reflection
            assert pdf == 1.0 and inside_after == inside
            assert abs(brdf - 0.99) < 1e-6
            assert np.allclose(nxt, d - 2 * np.dot(d, n) * n, atol=2e-6)
            continue
        ct = np.sqrt(1 - s2t)
        fr = ol.glass_fresnel(np.float32(ci), np.float32(ct), np.float32(eta))
        fr_sum += fr
        trials += 1
        if inside_after == inside:            # reflection
            refl += 1
            assert abs(pdf - fr) < 1e-4 and abs(brdf / pdf - 1.0) < 1e-4
            assert np.allclose(nxt, d - 2 * np.dot(d, n) * n, atol=2e-6)
        else:                                 # refraction: Snell, plane of incidence, far side
            t = np.asarray(nxt, np.float64)
            assert abs(np.linalg.norm(t) - 1.0) < 1e-5
            assert np.dot(t, n) < 0
            assert abs(np.dot(np.cross(d, n), t)) < 1e-5
            sin_t = np.linalg.norm(np.cross(t, n))
            assert abs(sin_t - eta * np.sqrt(1 - ci * ci)) < 2e-5
            assert abs((brdf / pdf) - eta * eta * ct / ci) < 2e-4 * eta * eta * ct / ci
    assert trials > 1000
    # Reflections are drawn with probability Fr: binomial count within 5 sigma.
    assert abs(refl - fr_sum) < 5 * np.sqrt(fr_sum) + 5


def test_transmittance_inside_the_medium():
    # Inside the medium the weight carries the Beer-Lambert factor over the
    # distance from the vertex the ray left, at the wavelength's response.
    n = np.array([0.0, 0.0, 1.0])
    d = unit([0.2, 0.1, 1.0])           # leaving through the top face from inside
    for lam in (450.0, 550.0, 650.0):
        for albedo, sa, tilt in (((0.8, 0.5, 0.3), 2.0, 0.0), ((0.3, 0.9, 0.6), 0.7, 0.5)):
            nxt, pdf, brdf, _, _ = ol.glass_bsdf(d, n, lam, ior=1.5, sigma_a=sa, tilt=tilt, albedo=albedo,
                                                 inside=True, vertex=(0.0, 0.0, -1.25), hit=(0.0, 0.0, 0.0), seed=11)
            _, pdf0, brdf0, _, _ = ol.glass_bsdf(d, n, lam, ior=1.5, sigma_a=sa, tilt=tilt, albedo=(1, 1, 1),
                                                 inside=True, vertex=(0.0, 0.0, -1.25), hit=(0.0, 0.0, 0.0), seed=11)
            resp = float(ol.interact([lam], [albedo])[0])
            want = max(resp, 1e-12) ** (1.25 * ol.glass_sigma(sa, lam, tilt))
            assert pdf == pdf0
            assert abs(brdf / brdf0 - want) < 1e-5 * max(want, 1e-6)


def slab_scene(with_slab):
    light = scenes.cubes([[0.0, 0.0, -3.0]], [4.0, 4.0, 0.1], emissive=[3.0, 3.0, 3.0])
    parts = [light]
    if with_slab:
        slab = scenes.cubes([[0.0, 0.0, 0.0]], [2.0, 2.0, 0.2])
        slab["shading_model"] = ap.MAT_GLASS
        slab["ior"] = 1.5
        slab["abbe_number"] = 0.0     # no dispersion: one Fresnel term for every wavelength
        slab["sigma_a"] = 3.0         # a white albedo absorbs nothing
        parts.append(slab)
    return scenes.concat(parts)


def test_slab_passes_one_minus_f_over_one_plus_f():
    cam = ap.Camera.make([0, 0, 2.0], [0, 0, -1], 0.5, 8, 8)
    prm = ap.RenderParams.defaults(spp=3000, max_depth=40, rr_min_depth=4, seed=3, sky=0)
    with_slab = ol.OracleBVH(slab_scene(True)).render(cam, prm)[0].mean(axis=(0, 1))
    without = ol.OracleBVH(slab_scene(False)).render(cam, prm)[0].mean(axis=(0, 1))
    F = ((1.5 - 1.0) / (1.5 + 1.0)) ** 2
    ratio = with_slab / without
    print("slab ratio", ratio, "want", (1 - F) / (1 + F))
    assert np.all(np.abs(ratio - (1 - F) / (1 + F)) < 0.012)


def test_ior_one_sphere_is_invisible_in_expectation():
    box = ap.synth_box_scene(1, 2, 15.0)
    sph = scenes.spheres([[0.0, -0.4, 0.2]], [0.45])
    sph["shading_model"] = ap.MAT_GLASS
    sph["ior"] = 1.0
    sph["sigma_a"] = 5.0
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 16, 16)
    prm = ap.RenderParams.defaults(spp=600, max_depth=40, rr_min_depth=4, seed=9)
    a = ol.OracleBVH(scenes.concat([box, sph])).render(cam, prm)[0].mean(axis=(0, 1))
    b = ol.OracleBVH(box).render(cam, prm)[0].mean(axis=(0, 1))
    print("IOR-1 sphere", a, "without", b)
    assert np.all(np.abs(a / b - 1.0) < 0.02)
