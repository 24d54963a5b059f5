"""GPU checks of glass and media against the oracle (oracle/or_render.cpp):
dielectric surfaces with Abbe dispersion and tilted Beer-Lambert absorption,
the solo trace of a ray inside a medium, misses inside a medium (which
restart at the world origin) and total internal reflection.

Bars: images, FP64 sums and primary ids bit-identical to the oracle's for a
closed box holding a dispersive glass sphere, a rotated high-IOR glass box,
a glass torus and a thin glass quad, in both triangle material modes, with
one and four wavelengths; the same with the camera inside a glass sphere and
inside a glass sphere that encloses the whole box; under directional guiding
over several frames; in a batch next to other scenes (bit-identical to the
scene-by-scene renders); with the radiance cache (frozen), the oracle's
cache renderer to the bars of its own tests. With four
wavelengths, whose secondaries end at a dispersive interface, the image
agrees in expectation with the one-wavelength estimator.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes
import test_batch_gpu as bg
import test_guiding_gpu as gg
import test_nrc_render_gpu as nr
import test_sdf_cpu as sc

pytestmark = pytest.mark.gpu


def make_glass(o, ior, sigma_a, albedo, abbe, tilt):
    o["shading_model"] = ap.MAT_GLASS
    o["ior"] = ior
    o["sigma_a"] = sigma_a
    o["albedo"] = albedo
    o["abbe_number"] = abbe
    o["tilt_number"] = tilt
    return o


def glass_objects():
    sph = make_glass(scenes.spheres([[-0.4, -0.5, -0.1]], [0.33]), 1.5, 1.5, [0.85, 0.95, 1.0], 20.0, 0.0)
    box = scenes.cubes([[0.45, -0.55, 0.15]], [0.22, 0.3, 0.18])
    r = sc.rot([0.3, 1.0, 0.2], 0.7)
    box["object_to_world"] = r
    box["world_to_object"] = r.T
    box = make_glass(box, 2.4, 3.0, [1.0, 0.7, 0.4], 0.0, 0.6)    # strong internal reflection, no dispersion
    tor = make_glass(scenes.sdfs([[0.05, 0.2, -0.3]], [0.3, 0.3, 0.3], 0, (0.35, 0.12, 0, 0),
                                 rotations=[sc.rot([1, 0, 0], 0.9)]), 1.33, 0.5, [0.9, 1.0, 0.9], 45.0, -0.4)
    quad = scenes.triangles([[-0.7, 0.1, 0.3], [-0.7, 0.1, 0.3]], [[-0.1, 0.1, 0.3], [-0.1, 0.7, 0.3]],
                            [[-0.1, 0.7, 0.3], [-0.7, 0.7, 0.3]])
    quad = make_glass(quad, 1.45, 2.0, [1.0, 1.0, 1.0], 30.0, 0.0)   # a thin surface: misses inside the medium
    return scenes.concat([sph, box, tor, quad])


def glass_scene():
    return scenes.concat([ap.synth_box_scene(1, 2, 15.0), glass_objects()])


def enclosed_scene():
    """The box inside a large glass sphere: the camera starts inside the medium."""
    shell = make_glass(scenes.spheres([[0.0, 0.0, 0.0]], [5.0]), 1.2, 0.05, [0.9, 0.9, 1.0], 25.0, 0.0)
    return scenes.concat([ap.synth_box_scene(1, 2, 15.0), glass_objects(), shell])


def gpu_render(objs, cam, prm):
    scene = ap.SceneRenderer(objs)
    n = cam.width * cam.height
    hdr = ap.DeviceBuffer(n * 12)
    hdr.zero()
    sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
    prim = ap.DeviceBuffer.from_numpy(np.full(n, -1, np.int32))
    scene.render(cam, prm, hdr, sum_xyz=sums, primary_id=prim)
    ap.synchronize()
    out = (hdr.to_numpy(np.float32, n * 3).reshape(cam.height, cam.width, 3),
           sums.to_numpy(np.float64, n * 3).reshape(cam.height, cam.width, 3),
           prim.to_numpy(np.int32, n).reshape(cam.height, cam.width))
    margin = np.float32(scene.info().line_margin)
    scene.close()
    return out, margin


def parity(objs, cam, prm, label, n_glass_first=0):
    (g_hdr, g_sums, g_prim), margin = gpu_render(objs, cam, prm)
    o_hdr, o_sums, o_prim = ol.OracleBVH(objs).render(cam, prm, cull_margin=margin)
    glass_ids = np.flatnonzero(objs["shading_model"] == ap.MAT_GLASS)
    seen = int(np.isin(g_prim, glass_ids).sum())
    ident = float(np.all(g_sums == o_sums, axis=-1).mean())
    print(f"[glass] {label}: {seen} pixels see glass first, pixels bit-identical {ident:.4f}, "
          f"mean {g_hdr.mean(axis=(0, 1))}", flush=True)
    assert np.array_equal(g_prim, o_prim)
    assert np.array_equal(g_hdr, o_hdr) and np.array_equal(g_sums, o_sums)
    assert seen >= n_glass_first
    assert np.isfinite(g_hdr).all() and g_hdr.max() > 0.0


@pytest.mark.parametrize("tri_mode,wavelengths", [(0, 1), (1, 1), (0, 4)])
def test_glass_scene_matches_the_oracle(gpu, tri_mode, wavelengths):
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 48, 40)
    prm = ap.RenderParams.defaults(spp=3, max_depth=14, rr_min_depth=4, seed=21, wavelengths=wavelengths,
                                   triangle_material=tri_mode)
    parity(glass_scene(), cam, prm, f"box with glass, triangle mode {tri_mode}, {wavelengths} wavelengths", 200)


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_cameras_inside_a_medium_match_the_oracle(gpu, wavelengths):
    prm = ap.RenderParams.defaults(spp=3, max_depth=14, rr_min_depth=4, seed=5, wavelengths=wavelengths)
    cam = ap.Camera.make([-0.4, -0.5, -0.1], [0.3, 0.2, -1], 70.0, 32, 24)   # inside the glass sphere
    parity(glass_scene(), cam, prm, f"camera inside a glass sphere, {wavelengths} wavelengths", 200)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 32, 24)              # inside the enclosing shell
    parity(enclosed_scene(), cam, prm, f"box inside a glass shell, {wavelengths} wavelengths")
    # Inside a diffuse shell the path starts inside a medium: after the first
    # hit the next ray meets that object alone and NEE is skipped
    # (render.cuh:89-92, 782-785); a miss against it leaves the medium and
    # restarts at the world origin (render.cuh:109-126).
    shell = scenes.spheres([[0.0, 0.0, 0.0]], [5.0])
    shell["albedo"] = [0.7, 0.7, 0.7]
    parity(scenes.concat([ap.synth_box_scene(1, 2, 15.0), shell]), cam, prm,
           f"box inside a diffuse shell, {wavelengths} wavelengths")


def sphere_light_glass_scene():
    """Glass balls around a small sphere light with GGX plates behind them,
    inside the closed box without its ceiling light (the layout of the
    reference's CreateGlassTestScene, scene_helper.h:176-262, scaled 2x)."""
    box = ap.synth_box_scene(1, 2, 15.0)
    box = box[box["shading_model"] != ap.MAT_LIGHT]
    balls = scenes.spheres([[0.0, 0.5, 0.25], [-0.5, 0.0, -0.25], [0.5, 0.0, 0.25], [0.0, 0.0, 0.0],
                            [0.0, -0.5, -0.25]], [0.25] * 5)
    balls = make_glass(balls, 1.5, 1.0, [[1, 1, 1], [1, 1, 0], [0, 1, 1], [1, 0.02, 0.02], [0.02, 0.02, 1]],
                       20.0, 0.0)
    balls["ior"] = [1.6, 1.5, 1.55, 1.45, 1.5]
    light = scenes.spheres([[0.25, 0.5, -0.25]], [0.1], emissive=[25.0, 25.0, 25.0])
    plates = scenes.cubes([[0.0, 0.0, -0.5], [0.0, -0.5, -0.6], [0.0, 0.5, -0.6]], [0.25, 0.25, 0.02])
    plates["shading_model"] = ap.MAT_GGX
    plates["albedo"] = [[1.0, 0.5, 0.0], [0.9, 0.9, 0.9], [0.5, 0.5, 0.5]]
    plates["roughness"] = [0.1, 0.1, 0.5]
    plates["metallic"] = [1.0, 0.0, 1.0]
    plates["specular"] = [1.0, 0.5, 1.0]
    return scenes.concat([box, balls, light, plates])


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_sphere_light_among_glass_matches_the_oracle(gpu, wavelengths):
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 40, 40)
    prm = ap.RenderParams.defaults(spp=3, max_depth=14, rr_min_depth=4, seed=8, wavelengths=wavelengths)
    parity(sphere_light_glass_scene(), cam, prm, f"glass balls around a sphere light, {wavelengths} wavelengths",
           200)


def test_guided_glass_matches_the_oracle(gpu):
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 32, 24)
    gg.guided_parity(glass_scene(), cam, dict(spp=3, max_depth=12, rr_min_depth=4), "box with glass", calls=2)


def test_batch_with_glass_equals_single_scene_renders(gpu):
    objs = [glass_scene(), ap.synth_box_scene(1, 2, 15.0), enclosed_scene()]
    cams = [ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 40, 32), ap.Camera.make([0.2, 0.1, 3.0], [0, 0, -1], 50.0, 16, 16),
            ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 24, 20)]
    scenes_ = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes_, cams)
    stream = ap.Stream()
    for wl in (1, 4):
        prm = ap.RenderParams.defaults(spp=2, max_depth=12, rr_min_depth=4, seed=6, wavelengths=wl)
        got, _ = bg.batch_render(batch, cams, prm, stream)
        bg.check_against_single(scenes_, cams, prm, got, f"glass batch, {wl} wavelengths")
    batch.close()
    for s in scenes_:
        s.close()


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_radiance_cache_with_glass_matches_the_oracle(gpu, wavelengths):
    # Glass vertices neither end into the cache nor record training samples,
    # and the spread sum skips the segment after one (its pdf is a delta).
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 48, 40)
    prm = ap.RenderParams.defaults(spp=2, max_depth=12, rr_min_depth=4, seed=7, wavelengths=wavelengths)
    scene, cache = nr.frozen_parity(glass_scene(), cam, prm, f"box with glass, {wavelengths} wavelengths")
    scene.close()
    cache.close()


def test_hero_wavelengths_agree_in_expectation(gpu):
    # Dispersive glass ends the secondary wavelengths; the four-wavelength
    # estimate must keep the one-wavelength estimator's expectation. Eight
    # independent renders per mode give the spread of the image mean.
    objs = glass_scene()
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 48, 40)
    means = {1: [], 4: []}
    for wl in (1, 4):
        for k in range(8):
            prm = ap.RenderParams.defaults(spp=48, max_depth=24, rr_min_depth=4, seed=100 + k, wavelengths=wl)
            (hdr, _, _), _ = gpu_render(objs, cam, prm)
            means[wl].append(hdr.mean(axis=(0, 1)))
    m1, m4 = np.array(means[1]), np.array(means[4])
    se = np.sqrt(m1.var(axis=0, ddof=1) / 8 + m4.var(axis=0, ddof=1) / 8)
    print(f"[glass] image mean, 1 wavelength {m1.mean(0)}, 4 wavelengths {m4.mean(0)}, standard error {se}",
          flush=True)
    assert np.all(np.abs(m1.mean(0) - m4.mean(0)) < 5 * se + 2e-3)
