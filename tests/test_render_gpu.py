"""GPU parity of the wavefront integrator (ap_render) against the oracle's
restatement of CalculateRadiance (core/render/render.cuh:460-999) on the same
seeds.

Bars:
  * primary hit ids: bit-exact (camera rays are pure FP32 arithmetic with the
    same operation order on both sides);
  * images: the GPU and the oracle draw the same random numbers per path,
    follow the same steps in the same FP32 operation order and take their
    transcendentals from the same portable implementation (ap_fmath.h), so
    the running mean and the FP64 per-pixel sums are bit-identical. The
    relMSE against a second seed is printed beside it for scale;
  * the FP64 per-pixel sums equal spp x the running mean to FP32 rounding,
    and odd/even halves average to the full mean;
  * progressive calls (frame_index advancing) equal one call of the summed spp;
  * the gray floor under a square light reproduces the closed-form irradiance;
  * hero-wavelength sampling (four wavelengths per path) is bit-identical to
    the oracle's restatement of it as well.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes
import test_render_cpu as cpu

pytestmark = pytest.mark.gpu


def rel_mse(a, b):
    a = a.reshape(-1, 3).astype(np.float64)
    b = b.reshape(-1, 3).astype(np.float64)
    return float((((a - b) ** 2) / (b ** 2 + 1e-2)).mean())


def render_gpu(scene, cam, prm, odd_even=False, hdr=None, sums=None):
    n = cam.width * cam.height
    d_hdr = ap.DeviceBuffer(n * 12)
    if hdr is None:
        d_hdr.zero()
    else:
        d_hdr = ap.DeviceBuffer.from_numpy(np.ascontiguousarray(hdr, np.float32))
    d_sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64) if sums is None else sums.reshape(-1))
    d_prim = ap.DeviceBuffer.from_numpy(np.full(n, -7, np.int32))
    d_odd = d_even = None
    if odd_even:
        d_odd, d_even = ap.DeviceBuffer(n * 12), ap.DeviceBuffer(n * 12)
        d_odd.zero()
        d_even.zero()
    scene.render(cam, prm, d_hdr, d_odd, d_even, d_sums, d_prim)
    ap.synchronize()
    shp = (cam.height, cam.width)
    out = {"hdr": d_hdr.to_numpy(np.float32, n * 3).reshape(*shp, 3),
           "sums": d_sums.to_numpy(np.float64, n * 3).reshape(*shp, 3),
           "prim": d_prim.to_numpy(np.int32, n).reshape(shp)}
    if odd_even:
        out["odd"] = d_odd.to_numpy(np.float32, n * 3).reshape(*shp, 3)
        out["even"] = d_even.to_numpy(np.float32, n * 3).reshape(*shp, 3)
    return out


def compare_with_oracle(objs, cam, prm, label):
    print(f"[render] {label}: {len(objs)} objects, {cam.width}x{cam.height}, spp {prm.spp}, depth {prm.max_depth}",
          flush=True)
    scene = ap.SceneRenderer(objs)
    margin = np.float32(scene.info().line_margin)
    g = render_gpu(scene, cam, prm)
    bvh = ol.OracleBVH(objs)
    o_hdr, o_sums, o_prim = bvh.render(cam, prm, cull_margin=margin)
    prm2 = ap.RenderParams.from_buffer_copy(prm)
    prm2.seed = prm.seed + 1
    o2_hdr, _, _ = bvh.render(cam, prm2, cull_margin=margin)
    assert np.array_equal(g["prim"], o_prim), f"{label}: primary ids differ"
    same = rel_mse(g["hdr"], o_hdr)
    noise = rel_mse(o2_hdr, o_hdr)
    close = float(np.all(np.abs(g["hdr"] - o_hdr) <= 1e-4 * np.abs(o_hdr) + 1e-6, axis=-1).mean())
    ident = float(np.all(g["hdr"] == o_hdr, axis=-1).mean())
    print(f"[render] {label}: relMSE same-seed {same:.3g}, seed-to-seed {noise:.3g}, pixels within 1e-4 {close:.4f}, "
          f"bit-identical {ident:.4f}", flush=True)
    assert noise > 0.0
    assert np.array_equal(g["hdr"], o_hdr), (same, close)
    assert np.array_equal(g["sums"], o_sums)
    scene.close()
    return g


def box_params(spp=16, depth=8, **kw):
    return ap.RenderParams.defaults(spp=spp, max_depth=depth, **kw)


BOX_CAM = dict(origin=[0.0, 0.0, 3.4], direction=[0.0, 0.0, -1.0], fov_deg=40.0)


def test_box_scene_matches_oracle(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 64, 64)
    compare_with_oracle(objs, cam, box_params(), "box scene, object materials")


def test_box_scene_reference_triangle_material(gpu):
    objs = ap.synth_box_scene(3, 1, 15.0)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 48, 48)
    prm = box_params(triangle_material=ap.TRIANGLE_MATERIAL_REFERENCE)
    compare_with_oracle(objs, cam, prm, "box scene, reference triangle GBuffer")


def test_box_scene_variants(gpu):
    objs = ap.synth_box_scene(5, 1, 15.0)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 32, 32)
    for kw, label in ((dict(use_nee=0), "no NEE"), (dict(nee_always=1, seed=9), "NEE every vertex"),
                      (dict(sky=0, rr_min_depth=1), "sky off, early roulette")):
        compare_with_oracle(objs, cam, box_params(spp=16, depth=8, **kw), label)


def test_deep_paths_and_ggx(gpu):
    objs = ap.synth_box_scene(7, 1, 15.0)
    ggx = np.arange(len(objs)) % 3 == 0
    ggx &= objs["shading_model"] == ap.MAT_DIFFUSE
    objs["shading_model"][ggx] = ap.MAT_GGX
    objs["roughness"][ggx] = np.linspace(0.05, 1.0, int(ggx.sum()), dtype=np.float32)
    objs["metallic"][ggx] = 0.3
    objs["specular"][ggx] = 0.5
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 32, 32)
    compare_with_oracle(objs, cam, box_params(spp=16, depth=48), "GGX mix, depth 48")


def test_heightfield_scene_matches_oracle(gpu):
    """The heightfield workload at a small grid: fBm field, icospheres,
    Lambert/GGX materials, four lights, depth 16 with roulette from depth 4."""
    objs = ap.synth_heightfield_scene(64, 16, 4, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 36)
    compare_with_oracle(objs, cam, ap.RenderParams.defaults(spp=8, max_depth=16, rr_min_depth=4),
                        "heightfield 64^2 + 16 spheres")


def test_sums_odd_even_and_progressive(gpu):
    objs = ap.synth_box_scene(1, 1, 15.0)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 40, 24)
    scene = ap.SceneRenderer(objs)
    full = render_gpu(scene, cam, box_params(spp=16), odd_even=True)
    assert np.allclose(full["sums"] / 16.0, full["hdr"], rtol=1e-4, atol=1e-5)
    assert np.allclose(0.5 * (full["odd"] + full["even"]), full["hdr"], rtol=1e-4, atol=1e-5)
    first = render_gpu(scene, cam, box_params(spp=8))
    second = render_gpu(scene, cam, box_params(spp=8, frame_index=8), hdr=first["hdr"], sums=first["sums"])
    assert np.array_equal(second["hdr"], full["hdr"])
    assert np.array_equal(second["sums"], full["sums"])
    scene.close()


def test_floor_irradiance_closed_form_on_gpu(gpu):
    objs = cpu.floor_and_light(0.5, 10.0, a=0.25, h=1.0)
    expected = 0.5 * 10.0 * 4 * cpu.corner_form_factor(0.25, 0.25)
    cam = ap.Camera.make([0.0, 0.5, 0.0], [0.0, -1.0, 0.001], 1.0, 16, 16)
    scene = ap.SceneRenderer(objs)
    for use_nee, spp in ((1, 2048), (0, 16384)):
        out = render_gpu(scene, cam, ap.RenderParams.defaults(spp=spp, max_depth=2, sky=0, use_nee=use_nee,
                                                              nee_always=use_nee))
        assert (out["prim"] >= 0).all() and (out["prim"] < len(objs) - 1).all()
        y = float(out["hdr"][..., 1].mean())
        print(f"[render] floor use_nee={use_nee}: Y {y:.5f} expected {expected:.5f}", flush=True)
        assert abs(y / expected - 1.0) < 0.03, (use_nee, y, expected)
    scene.close()


def test_render_rejects_unsupported_and_bad_arguments(gpu):
    objs = ap.synth_box_scene(1, 0, 15.0)
    scene = ap.SceneRenderer(objs)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 8, 8)
    hdr = ap.DeviceBuffer(8 * 8 * 12)
    with pytest.raises(ap.ApertureError) as e:
        scene.render(cam, ap.RenderParams.defaults(triangle_material=7), hdr)
    assert e.value.code == ap.ERR_INVALID_ARG
    bad = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 0, 8)
    with pytest.raises(ap.ApertureError):
        scene.render(bad, ap.RenderParams.defaults(), hdr)
    textured = objs.copy()
    textured["textures"][0] = (1, 0, 0, 0)   # an albedo texture handle, and no texture cache bound
    gscene = ap.SceneRenderer(textured)
    with pytest.raises(ap.ApertureError) as e:
        gscene.render(cam, ap.RenderParams.defaults(), hdr)
    assert e.value.code == ap.ERR_INVALID_ARG
    with pytest.raises(ap.ApertureError) as e:
        scene.render(cam, ap.RenderParams.defaults(wavelengths=3), hdr)
    assert e.value.code == ap.ERR_INVALID_ARG
    gscene.close()
    scene.close()


def test_hero_wavelengths_match_oracle(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 48, 48)
    compare_with_oracle(objs, cam, box_params(spp=8, wavelengths=4), "box scene, four wavelengths per path")
    hf = ap.synth_heightfield_scene(64, 16, 4, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 48, 27)
    compare_with_oracle(hf, cam, ap.RenderParams.defaults(spp=4, max_depth=16, rr_min_depth=4, wavelengths=4),
                        "heightfield 64^2, four wavelengths per path")


def test_sphere_lights_match_oracle(gpu):
    """Emissive spheres as the only lights (NEESampleSphereCapArea and the
    paired NEEPDFOmegaFromHit_Sphere, nee.cuh:254-310), with GGX surfaces; then
    a table of the cube light and two spheres, one and four wavelengths."""
    cam = ap.Camera.make(BOX_CAM["origin"], BOX_CAM["direction"], BOX_CAM["fov_deg"], 48, 48)
    objs = scenes.sphere_lit_box(2)
    compare_with_oracle(objs, cam, box_params(spp=8, depth=8), "sphere lights only")
    mixed = scenes.sphere_lit_box(3, keep_cube_light=True)
    compare_with_oracle(mixed, cam, box_params(spp=8, depth=8, seed=4), "cube and sphere lights")
    compare_with_oracle(mixed, cam, box_params(spp=6, depth=8, wavelengths=4, nee_always=1),
                        "cube and sphere lights, four wavelengths, NEE at every vertex")
