"""GPU checks of mesh and texture ingestion against the oracle.

Bars: the device texture sampler (ap_texture_sample) returns the oracle's
samples bit for bit at random, seam, out-of-range and non-finite UVs for
images of 1x1, 3x5 and 64x32 texels; a box holding a mesh read by
ap_mesh_load_obj (quads, UVs, normals) with an albedo texture, a textured
cube (which samples UV (0, 0)) and a textured glass sphere renders images,
FP64 sums and primary ids bit-identical to the oracle in both triangle
material modes, with one and four wavelengths; the same scene in a batch is
bit-identical to its single-scene render; a material naming a texture the
bound cache does not hold is refused.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes
import test_batch_gpu as bg
import test_glass_gpu as tg
import test_mesh_cpu as mc

pytestmark = pytest.mark.gpu


def images():
    return [mc.checker(1, 1, 1), mc.checker(3, 5, 2), mc.checker(64, 32, 3)]


def test_texture_sampler_matches_the_oracle(gpu):
    imgs = images()
    cache = ap.TextureCache()
    handles = [cache.add_rgba8(i) for i in imgs]
    assert handles == [1, 2, 3] and cache.count() == 3
    ol.texture_set(imgs)
    g = np.random.default_rng(8)
    uv = np.concatenate([g.uniform(-4, 4, (4000, 2)), g.uniform(0, 1, (4000, 2)),
                         [[0, 0], [1, 1], [-0.0, 1e-9], [0.999999, 0.5], [np.nan, 0.2], [np.inf, -np.inf],
                          [1e30, -1e30], [-1e-9, -1e-9]]]).astype(np.float32)
    for k, h in enumerate(handles):
        got = cache.sample(h, uv)
        want = ol.texture_sample(k, uv)
        print(f"[mesh] texture {imgs[k].shape[1]}x{imgs[k].shape[0]}: {int((got != want).any(1).sum())} "
              f"of {len(uv)} samples differ", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with pytest.raises(ap.ApertureError):
        cache.sample(4, uv[:4])
    with pytest.raises(ap.ApertureError):
        cache.add_rgba8(np.zeros((0, 4, 4), np.uint8))     # no rows
    cache.close()
    ol.texture_set([])


def grid_obj(n=6):
    """An n x n grid of quads, a panel before the box's back wall, with UVs and normals."""
    lines = []
    for j in range(n + 1):
        for i in range(n + 1):
            x, y = -0.9 + 1.8 * i / n, -0.9 + 1.8 * j / n
            z = -0.93 + 0.04 * np.sin(i * 1.3 + j * 0.7)
            lines.append(f"v {x:.6f} {y:.6f} {z:.6f}")
            lines.append(f"vt {1.7 * i / n:.6f} {1.3 * j / n:.6f}")
    lines.append("vn 0 0 1")
    lines.append("vn 0.1 0 0.99")
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i + 1
            b, c, d = a + 1, a + n + 2, a + n + 1
            lines.append(f"f {a}/{a}/{1 + (i + j) % 2} {d}/{d}/1 {c}/{c}/2 {b}/{b}/1")
    return "\n".join(lines) + "\n"


def textured_scene(tmp_path):
    img = mc.checker(64, 32, 3)
    path = tmp_path / "grid.obj"
    path.write_text(grid_obj())
    mesh = ap.load_obj(str(path), (0.0, 0.0, 0.0), 1.0, (1, 0, 0, 0))
    cube = scenes.cubes([[0.5, -0.6, -0.3]], [0.2, 0.25, 0.2])
    cube["textures"][:, 0] = 1
    sph = tg.make_glass(scenes.spheres([[-0.4, -0.55, 0.1]], [0.3]), 1.5, 2.0, [1, 1, 1], 20.0, 0.0)
    sph["textures"][:, 0] = 1
    objs = scenes.concat([ap.synth_box_scene(1, 2, 15.0), mesh, cube, sph])
    return objs, img, len(mesh)


def render_with_texture(objs, img, cam, prm):
    cache = ap.TextureCache()
    cache.add_rgba8(img)
    scene = ap.SceneRenderer(objs)
    scene.set_texture_cache(cache)
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
    cache.close()
    return out, margin


@pytest.mark.parametrize("tri_mode,wavelengths", [(0, 1), (1, 1), (0, 4)])
def test_textured_mesh_matches_the_oracle(gpu, tmp_path, tri_mode, wavelengths):
    objs, img, n_mesh = textured_scene(tmp_path)
    ol.texture_set([img])
    cam = ap.Camera.make([0, 0.2, 3.4], [0, -0.25, -1], 45.0, 48, 40)
    prm = ap.RenderParams.defaults(spp=3, max_depth=12, rr_min_depth=4, seed=31, wavelengths=wavelengths,
                                   triangle_material=tri_mode)
    (g_hdr, g_sums, g_prim), margin = render_with_texture(objs, img, cam, prm)
    o_hdr, o_sums, o_prim = ol.OracleBVH(objs).render(cam, prm, cull_margin=margin)
    n_box = len(objs) - n_mesh - 2
    mesh_px = int(((g_prim >= n_box) & (g_prim < n_box + n_mesh)).sum())
    ident = float(np.all(g_sums == o_sums, axis=-1).mean())
    print(f"[mesh] textured grid ({n_mesh} triangles), triangle mode {tri_mode}, {wavelengths} wavelengths: "
          f"{mesh_px} pixels see the mesh first, pixels bit-identical {ident:.4f}", flush=True)
    assert mesh_px > 300
    assert np.array_equal(g_prim, o_prim)
    assert np.array_equal(g_hdr, o_hdr) and np.array_equal(g_sums, o_sums)
    ol.texture_set([])


def test_batch_with_textures_equals_single_render(gpu, tmp_path):
    objs, img, _ = textured_scene(tmp_path)
    cache = ap.TextureCache()
    cache.add_rgba8(img)
    scs = [ap.SceneRenderer(objs), ap.SceneRenderer(ap.synth_box_scene(1, 2, 15.0))]
    scs[0].set_texture_cache(cache)
    cams = [ap.Camera.make([0, 0.2, 3.4], [0, -0.25, -1], 45.0, 40, 32),
            ap.Camera.make([0.2, 0.1, 3.0], [0, 0, -1], 50.0, 16, 16)]
    batch = ap.Batch(scs, cams)
    stream = ap.Stream()
    prm = ap.RenderParams.defaults(spp=2, max_depth=10, rr_min_depth=4, seed=6)
    got, _ = bg.batch_render(batch, cams, prm, stream)
    bg.check_against_single(scs, cams, prm, got, "textured batch")
    batch.close()
    for s in scs:
        s.close()
    cache.close()


def test_missing_texture_is_refused(gpu, tmp_path):
    objs, img, _ = textured_scene(tmp_path)
    scene = ap.SceneRenderer(objs)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 8, 8)
    hdr = ap.DeviceBuffer(8 * 8 * 12)
    with pytest.raises(ap.ApertureError) as e:
        scene.render(cam, ap.RenderParams.defaults(), hdr)    # no texture cache bound
    assert e.value.code == ap.ERR_INVALID_ARG
    empty = ap.TextureCache()
    scene.set_texture_cache(empty)
    with pytest.raises(ap.ApertureError):
        scene.render(cam, ap.RenderParams.defaults(), hdr)    # the cache lacks handle 1
    scene.close()
    empty.close()
