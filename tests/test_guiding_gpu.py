"""GPU parity of directional path guiding (use_guiding = 1) against the
oracle's restatement (oracle/or_guiding.h + the guided loop of
oracle/or_render.cpp).

Frames run one at a time; after each, both sides fold the frame's samples
into the probes (unique voxels in (x, y, z, face) order, per-texel sums in
sample order, FP64 merge, mips), and the next frame samples with them. Bar:
after several frames the images, FP64 sums, probe texels and counts, and
the index volume are bit-identical to the oracle's. Frames start at index
48 so that the guiding probability is already near its 0.5 cap. The build's
compaction, sorts and unique are the project's own scan and stable radix sort;
a 256x160 frame gives them tens of 4096-key tiles and long runs of equal keys.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import test_render_gpu as rg

pytestmark = pytest.mark.gpu


def guided_parity(objs, cam, prm_kw, label, calls=2):
    print(f"[guiding] {label}: {len(objs)} objects, {cam.width}x{cam.height}", flush=True)
    scene = ap.SceneRenderer(objs)
    margin = np.float32(scene.info().line_margin)
    bvh = ol.OracleBVH(objs)
    guide = ol.Guide()
    n = cam.width * cam.height
    d_hdr = ap.DeviceBuffer(n * 12)
    d_hdr.zero()
    d_sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
    o_hdr = np.zeros((cam.height, cam.width, 3), np.float32)
    o_sums = np.zeros((cam.height, cam.width, 3), np.float64)
    frame = 48
    for c in range(calls):
        prm = ap.RenderParams.defaults(frame_index=frame, use_guiding=1, **prm_kw)
        scene.render(cam, prm, d_hdr, sum_xyz=d_sums)
        ap.synchronize()
        o_hdr, o_sums, _ = bvh.render(cam, prm, cull_margin=margin, hdr=o_hdr, guide=guide, sums=o_sums)
        frame += prm.spp
        g = d_hdr.to_numpy(np.float32, n * 3).reshape(cam.height, cam.width, 3)
        gs = d_sums.to_numpy(np.float64, n * 3).reshape(cam.height, cam.width, 3)
        gdata, gcount, gvol, gbounds = scene.guiding_export()
        odata, ocount, ovol, obounds = guide.export()
        same = float(np.all(g == o_hdr, axis=-1).mean())
        print(f"[guiding] {label} call {c}: {len(gcount)} probes (oracle {len(ocount)}), "
              f"identical pixels {same:.4f}", flush=True)
        assert np.array_equal(gbounds, obounds), (gbounds, obounds)
        assert np.array_equal(gvol, ovol)
        assert np.array_equal(gcount, ocount)
        assert np.array_equal(gdata.view(np.uint32), odata.view(np.uint32))
        assert np.array_equal(g, o_hdr)
        assert np.array_equal(gs, o_sums)
    assert len(gcount) > 0
    scene.guiding_reset()
    assert scene.guiding_export()[1].size == 0
    scene.close()


def test_guided_box_scene(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 48, 48)
    guided_parity(objs, cam, dict(spp=4, max_depth=8), "box scene")


def test_guided_heightfield_ggx(gpu):
    objs = ap.synth_heightfield_scene(64, 16, 4, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 36)
    guided_parity(objs, cam, dict(spp=3, max_depth=16, rr_min_depth=4), "heightfield 64^2 + 16 spheres")


def test_guided_four_wavelengths(gpu):
    """Guiding with four hero wavelengths per path: per-wavelength collection,
    the throughput's XYZ as the probe query, sum_j xyz_j L_j / psum as the
    sample radiance, all bit-identical to the oracle."""
    objs = ap.synth_heightfield_scene(48, 8, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 48, 27)
    guided_parity(objs, cam, dict(spp=3, max_depth=16, rr_min_depth=4, wavelengths=4), "heightfield, four wavelengths")


def test_guided_sphere_lights(gpu):
    """Guided sampling mixed with the sphere-cap light pdf in NEE's MIS."""
    import scenes

    objs = scenes.sphere_lit_box(4, keep_cube_light=True)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 40, 40)
    guided_parity(objs, cam, dict(spp=3, max_depth=8), "box with sphere lights")
    guided_parity(objs, cam, dict(spp=2, max_depth=8, wavelengths=4), "box with sphere lights, four wavelengths")


def test_guided_full_heightfield_workload_crop(gpu):
    """The full 1080p heightfield workload (2,097,152 + 81,920 triangles, 4
    lights, depth 16, roulette from depth 4) seen through a 48x48 centre crop
    of its camera, guided over two frames with one and four wavelengths."""
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0 * 48 / 1080, 48, 48)
    guided_parity(objs, cam, dict(spp=2, max_depth=16, rr_min_depth=4), "full heightfield, centre crop")
    guided_parity(objs, cam, dict(spp=2, max_depth=16, rr_min_depth=4, wavelengths=4),
                  "full heightfield, centre crop, four wavelengths")


def test_guided_multi_tile_sorts(gpu):
    """About 200 K guiding samples per frame: the voxel and texel sorts span
    many tiles, and the texel sums depend on the sort keeping equal keys in
    slot order."""
    objs = ap.synth_box_scene(2, 2, 15.0)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 256, 160)
    guided_parity(objs, cam, dict(spp=1, max_depth=8), "box 256x160", calls=3)


def test_guided_groups_of_frames(gpu):
    """guide_frames > 1: the frames of a group are traced together with the
    probes built before the group, then one update folds the group's samples
    in (frame-major slot order); groups of 2 over five frames (the last group
    shorter) and of 4 with four wavelengths, bit-identical to the oracle's
    groups (images, sums, probes, index volume)."""
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 48, 40)
    guided_parity(objs, cam, dict(spp=5, max_depth=8, guide_frames=2), "box scene, groups of 2 frames")
    objs = ap.synth_heightfield_scene(48, 8, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 48, 27)
    guided_parity(objs, cam, dict(spp=4, max_depth=16, rr_min_depth=4, wavelengths=4, guide_frames=4),
                  "heightfield, four wavelengths, groups of 4 frames")


def test_guided_groups_are_refused_where_unsupported(gpu):
    """Groups of frames with the cache or in a batch are refused, as is a
    negative group size."""
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make(rg.BOX_CAM["origin"], rg.BOX_CAM["direction"], rg.BOX_CAM["fov_deg"], 16, 16)
    scene = ap.SceneRenderer(objs)
    hdr = ap.DeviceBuffer(16 * 16 * 12)
    with pytest.raises(ap.ApertureError):
        scene.render(cam, ap.RenderParams.defaults(spp=2, use_guiding=1, guide_frames=-1), hdr)
    cache = ap.RadianceCache(seed=1)
    with pytest.raises(ap.ApertureError):
        scene.render_nrc(cam, ap.RenderParams.defaults(spp=2, use_guiding=1, guide_frames=2), cache,
                         ap.NrcRenderParams.defaults(), hdr)
    cache.close()
    scene.close()
