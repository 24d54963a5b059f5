"""GPU checks of signed-distance objects (SDF tori and volumes) against the
oracle (oracle/or_bvh.cpp, or_render.cpp).

Bars: volumes filled on the device from the torus and diamond functors equal
the oracle's fill bit for bit; closest-hit and any-hit traversal over scenes
of rotated tori, volume objects, undefined SDF types, missing volumes and
triangles returns the oracle's ids and distances bit for bit, with equal
per-ray visit counts, in both traversal modes; images and FP64 sums of a
closed box holding a torus and a volume object are bit-identical to the
oracle's.
"""
import numpy as np
import pytest

import aperture as ap
import gpu_util
import oracle_lib as ol
import scenes
import test_sdf_cpu as sc

pytestmark = pytest.mark.gpu


def test_volume_fill_matches_the_oracle(gpu):
    cache = ap.SdfCache()
    for functor, info, dims in ((ap.SDF_FUNCTOR_TORUS, (0.3, 0.1, 0, 0), (17, 13, 21)),
                                (ap.SDF_FUNCTOR_DIAMOND, (0.94, 0.5, 0.3, 0.45), (40, 33, 36)),
                                (ap.SDF_FUNCTOR_DIAMOND, (0.0, 0.0, 0.0, 0.0), (8, 8, 8))):
        idx = cache.add_functor(functor, info, dims)
        got = cache.volume(idx)
        want = ol.sdf_fill(functor, info, dims)
        print(f"[sdf] fill functor {functor} dims {dims}: {int((got != want).sum())} samples differ", flush=True)
        assert got.shape == want.shape and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with pytest.raises(ap.ApertureError):
        cache.add_functor(7, (0, 0, 0, 0), (4, 4, 4))
    cache.close()


def sdf_scene(seed=3, n=48):
    g = np.random.default_rng(seed)
    centers = g.uniform(-3, 3, (n, 3)).astype(np.float32)
    ext = g.uniform(0.2, 0.6, (n, 3)).astype(np.float32)
    rots = [sc.rot(g.normal(size=3), g.uniform(0, np.pi)) for _ in range(n)]
    kinds = g.choice([0, 1, 1, 2], n)                     # torus, volume, undefined
    vols = g.choice([0, 1, 5], n)                         # volume 5 is missing
    parts = [scenes.sdfs([centers[i]], [ext[i]], int(kinds[i]), (0.35, 0.12, 0, 0), volume=int(vols[i]),
                         rotations=[rots[i]]) for i in range(n)]
    tri = scenes.triangles(g.uniform(-3, 3, (80, 3)), g.uniform(-3, 3, (80, 3)), g.uniform(-3, 3, (80, 3)))
    return scenes.concat(parts + [tri])


def volumes():
    return [((ap.SDF_FUNCTOR_TORUS, (0.35, 0.12, 0, 0), (32, 24, 32))),
            ((ap.SDF_FUNCTOR_DIAMOND, (0.94, 0.5, 0.3, 0.45), (40, 40, 40)))]


def test_traversal_matches_the_oracle(gpu):
    objs = sdf_scene()
    cache = ap.SdfCache()
    host = []
    for f, info, dims in volumes():
        cache.add_functor(f, info, dims)
        host.append(ol.sdf_fill(f, info, dims))
    ol.sdf_set_volumes(host)
    scene = ap.SceneRenderer(objs)
    scene.set_sdf_cache(cache)
    bvh = ol.OracleBVH(objs)
    rays = scenes.random_rays(20000, seed=9, lo=-4, hi=4)
    rays["tmin"] = np.where(np.arange(len(rays)) % 3 == 0, 0.0, 1e-3 * np.exp(np.arange(len(rays)) % 5 * 0.5))
    margin = np.float32(scene.info().line_margin)
    ref = bvh.trace(rays, stats=True)
    cul = bvh.trace(rays, stats=True, cull_margin=margin)
    for mode, want in ((ap.TRAVERSE_REFERENCE, ref), (ap.TRAVERSE_CULLED, cul)):
        hits, st = gpu_util.trace(scene, rays, stats=True, mode=mode)
        sdf_hits = int(np.isin(hits["id"], np.arange(48)).sum())
        print(f"[sdf] traversal mode {mode}: {int((hits['id'] >= 0).sum())} hits, {sdf_hits} on SDF objects; "
              f"id mismatches {int((hits['id'] != ref[0]).sum())}", flush=True)
        assert np.array_equal(hits["id"], ref[0])
        assert np.array_equal(hits["t"].view(np.uint32), ref[1].view(np.uint32))
        assert np.array_equal(st["nodes"], want[2]) and np.array_equal(st["prims"], want[3])
        anyh = gpu_util.trace(scene, rays, any_hit=True, mode=mode)
        aref = bvh.trace(rays, any_hit=True)
        assert np.array_equal(anyh["id"], aref[0]) and np.array_equal(anyh["t"].view(np.uint32),
                                                                      aref[1].view(np.uint32))
        assert sdf_hits > 300
    scene.close()
    cache.close()
    ol.sdf_set_volumes([])


def test_render_with_sdf_objects_matches_the_oracle(gpu):
    box = ap.synth_box_scene(1, 2, 15.0)
    extra = scenes.concat([scenes.sdfs([[-0.35, -0.55, -0.2]], [0.35, 0.35, 0.35], 0, (0.375, 0.125, 0, 0),
                                       rotations=[sc.rot([1, 0, 0], 0.6)]),
                           scenes.sdfs([[0.4, -0.6, 0.1]], [0.3, 0.3, 0.3], 1, volume=1)])
    extra["albedo"] = [[0.8, 0.3, 0.2], [0.2, 0.6, 0.9]]
    extra["shading_model"] = [ap.MAT_DIFFUSE, ap.MAT_GGX]
    extra["roughness"] = 0.4
    objs = scenes.concat([box, extra])
    cache = ap.SdfCache()
    host = []
    for f, info, dims in volumes():
        cache.add_functor(f, info, dims)
        host.append(ol.sdf_fill(f, info, dims))
    ol.sdf_set_volumes(host)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 48, 40)
    scene = ap.SceneRenderer(objs)
    scene.set_sdf_cache(cache)
    margin = np.float32(scene.info().line_margin)
    for wavelengths in (1, 4):
        prm = ap.RenderParams.defaults(spp=3, max_depth=10, seed=21, wavelengths=wavelengths)
        n = cam.width * cam.height
        hdr = ap.DeviceBuffer(n * 12)
        hdr.zero()
        sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
        prim = ap.DeviceBuffer.from_numpy(np.full(n, -1, np.int32))
        scene.render(cam, prm, hdr, sum_xyz=sums, primary_id=prim)
        ap.synchronize()
        g_hdr = hdr.to_numpy(np.float32, n * 3).reshape(cam.height, cam.width, 3)
        g_sums = sums.to_numpy(np.float64, n * 3).reshape(cam.height, cam.width, 3)
        g_prim = prim.to_numpy(np.int32, n).reshape(cam.height, cam.width)
        o_hdr, o_sums, o_prim = ol.OracleBVH(objs).render(cam, prm, cull_margin=margin)
        sdf_px = int(np.isin(g_prim, [len(box), len(box) + 1]).sum())
        ident = float(np.all(g_sums == o_sums, axis=-1).mean())
        print(f"[sdf] render, {wavelengths} wavelengths: {sdf_px} pixels see an SDF object first, "
              f"pixels bit-identical {ident:.4f}", flush=True)
        assert sdf_px > 50
        assert np.array_equal(g_prim, o_prim)
        assert np.array_equal(g_hdr, o_hdr) and np.array_equal(g_sums, o_sums)
    scene.close()
    cache.close()
    ol.sdf_set_volumes([])
