"""GPU parity of the HIP traversal (ap_trace_closest / ap_trace_any) against
the oracle's restatement of TraverseBVH (core/render/intersection.cuh:235-370).

Bar: hit ids and distances bit-exact; the per-ray visit counts of the GPU's
visit order equal the oracle's, in both traversal modes. At the 1M-triangle
benchmark size the full ray sets are checked on samples plus size-independent
properties (hit iff any-hit, any-hit distance >= closest distance).
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes

pytestmark = pytest.mark.gpu


def margin_of(scene):
    return np.float32(scene.info().line_margin)


def check_parity(objs, rays, label):
    print(f"[parity] {label}: {len(objs)} objects, {len(rays)} rays", flush=True)
    import gpu_util

    scene = ap.SceneRenderer(objs)
    bvh = ol.OracleBVH(objs)
    ref = bvh.trace(rays, stats=True)
    cul = bvh.trace(rays, stats=True, cull_margin=margin_of(scene))
    for mode, want in ((ap.TRAVERSE_REFERENCE, ref), (ap.TRAVERSE_CULLED, cul)):
        hits, st = gpu_util.trace(scene, rays, stats=True, mode=mode)
        assert np.array_equal(hits["id"], ref[0]), f"{label} mode {mode}: ids differ"
        assert np.array_equal(hits["t"].view(np.uint32), ref[1].view(np.uint32)), f"{label} mode {mode}: t differ"
        assert np.array_equal(st["nodes"], want[2]), f"{label} mode {mode}: node visit counts differ"
        assert np.array_equal(st["prims"], want[3]), f"{label} mode {mode}: primitive test counts differ"
        plain = gpu_util.trace(scene, rays, mode=mode)
        assert np.array_equal(plain, hits)
        anyh = gpu_util.trace(scene, rays, any_hit=True, mode=mode)
        aref = bvh.trace(rays, any_hit=True)
        assert np.array_equal(anyh["id"], aref[0]) and np.array_equal(anyh["t"], aref[1])
    scene.close()


def test_soup_camera_and_random_rays(gpu):
    objs = ap.synth_triangle_soup(20000, seed=2)
    rays = np.concatenate([ap.synth_camera_rays(96, 96, [0.5, 0.5, -1.5], [0, 0, 1], 40.0),
                           scenes.random_rays(8192, seed=5)])
    check_parity(objs, rays, "soup20k")


def test_secondary_rays(gpu):
    import gpu_util

    objs = ap.synth_triangle_soup(20000, seed=2, edge_min=2e-3, edge_max=2e-2)
    scene = ap.SceneRenderer(objs)
    prim = ap.synth_camera_rays(128, 128, [0.5, 0.5, -1.5], [0, 0, 1], 40.0)
    hits = gpu_util.trace(scene, prim)
    sec = scene.secondary_rays(prim, hits, seed=3)
    assert len(sec) == int((hits["id"] >= 0).sum()) > 0
    check_parity(objs, sec, "secondaries")


def test_shared_edges_and_vertices(gpu):
    grid = scenes.quad_grid(16)
    xs = np.linspace(0.0, 1.0, 17, dtype=np.float32)
    gx, gy = np.meshgrid(xs, xs)
    o = np.stack([gx.ravel(), gy.ravel(), np.ones(gx.size, np.float32)], 1)
    r = np.concatenate([scenes.rays(o, (0, 0, -1)), scenes.rays(o + [0.01, 0.0, 0.0], (0, 0, -1)),
                        scenes.random_rays(4096, seed=9)])
    check_parity(grid, r, "quad_grid")


def test_cubes_spheres_and_invisible(gpu):
    grid = scenes.quad_grid(6)
    grid["visible"][::3] = 0
    objs = scenes.concat([grid, scenes.cubes([(0.5, 0.5, 1.0), (0.2, 0.2, 0.3)], (0.25, 0.01, 0.25), emissive=15.0),
                          scenes.spheres([(0.2, 0.7, 0.4), (0.8, 0.3, 0.2)], [0.1, 0.15])])
    check_parity(objs, scenes.random_rays(8192, seed=13), "mixed")


def test_windows_and_degenerate_rays(gpu):
    objs = ap.synth_triangle_soup(3000, seed=4, edge_min=0.02, edge_max=0.1)
    r = scenes.random_rays(4096, seed=21)
    g = np.random.default_rng(1)
    r["tmin"] = g.uniform(0.0, 0.5, len(r)).astype(np.float32)
    r["tmax"] = g.uniform(0.5, 3.0, len(r)).astype(np.float32)
    r["tmax"][:64] = r["tmin"][:64]          # empty window
    r["d"][64:96] = 0.0                        # zero direction
    check_parity(objs, r, "windows")


def test_empty_and_single(gpu):
    import gpu_util

    empty = ap.SceneRenderer(np.zeros(0, dtype=ap.SCENE_OBJECT_DTYPE))
    h = gpu_util.trace(empty, scenes.rays([(0, 0, 0)], [(0, 0, 1)], tmax=5.0))
    assert h["id"][0] == -1 and h["t"][0] == np.inf
    empty.close()
    one = scenes.triangles([(0, 0, 0)], [(1, 0, 0)], [(0, 1, 0)])
    check_parity(one, scenes.random_rays(2048, seed=2, lo=-0.2, hi=0.8), "single")
    h = gpu_util.trace(ap.SceneRenderer(one), np.zeros(0, dtype=ap.RAY_DTYPE))
    assert len(h) == 0


def test_benchmark_size_sampled(gpu):
    """1M-triangle soup with 4096^2 primaries and their secondaries: sampled
    bit-exact parity plus whole-set properties."""
    import gpu_util

    objs = ap.synth_triangle_soup(1_000_000, seed=2)
    scene = ap.SceneRenderer(objs)
    prim = ap.synth_camera_rays(4096, 4096, [0.5, 0.5, -1.5], [0, 0, 1], 40.0)
    hits = gpu_util.trace(scene, prim)
    sec = scene.secondary_rays(prim, hits, seed=3)
    shits = gpu_util.trace(scene, sec)
    bvh = ol.OracleBVH(objs)
    g = np.random.default_rng(0)
    for rays, h, name in ((prim, hits, "primary"), (sec, shits, "secondary")):
        idx = g.choice(len(rays), 32768, replace=False)
        ids, ts = bvh.trace(rays[idx], cull_margin=margin_of(scene))
        assert np.array_equal(h["id"][idx], ids), name
        assert np.array_equal(h["t"][idx].view(np.uint32), ts.view(np.uint32)), name
        small = idx[:48]  # the reference order is ~1e5 node visits per ray here
        rid, rts = bvh.trace(rays[small])
        assert np.array_equal(h["id"][small], rid) and np.array_equal(h["t"][small], rts), name
        anyh = gpu_util.trace(scene, rays, any_hit=True)
        assert np.array_equal(anyh["id"] >= 0, h["id"] >= 0), name
        assert (anyh["t"][h["id"] >= 0] >= h["t"][h["id"] >= 0]).all(), name
    scene.close()


def test_launch_shapes_and_autotuner_keep_results(gpu):
    """Every launch shape (rays per wave, refill threshold) returns the same
    hits and visit counts; the autotuner picks one of its candidates and its
    hits are an ordinary trace's."""
    import gpu_util

    objs = ap.synth_triangle_soup(30000, seed=4)
    rays = np.concatenate([ap.synth_camera_rays(64, 64, [0.5, 0.5, -1.5], [0, 0, 1], 40.0),
                           scenes.random_rays(5000, seed=8)])
    scene = ap.SceneRenderer(objs)
    print(f"[launch] {len(objs)} triangles, {len(rays)} rays", flush=True)
    ref = gpu_util.trace(scene, rays, stats=True)
    for wr, rf in ((64, 1), (64, 64), (128, 4), (1024, 16), (4096, 64), (65536, 1)):
        scene.set_launch(wr, rf)
        assert scene.get_launch() == (wr, rf)
        got = gpu_util.trace(scene, rays, stats=True)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (wr, rf)
        anyh = gpu_util.trace(scene, rays, any_hit=True)
        assert np.array_equal(anyh["id"] >= 0, ref[0]["id"] >= 0)
    d_rays = ap.DeviceBuffer.from_numpy(np.ascontiguousarray(rays, dtype=ap.RAY_DTYPE))
    d_hits = ap.DeviceBuffer(len(rays) * 8)
    wr, rf, ms = scene.autotune(d_rays, d_hits, n=len(rays))
    assert wr in (128, 192, 256, 320, 384, 448, 512, 640, 768, 1024) and rf in (4, 8, 16) and ms > 0.0
    assert scene.get_launch() == (wr, rf)
    assert np.array_equal(d_hits.to_numpy(ap.HIT_DTYPE, len(rays)), ref[0])
    for bad in ((63, 0), (-64, 0), (0, 65), (131072, 0)):
        with pytest.raises(ap.ApertureError):
            scene.set_launch(*bad)
    scene.close()
