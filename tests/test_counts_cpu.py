"""CPU tests of the oracle's visit accounting (oracle/or_count.h), the counts
that price an integrator frame in algorithmic bytes:

  * counting changes no output bit;
  * with depth 1 and no NEE every camera path traces exactly one closest-hit
    ray and no shadow ray, and the shaded vertices are the camera rays with a
    primary hit;
  * NEE always at depth 1 in the closed box: one shadow ray at most per
    shaded vertex; probe queries only when guiding, found <= queried <=
    shaded vertices;
  * the node and primitive totals of a camera's rays equal the per-ray
    statistics of the traversal (or_trace, CULLED order) for the same ray;
  * the committed fixture tests/golden/integrator_counts_heightfield.json
    (made by tests/golden/make_integrator_counts.py) is what the oracle
    counts today on the 2.18 M-triangle heightfield.
"""
import json
import os

import numpy as np

import aperture as ap
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
BOX_CAM = ([0.0, 0.0, 3.4], [0.0, 0.0, -1.0], 40.0)


def box(w=24, h=24):
    objs = ap.synth_box_scene(1, 2, 15.0)
    return objs, ol.OracleBVH(objs), ap.Camera.make(*BOX_CAM, w, h)


def test_counting_changes_nothing():
    objs, bvh, cam = box()
    prm = ap.RenderParams.defaults(spp=2, max_depth=8, nee_always=1, wavelengths=4)
    a = bvh.render(cam, prm, cull_margin=bvh.line_margin())
    with ol.counting() as c:
        b = bvh.render(cam, prm, cull_margin=bvh.line_margin())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    t = c.totals
    assert t["paths"] == 2 * 24 * 24
    assert t["ext_rays"] >= t["paths"] and t["shadow_rays"] > 0 and t["interacts"] > 0


def test_depth_one_counts_camera_rays():
    objs, bvh, cam = box(20, 16)
    prm = ap.RenderParams.defaults(spp=1, max_depth=1, use_nee=0)
    with ol.counting() as c:
        _, _, prim = bvh.render(cam, prm, cull_margin=bvh.line_margin())
    t = c.totals
    assert t["paths"] == 20 * 16
    assert t["ext_rays"] == t["paths"]
    assert t["shadow_rays"] == 0 and t["shadow_nodes"] == 0 and t["shadow_prims"] == 0
    assert t["hits"] == int((prim >= 0).sum())
    assert t["ext_nodes"] >= t["ext_rays"] and t["ext_prims"] > 0


def test_shadow_rays_and_probe_queries_are_bounded():
    objs, bvh, cam = box(32, 32)
    prm = ap.RenderParams.defaults(spp=1, max_depth=1, nee_always=1)
    with ol.counting() as c:
        bvh.render(cam, prm, cull_margin=bvh.line_margin())
    t = c.totals
    assert 0 < t["shadow_rays"] <= t["hits"]
    assert t["probe_queries"] == 0 and t["probe_found"] == 0
    guide = ol.Guide()
    kw = dict(spp=1, max_depth=8, use_guiding=1)
    bvh.render(cam, ap.RenderParams.defaults(frame_index=48, **kw), cull_margin=bvh.line_margin(), guide=guide)
    with ol.counting() as c:
        bvh.render(cam, ap.RenderParams.defaults(frame_index=49, **kw), cull_margin=bvh.line_margin(), guide=guide)
    t = c.totals
    assert 0 < t["probe_found"] <= t["probe_queries"] <= t["hits"]


def test_node_totals_equal_per_ray_statistics():
    """Rays traced by or_trace do not pass through the accounting, so the
    totals of a depth-1 render equal or_trace's per-ray statistics of its
    camera rays. The camera has one pixel of a 1e-4 degree field of view
    aimed at a generic point of the box's back wall (away from every edge
    and split plane): its 64 jittered rays cross the same nodes and leaves
    in the same order as the pixel's central ray."""
    objs, bvh, _ = box()
    origin = np.array([0.0, 0.0, 3.4], np.float32)
    target = np.array([0.23, -0.31, -1.0], np.float32)
    d = (target - origin) / np.linalg.norm(target - origin)
    cam = ap.Camera.make(origin.tolist(), d.tolist(), 1e-4, 1, 1)
    prm = ap.RenderParams.defaults(spp=64, max_depth=1, use_nee=0)
    m = bvh.line_margin()
    with ol.counting() as c:
        bvh.render(cam, prm, cull_margin=m)
    t = c.totals
    assert t["ext_rays"] == 64 and t["hits"] == 64
    ray = np.zeros(1, ap.RAY_DTYPE)
    ray["o"] = origin
    ray["d"] = d.astype(np.float32)
    ray["tmin"] = np.float32(1e-3)  # INTERSECT_EPS at depth 0
    ray["tmax"] = np.inf
    _, _, nodes, prims = bvh.trace(ray, stats=True, cull_margin=m)
    assert t["ext_nodes"] == 64 * int(nodes[0]), (t, nodes)
    assert t["ext_prims"] == 64 * int(prims[0]), (t, prims)


def test_fixture_matches_the_oracle():
    with open(os.path.join(HERE, "golden", "integrator_counts_heightfield.json")) as f:
        fx = json.load(f)
    assert fx["counts"] == list(ol.COUNT_NAMES)
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    bvh = ol.OracleBVH(objs)
    margin = bvh.line_margin()
    view = fx["views"]["view128x72"]
    cam = ap.Camera.make(fx["camera"]["origin"], fx["camera"]["direction"], view["fov_deg"], view["width"],
                         view["height"])
    kw = dict(spp=1, max_depth=16, rr_min_depth=4, wavelengths=4, use_guiding=1)
    guide = ol.Guide()
    bvh.render(cam, ap.RenderParams.defaults(frame_index=48, **kw), cull_margin=margin, guide=guide)
    with ol.counting() as c:
        bvh.render(cam, ap.RenderParams.defaults(frame_index=49, **kw), cull_margin=margin, guide=guide)
    assert c.totals == view["guided_4"]
