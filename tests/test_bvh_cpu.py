"""CPU tests of the BVH builder and of the oracle traversal.

The oracle (oracle/or_bvh.cpp) restates BuildBVH (core/bvh/bvh.cuh:142-422)
and TraverseBVH (core/render/intersection.cuh:235-370). It is pinned here by
what the mathematics fixes, independently of its own formulas:
  * builder invariants (partition of the primitives, bounds = union, leaf
    sizes) and equality with the product builder, byte for byte;
  * traversal == brute force over every object (Intersect, :176-205): the same
    closest distance on every ray, the same id except on exact distance ties;
  * closed-form rays against axis-aligned geometry (known t, parallel rays,
    the t > 1e-5 rule, the [tmin, tmax) window, shared edges and vertices);
  * the core's CULLED order returns exactly the reference order's results.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes


def soups():
    yield "soup1", ap.synth_triangle_soup(1, seed=7)
    yield "soup9", ap.synth_triangle_soup(9, seed=8)
    yield "soup3000", ap.synth_triangle_soup(3000, seed=2)
    yield "soup_big_edges", ap.synth_triangle_soup(2000, seed=5, edge_min=0.05, edge_max=0.3)
    yield "grid", scenes.quad_grid(12)
    same = scenes.triangles(np.zeros((20, 3)), np.tile([1, 0, 0], (20, 1)), np.tile([0, 1, 0], (20, 1)),
                            center=(0.1, 0.2, 0.3))
    yield "identical_centroids", same
    mix = scenes.concat([scenes.quad_grid(4), scenes.cubes([(0.5, 0.5, 1.0)], (0.25, 0.01, 0.25), emissive=15.0),
                          scenes.spheres([(0.2, 0.7, 0.4)], [0.1])])
    yield "mixed", mix


@pytest.mark.parametrize("name,objs", list(soups()), ids=[n for n, _ in soups()])
def test_product_builder_matches_oracle_builder(name, objs):
    nodes, idx = ap.build_bvh(objs)
    onodes, oidx = ol.OracleBVH(objs).export()
    assert nodes.tobytes() == onodes.tobytes()
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("name,objs", list(soups()), ids=[n for n, _ in soups()])
def test_builder_invariants(name, objs):
    nodes, idx = ap.build_bvh(objs)
    n = len(objs)
    assert sorted(idx.tolist()) == list(range(n))
    leaves = nodes[nodes["count"] > 0]
    covered = np.zeros(n, np.int32)
    for lf in leaves:
        covered[lf["first"]:lf["first"] + lf["count"]] += 1
    assert (covered == 1).all(), "every slot in exactly one leaf"
    pmin, pmax = objs["proxy_min"], objs["proxy_max"]
    for k, nd in enumerate(nodes):
        if nd["count"] > 0:
            prim = idx[nd["first"]:nd["first"] + nd["count"]]
            assert np.array_equal(nd["bound_min"], pmin[prim].min(0))
            assert np.array_equal(nd["bound_max"], pmax[prim].max(0))
            cen = objs["proxy_center"][prim]
            degenerate = (cen.max(0) - cen.min(0) <= 1e-6).all()
            assert nd["count"] <= 8 or degenerate
            assert nd["left"] == -1 and nd["right"] == -1
        else:
            l, r = nodes[nd["left"]], nodes[nd["right"]]
            assert nd["left"] > k and nd["right"] > nd["left"], "DFS pre-order numbering"
            assert np.array_equal(nd["bound_min"], np.minimum(l["bound_min"], r["bound_min"]))
            assert np.array_equal(nd["bound_max"], np.maximum(l["bound_max"], r["bound_max"]))


def test_identical_centroids_make_one_leaf():
    same = scenes.triangles(np.zeros((20, 3)), np.tile([1, 0, 0], (20, 1)), np.tile([0, 1, 0], (20, 1)),
                            center=(0.1, 0.2, 0.3))
    nodes, _ = ap.build_bvh(same)
    assert len(nodes) == 1 and nodes[0]["count"] == 20


def test_empty_build():
    nodes, _ = ap.build_bvh(np.zeros(0, dtype=ap.SCENE_OBJECT_DTYPE))
    assert len(nodes) == 0


def _agree_with_brute(objs, rays, ids, ts):
    bid, bts = ol.brute_closest(objs, rays)
    assert np.array_equal(ts.view(np.uint32), bts.view(np.uint32)), "closest distance must be exact"
    diff = np.nonzero(ids != bid)[0]
    for i in diff:  # only exact ties may pick a different object
        assert ids[i] >= 0 and bid[i] >= 0
        assert ol.object_t(objs, ids[i], rays[i]) == ts[i] == ol.object_t(objs, bid[i], rays[i])
    return len(diff)


@pytest.mark.parametrize("name,objs", list(soups()), ids=[n for n, _ in soups()])
def test_oracle_traversal_equals_brute_force(name, objs):
    rays = np.concatenate([scenes.random_rays(3000, seed=11), ap.synth_camera_rays(32, 32, [0.5, 0.5, -1.5],
                                                                                     [0, 0, 1], 40.0)])
    ids, ts = ol.OracleBVH(objs).trace(rays)
    _agree_with_brute(objs, rays, ids, ts)
    aid, ats = ol.OracleBVH(objs).trace(rays, any_hit=True)
    hit = ids >= 0
    assert np.array_equal(aid >= 0, hit), "any-hit finds a hit iff closest-hit does"
    assert (ats[hit] >= ts[hit]).all()


def test_closed_form_axis_aligned():
    grid = scenes.quad_grid(4)  # plane z = 0 over [0,1]^2
    o = [(0.3, 0.6, 2.0), (0.3, 0.6, -1.0), (0.3, 0.6, 1e-6), (0.3, 0.6, 1.0), (2.0, 0.5, 0.0), (0.3, 0.6, 1.0)]
    d = [(0, 0, -1), (0, 0, 1), (0, 0, -1), (0, 0, -1), (-1, 0, 0), (0, 0, -1)]
    r = scenes.rays(o, d)
    r["tmax"][3] = 1.0    # exactly at the plane: t >= tmax is rejected
    r["tmin"][5] = 1.5    # window starts beyond the plane
    ids, ts = ol.OracleBVH(grid).trace(r)
    assert ts[0] == 2.0 and ids[0] >= 0
    assert ts[1] == 1.0 and ids[1] >= 0
    assert ids[2] == -1, "t <= 1e-5 is not a hit"
    assert ids[3] == -1 and ts[3] == 1.0
    assert ids[4] == -1, "a ray in the plane is parallel (|a| < 1e-6)"
    assert ids[5] == -1 and ts[5] == np.inf


def test_shared_vertex_first_in_visit_order():
    """A ray through a vertex shared by several triangles hits all of them at
    the same t; the reference keeps the first one it visits."""
    grid = scenes.quad_grid(4)
    r = scenes.rays([(0.5, 0.5, 1.0), (0.25, 0.25, 1.0), (0.5, 0.3, 1.0)], [(0, 0, -1)] * 3)
    bvh = ol.OracleBVH(grid)
    ids, ts = bvh.trace(r)
    assert (ts == 1.0).all() and (ids >= 0).all()
    bid, _ = ol.brute_closest(grid, r)
    # brute force keeps the lowest index; both must be tied candidates
    for i in range(len(r)):
        assert ol.object_t(grid, ids[i], r[i]) == 1.0 == ol.object_t(grid, bid[i], r[i])


def test_empty_scene_returns_inf():
    objs = np.zeros(0, dtype=ap.SCENE_OBJECT_DTYPE)
    ids, ts = ol.OracleBVH(objs).trace(scenes.rays([(0, 0, 0)], [(0, 0, 1)], tmax=5.0))
    assert ids[0] == -1 and ts[0] == np.inf


def test_invisible_triangles_never_hit():
    grid = scenes.quad_grid(2)
    grid["visible"] = 0
    ids, _ = ol.OracleBVH(grid).trace(scenes.rays([(0.3, 0.6, 1.0)], [(0, 0, -1)]))
    assert ids[0] == -1


def test_cube_and_sphere_closed_forms():
    objs = scenes.concat([scenes.cubes([(0, 0, 0)], (1, 2, 3)), scenes.spheres([(10, 0, 0)], [2.0])])
    r = scenes.rays([(0, 0, 10), (0, 0, 0), (10, 0, 10), (10, 0, 0)], [(0, 0, -1)] * 4)
    ids, ts = ol.OracleBVH(objs).trace(r)
    assert ids.tolist() == [0, 0, 1, 1]
    assert ts.tolist() == [7.0, 3.0, 8.0, 2.0]  # outside hits enter, inside hits exit


@pytest.mark.parametrize("name,objs", list(soups()), ids=[n for n, _ in soups()])
def test_culled_order_matches_reference_order(name, objs):
    rays = np.concatenate([scenes.random_rays(4000, seed=3),
                           ap.synth_camera_rays(48, 48, [0.5, 0.5, -1.5], [0, 0, 1], 40.0)])
    bvh = ol.OracleBVH(objs)
    margin = np.float32(max(1.0, np.abs(np.concatenate([objs["proxy_min"], objs["proxy_max"]])).max())) * \
        np.float32(2.0 ** -12)
    a = bvh.trace(rays, stats=True)
    b = bvh.trace(rays, stats=True, cull_margin=margin)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert (b[2] <= a[2]).all() and (b[3] <= a[3]).all()
    aa = bvh.trace(rays, any_hit=True)
    bb = bvh.trace(rays, any_hit=True, cull_margin=margin)
    assert np.array_equal(aa[0], bb[0]) and np.array_equal(aa[1], bb[1])
