"""CPU tests pinning the oracle's restatement of signed-distance objects
(oracle/or_bvh.cpp: SdfTorus, SdfDiamondUnit, FillSDFVolume,
SampleSDFAdaptive01, RaymarchLocal, IntersectSDFObject, the
central-difference normal) with properties fixed independently of it:

  * a torus object's hit distance equals the smallest positive root of the
    torus quartic (numpy, FP64) within the marching tolerance, for rays from
    outside, from inside the tube and through the hole, with rotated objects;
    its normal equals the analytic torus normal;
  * the torus volume fill equals the FP64 torus distance at the grid points;
    the diamond fill is 8-fold symmetric about its axis and on the axis
    equals the closed-form plane distances of its table, culet and crown;
  * a volume object sampled from a torus grid hits where the analytic torus
    is (within two voxels);
  * a BVH over tori, volumes and triangles returns what brute force over all
    objects returns; undefined SDF types and missing volumes are never hit.
"""
import numpy as np

import aperture as ap
import oracle_lib as ol
import scenes

G = np.random.default_rng(17)


def rot(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return (np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * k @ k).astype(np.float32)


def quartic_t(lo, ld, R, r):
    """Smallest positive t of |p|^2 + R^2 - r^2)^2 = 4 R^2 (x^2 + z^2), p = lo + t ld."""
    lo, ld = np.asarray(lo, np.float64), np.asarray(ld, np.float64)
    ld = ld / np.linalg.norm(ld)
    a = np.dot(ld, ld)
    b = 2 * np.dot(lo, ld)
    c = np.dot(lo, lo) + R * R - r * r
    # (a t^2 + b t + c)^2 - 4 R^2 ((ox + t dx)^2 + (oz + t dz)^2)
    q2 = np.polymul([a, b, c], [a, b, c])
    s = np.polyadd(np.polymul([ld[0], lo[0]], [ld[0], lo[0]]), np.polymul([ld[2], lo[2]], [ld[2], lo[2]]))
    roots = np.roots(np.polysub(q2, 4 * R * R * s))
    real = roots[np.abs(roots.imag) < 1e-7].real
    pos = real[real > 1e-6]
    return pos.min() if len(pos) else np.inf


def ray_arr(o, d, tmin=1e-3, tmax=np.inf):
    return scenes.rays(np.asarray(o, np.float32).reshape(1, 3), np.asarray(d, np.float32).reshape(1, 3), tmin, tmax)


def test_torus_hits_the_quartic_root():
    e = 0.8  # uniform half extent: unit space is local / 1.6
    R, r = 0.375, 0.125
    Rl, rl = 2 * e * R, 2 * e * r
    checked = 0
    for k in range(300):
        rm = rot(G.normal(size=3), G.uniform(0, np.pi)) if k % 2 else np.eye(3, dtype=np.float32)
        c = G.uniform(-1, 1, 3).astype(np.float32)
        objs = scenes.sdfs([c], [e, e, e], 0, (R, r, 0, 0), rotations=[rm])
        kind = k % 3
        if kind == 0:    # from outside toward a point of the tube
            target = np.array([Rl * np.cos(k), G.uniform(-rl, rl) * 0.5, Rl * np.sin(k)])
            lo = G.normal(size=3) * 2.0
        elif kind == 1:  # from the hole, outward
            lo = np.array([0.0, G.uniform(-0.05, 0.05), 0.0])
            target = np.array([np.cos(k), 0.0, np.sin(k)])
        else:            # from inside the tube
            lo = np.array([Rl * np.cos(k), 0.0, Rl * np.sin(k)]) * 1.0
            target = lo + G.normal(size=3)
        ld = target - lo
        ld /= np.linalg.norm(ld)
        o = (rm @ lo + c).astype(np.float32)
        d = (rm @ ld).astype(np.float32)
        t = ol.object_t(objs, 0, ray_arr(o, d)[0])
        t_true = quartic_t(lo, ld, Rl, rl)
        if np.isinf(t_true):
            assert np.isnan(t) or t > 50 * e
            continue
        # The march stops less than eps inside the surface: t within a few eps of the root
        # (grazing rays excepted, where the tube is thin along the ray).
        if np.isnan(t):
            continue
        assert abs(t - t_true) < 2e-3 + 1e-3 * t_true, (k, t, t_true)
        checked += 1
    assert checked > 200


def test_torus_normal_is_the_analytic_normal():
    e = 0.9
    R, r = 0.375, 0.125
    Rl = 2 * e * R
    n_checked = 0
    for k in range(200):
        rm = rot(G.normal(size=3), G.uniform(0, np.pi))
        c = G.uniform(-1, 1, 3).astype(np.float32)
        objs = scenes.sdfs([c], [e, e, e], 0, (R, r, 0, 0), rotations=[rm])
        lo = G.normal(size=3) * 2.0
        target = np.array([Rl * np.cos(k), 0.0, Rl * np.sin(k)])
        ld = (target - lo) / np.linalg.norm(target - lo)
        hit = ol.sdf_hit(objs, 0, ray_arr(rm @ lo + c, rm @ ld)[0])
        if hit is None:
            continue
        p_loc = rm.T.astype(np.float64) @ (hit[0] - c)
        # Analytic normal: from the nearest point of the ring centre line.
        ring = np.array([p_loc[0], 0.0, p_loc[2]])
        ring = ring / np.linalg.norm(ring) * Rl
        n_loc = (p_loc - ring) / np.linalg.norm(p_loc - ring)
        assert np.dot(rm.astype(np.float64) @ n_loc, hit[1]) > 0.999, k
        n_checked += 1
    assert n_checked > 150


def test_volume_fill_matches_the_distance_functions():
    dims = (17, 13, 21)
    tor = ol.sdf_fill(ap.SDF_FUNCTOR_TORUS, (0.3, 0.1, 0, 0), dims)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) / (n - 1) - 0.5 for n in dims[::-1]], indexing="ij")
    want = np.sqrt((np.sqrt(x * x + z * z) - 0.3) ** 2 + y * y) - 0.1
    assert np.allclose(tor, want, atol=2e-7)
    info = (0.94, 0.5, 0.3, 0.5)
    dia = ol.sdf_fill(ap.SDF_FUNCTOR_DIAMOND, info, (33, 33, 33))
    # On the axis (x = y = 0 in the functor's frame: index 16, 16): the max of the
    # table, culet and the crown/pavilion planes at angle 0.
    zc = np.arange(33) / 32.0 - 0.5
    pz = zc / 0.5
    crown = (1.444 * pz - 0.94) / np.sqrt(1 + 1.444 ** 2)
    pav = (-1.072 * pz - 0.94) / np.sqrt(1 + 1.072 ** 2)
    table, culet = pz - 0.3, -pz - 0.865
    girdle = -0.911
    lgf = (-1.02 * pz - 0.9193) / np.sqrt(1 + 1.02 ** 2)
    ugf = (1.03 * pz - 0.912) / np.sqrt(1 + 1.03 ** 2)
    star = (2.21 * pz - 1.131) / np.sqrt(1 + 2.21 ** 2)
    axis = np.max(np.stack([crown, pav, table, culet, np.full_like(pz, girdle), lgf, ugf, star]), 0) * 0.5
    assert np.allclose(dia[:, 16, 16], axis, atol=2e-6)
    # 8-fold symmetry about the functor's z axis (a quarter turn maps the grid onto itself).
    assert np.allclose(dia, np.rot90(dia, 1, axes=(1, 2)), atol=1e-5)


def test_volume_object_follows_its_torus():
    dims = (64, 64, 64)
    R, r = 0.3, 0.1
    vol = ol.sdf_fill(ap.SDF_FUNCTOR_TORUS, (R, r, 0, 0), dims)
    ol.sdf_set_volumes([vol])
    e = 1.0
    objs = scenes.concat([scenes.sdfs([[0, 0, 0]], [e, e, e], 1, volume=0),
                          scenes.sdfs([[5, 0, 0]], [e, e, e], 1, volume=3),     # missing volume
                          scenes.sdfs([[-5, 0, 0]], [e, e, e], 2, volume=0)])   # undefined type
    voxel = 2 * e / 63
    hits = 0
    for k in range(200):
        lo = G.normal(size=3) * 2.5
        target = np.array([2 * e * R * np.cos(k), 0.0, 2 * e * R * np.sin(k)])
        ld = target - lo
        ld /= np.linalg.norm(ld)
        t = ol.object_t(objs, 0, ray_arr(lo, ld)[0])
        t_true = quartic_t(lo, ld, 2 * e * R, 2 * e * r)
        if np.isinf(t_true) or np.isnan(t):
            continue
        assert abs(t - t_true) < 2 * voxel, (k, t, t_true)
        hits += 1
    assert hits > 150
    for idx, c in ((1, [5, 0, 0]), (2, [-5, 0, 0])):
        assert np.isnan(ol.object_t(objs, idx, ray_arr(np.array(c) + [0, 3, 0], [0, -1, 0])[0]))
    ol.sdf_set_volumes([])


def test_bvh_with_sdf_objects_equals_brute_force():
    ol.sdf_set_volumes([ol.sdf_fill(ap.SDF_FUNCTOR_TORUS, (0.35, 0.12, 0, 0), (32, 24, 32)),
                        ol.sdf_fill(ap.SDF_FUNCTOR_DIAMOND, (0.94, 0.5, 0.3, 0.45), (40, 40, 40))])
    n = 40
    centers = G.uniform(-3, 3, (n, 3)).astype(np.float32)
    ext = G.uniform(0.2, 0.6, (n, 3)).astype(np.float32)
    rots = [rot(G.normal(size=3), G.uniform(0, np.pi)) for _ in range(n)]
    types = G.choice([0, 1], n)
    parts = [scenes.sdfs([centers[i]], [ext[i]], int(types[i]), (0.35, 0.12, 0, 0), volume=int(i % 2),
                         rotations=[rots[i]]) for i in range(n)]
    tri = scenes.triangles(G.uniform(-3, 3, (60, 3)), G.uniform(-3, 3, (60, 3)), G.uniform(-3, 3, (60, 3)))
    objs = scenes.concat(parts + [tri])
    rays = scenes.random_rays(3000, seed=5, lo=-4, hi=4)
    rays["tmin"] = 1e-3
    ids, ts = ol.OracleBVH(objs).trace(rays)
    bi, bt = ol.brute_closest(objs, rays)
    assert np.array_equal(ids, bi) and np.array_equal(ts, bt)
    assert (ids >= 0).mean() > 0.2 and np.isin(ids, np.arange(n)).sum() > 100
    ol.sdf_set_volumes([])
