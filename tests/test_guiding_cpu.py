"""CPU tests pinning the oracle's path-guiding restatement (oracle/or_guiding.h,
the guided path loop in oracle/or_render.cpp) with properties the
mathematics fixes, independent of the restated code:

  * the Clarberg map is an equal-area bijection: UVtoSphere(SphereToUV(d)) = d,
    uniform uv gives z and phi uniform (core/guiding/octahedron.cuh:13-145);
  * a probe's OctahedronPDF integrates to one over the sphere, its hemisphere
    pdf to one over the hemisphere, the hierarchical sampler returns the pdf
    of the direction it drew, and its histogram follows that pdf
    (octahedron.cuh:271-401);
  * GenerateMipmaps keeps the mip top equal to the mean of the leaves;
  * icosahedral face binning picks the nearest of 20 unit face normals;
  * one build step over hand-made samples gives the probe set, index volume
    and FP64 count-weighted means a numpy restatement of the merge expects
    (scene.cu:583-744, guiding.cuh:349-506);
  * guiding only changes the sampling pdf, so the gray floor under a square
    light still converges to its closed form (one and four wavelengths per
    path);
  * frames in a guiding group (guide_frames F) all sample the probes built
    before the group: their FP64 sums equal those of the same frames each
    rendered alone after the same earlier frames, and a group of F frames
    over empty probes equals the unguided render of those frames.
"""
import math

import numpy as np

import aperture as ap
import oracle_lib as ol
import test_render_cpu as rc

RNG = np.random.default_rng(21)


def unit(n):
    v = RNG.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_clarberg_round_trip_and_equal_area():
    d = unit(100000)
    back = ol.pg_uv_to_sphere(ol.pg_sphere_to_uv(d))
    assert np.abs(back - d).max() < 2e-5
    uv = RNG.random((400000, 2), dtype=np.float32)
    s = ol.pg_uv_to_sphere(uv)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=2e-6)
    hz, _ = np.histogram(s[:, 2], bins=20, range=(-1, 1))
    hp, _ = np.histogram(np.arctan2(s[:, 1], s[:, 0]), bins=20, range=(-math.pi, math.pi))
    for h in (hz, hp):
        e = len(s) / 20
        assert (((h - e) ** 2) / e).sum() < 60.0  # chi2 with 19 dof


def random_probe(seed, zero_fraction=0.3):
    g = np.random.default_rng(seed)
    leaf = g.gamma(0.5, 1.0, size=(64, 3)).astype(np.float32)
    leaf[g.random(64) < zero_fraction] = 0.0
    return ol.pg_probe(leaf)


def test_mipmap_top_is_leaf_mean():
    p = random_probe(1)
    leaf = p["data"][0, :64, :3].astype(np.float64)
    assert leaf.min() >= np.float32(1e-5)  # the PATH_GUIDING_PROBE_MIN_BRIGHTNESS floor
    assert np.allclose(p["data"][0, 84, :3], leaf.mean(0), rtol=1e-6)
    assert np.allclose(p["data"][0, 64:80, :3].reshape(4, 4, 3).mean((0, 1)), leaf.mean(0), rtol=1e-6)
    assert (p["data"][0, 85] == 0).all()


def test_pdf_integrates_to_one_and_matches_sampler():
    p = random_probe(2)
    q = np.array([0.2, 0.5, 0.3], np.float32)
    d = unit(400000)
    assert abs(ol.pg_pdf(p, d, q).mean() * 4 * math.pi - 1.0) < 0.01
    n = np.array([0.3, 0.8, -0.52], np.float32)
    n /= np.linalg.norm(n)
    hemi = ol.pg_pdf(p, d, q, normal=n)
    assert (hemi[d @ n <= 0] == 0).all()
    assert abs(hemi.mean() * 4 * math.pi - 1.0) < 0.01
    dirs, pdfs = ol.pg_sample(p, 200000, q, seed=5)
    ref = ol.pg_pdf(p, dirs, q)
    assert np.median(np.abs(pdfs / ref - 1.0)) < 1e-5
    assert np.mean(np.abs(pdfs / ref - 1.0) < 1e-4) > 0.99
    # Histogram of sampled leaf texels against the probe's texel probabilities.
    uv = ol.pg_sphere_to_uv(dirs)
    cell = np.clip((uv * 8).astype(int), 0, 7)
    counts = np.bincount(cell[:, 1] * 8 + cell[:, 0], minlength=64)
    w = p["data"][0, :64, :3].astype(np.float64) @ q.astype(np.float64)
    expect = w / w.sum() * len(dirs)
    m = expect > 50
    assert (((counts - expect) ** 2 / expect)[m]).sum() < 3.0 * m.sum()
    hd, hp = ol.pg_sample(p, 50000, q, seed=6, normal=n)
    assert (hd @ n >= 0).all()
    assert np.median(np.abs(hp / ol.pg_pdf(p, hd, q, normal=n) - 1.0)) < 1e-5


def test_face_binning_and_voxels():
    d = unit(50000)
    pos = (RNG.random((50000, 3)) * 4 - 2).astype(np.float32)
    v = ol.pg_voxel(pos, d)
    assert np.array_equal(v[:, :3], np.floor(pos.astype(np.float32) * np.float32(20.0)).astype(np.int32))
    phi = (1 + 5 ** 0.5) / 2
    # The 20 face normals of the icosahedron with vertices (0,+-1,+-phi) and
    # cyclic shifts: the normalised (+-1,+-1,+-1) and (0,+-phi,+-1/phi) with
    # cyclic shifts. The chosen face must have the largest dot.
    fam = [(a, b, c) for a in (-1, 1) for b in (-1, 1) for c in (-1, 1)]
    for s1 in (-1, 1):
        for s2 in (-1, 1):
            fam += [(0, s1 * phi, s2 / phi), (s2 / phi, 0, s1 * phi), (s1 * phi, s2 / phi, 0)]
    normals = np.array(fam, np.float64)
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    assert len(normals) == 20
    best = (d.astype(np.float64) @ normals.T).max(1)
    faces = np.zeros(20)
    # Each face index gathers the directions around one normal of the family.
    for f in range(20):
        sel = v[:, 3] == f
        assert sel.any()
        c = d[sel].mean(0)
        c /= np.linalg.norm(c)
        faces[f] = (normals @ c).max()
        assert np.all(np.abs((d[sel].astype(np.float64) @ normals.T).max(1) - best[sel]) < 1e-6)
    assert faces.min() > 0.95


def make_samples(specs):
    out = np.zeros(len(specs), ol.GUIDE_SAMPLE_DTYPE)
    for i, (valid, vox, direction, rad) in enumerate(specs):
        out[i]["valid"] = valid
        out[i]["voxel"] = vox
        out[i]["direction"] = direction
        out[i]["radiance"] = rad
    return out


def test_build_step_against_numpy_merge():
    g = ol.Guide()
    d1 = np.array([0, 0, 1], np.float32)
    d2 = np.array([1, 0, 0], np.float32)
    frame1 = make_samples([
        (1, (3, -2, 5, 7), d1, (1.0, 2.0, 3.0)),
        (0, (9, 9, 9, 1), d1, (5.0, 5.0, 5.0)),      # invalid slot: ignored
        (1, (3, -2, 5, 7), d1, (0.5, 0.25, 0.125)),
        (1, (-4, 1, 0, 19), d2, (2.0, 0.0, 1.0)),
    ])
    g.update(frame1)
    data, counts, vol, bounds = g.export()
    # Keys sorted (x, y, z, face): (-4,1,0,19) -> probe 0, (3,-2,5,7) -> probe 1.
    assert list(bounds[:6]) == [-4, -2, 0, 3, 1, 5] and list(bounds[6:]) == [8, 4, 6]
    assert list(counts) == [1, 2]
    sx, sy, sz = bounds[6:]

    def lookup(v):
        lx, ly, lz = v[0] - bounds[0], v[1] - bounds[1], v[2] - bounds[2]
        return vol[((lz * sy + ly) * sx + lx) * 20 + v[3]]

    assert lookup((-4, 1, 0, 19)) == 0 and lookup((3, -2, 5, 7)) == 1
    assert (vol >= 0).sum() == 2
    J = np.float32(4.0 * np.float32(3.141592654))
    leaf1 = ol.pg_sphere_to_uv(d1[None])[0]
    t1 = int(min(7, leaf1[1] * 8)) * 8 + int(min(7, leaf1[0] * 8))
    s = (np.float32(1.0) * J + np.float32(0.5) * J, np.float32(2.0) * J + np.float32(0.25) * J,
         np.float32(3.0) * J + np.float32(0.125) * J)
    mean1 = [np.float32(np.float64(v) / 2.0) for v in s]
    assert np.array_equal(data[1, t1, :3], np.array(mean1, np.float32))
    # Second frame: a new probe far away grows the volume; probe 1 gets a weighted merge.
    frame2 = make_samples([(1, (3, -2, 5, 7), d1, (4.0, 4.0, 4.0)), (1, (30, 2, -7, 0), d2, (1.0, 1.0, 1.0))])
    g.update(frame2)
    data2, counts2, vol2, bounds2 = g.export()
    assert list(bounds2[:6]) == [-4, -2, -7, 30, 2, 5]
    assert list(counts2) == [1, 3, 1]
    new = np.float32(4.0) * J
    # alpha = 2/3 and 1/den = 1/3 are formed in FP64 exactly as the merge forms them.
    inv = 1.0 / 3.0
    want = [np.float32(np.float64(mean1[k]) * (2.0 * inv) + np.float64(new) * inv) for k in range(3)]
    assert np.array_equal(data2[1, t1, :3], np.array(want, np.float32))
    # Untouched probe 0 keeps its texels; every probe keeps the mip invariant.
    assert np.array_equal(data2[0], data[0])
    assert np.allclose(data2[:, 84, :3], data2[:, :64, :3].mean(1), rtol=1e-6)


def test_guided_floor_matches_closed_form():
    """Guiding (alpha = 0.5 from frame 64 on) must leave the estimate unbiased,
    with one wavelength per path and with four."""
    objs = rc.floor_and_light(0.5, 10.0, a=0.25, h=1.0)
    expected = 0.5 * 10.0 * 4 * rc.corner_form_factor(0.25, 0.25)
    cam = ap.Camera.make([0.0, 0.5, 0.0], [0.0, -1.0, 0.001], 1.0, 4, 4)
    bvh = ol.OracleBVH(objs)
    for nl, spp in ((1, 2048), (4, 1024)):
        g = ol.Guide()
        prm = ap.RenderParams.defaults(spp=spp, frame_index=64, max_depth=2, sky=0, use_nee=0, use_guiding=1,
                                       wavelengths=nl)
        hdr, sums, prim = bvh.render(cam, prm, cull_margin=2 ** -12, guide=g)
        assert g.probe_count() > 0
        y = float((sums[..., 1] / spp).mean())
        assert abs(y / expected - 1.0) < 0.05, (nl, y, expected)


def test_guiding_group_frames_share_the_probes_before_the_group():
    """guide_frames = 3: frames 2, 3, 4 of a group sample the probes built
    from frames 0 and 1, so each equals that frame rendered alone (one frame,
    its own call) after the same two frames; the FP64 sums of the group are
    the three frames' sums added in frame order. A probe update inside the
    group (a frame seeing the probes of the frame before it) would fail it."""
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 20, 16)
    bvh = ol.OracleBVH(objs)
    kw = dict(max_depth=8, use_guiding=1, seed=9)

    def warm():
        g = ol.Guide()
        bvh.render(cam, ap.RenderParams.defaults(spp=2, frame_index=0, **kw), cull_margin=2 ** -12, guide=g)
        assert g.probe_count() > 0
        return g

    g = warm()
    _, group_sums, _ = bvh.render(cam, ap.RenderParams.defaults(spp=3, frame_index=2, guide_frames=3, **kw),
                                  cull_margin=2 ** -12, guide=g)
    alone = np.zeros_like(group_sums)
    for f in (2, 3, 4):
        _, s_f, _ = bvh.render(cam, ap.RenderParams.defaults(spp=1, frame_index=f, **kw), cull_margin=2 ** -12,
                               guide=warm())
        alone += s_f
    assert np.array_equal(group_sums, alone)
    # With guide_frames 1 the same frames differ: frames 3 and 4 see frame 2's samples.
    g1 = warm()
    _, seq_sums, _ = bvh.render(cam, ap.RenderParams.defaults(spp=3, frame_index=2, **kw), cull_margin=2 ** -12,
                                guide=g1)
    assert not np.array_equal(seq_sums, alone)
    # Over empty probes a group is the unguided render of its frames.
    g0 = ol.Guide()
    _, empty_sums, _ = bvh.render(cam, ap.RenderParams.defaults(spp=3, frame_index=0, guide_frames=4, **kw),
                                  cull_margin=2 ** -12, guide=g0)
    plain = dict(kw, use_guiding=0)
    _, plain_sums, _ = bvh.render(cam, ap.RenderParams.defaults(spp=3, frame_index=0, **plain), cull_margin=2 ** -12)
    assert np.array_equal(empty_sums, plain_sums)
    assert g0.probe_count() > 0
