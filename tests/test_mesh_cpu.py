"""CPU tests of mesh and texture ingestion.

ap_mesh_load_obj (host code of the library, csrc/ap_mesh.cpp) against an
independent reading of the same OBJ files here: the reference's
LoadTexturedMesh (core/helper/scene_helper.h:453-564) over tinyobjloader
2.0's reading (triangles as written, quads split along the shorter
diagonal, larger faces ear-clipped, negative indices, v / v/vt / v//vn /
v/vt/vn vertices, faces of fewer than three vertices skipped), field by
field and bit for bit; ear-clipped convex and concave polygons in each
projection plane are tiled by their triangles (areas add up).

The oracle's texture sampler (oracle/or_render.cpp TexSample) against the
CUDA-documented bilinear filter evaluated in FP64 with the same 8-bit
weights, exact texel values at texel centres, wrap addressing, and within
one weight quantum of the unquantised filter; and, in the integrator, a
constant texture renders exactly as the same albedo without a texture.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes

OBJ = """# a small mesh
mtllib none.mtl
o first
v 0.0 0.0 0.0
v 1.0 0.0 0.0
v 1.0 1.0 0.0
v 0.0 1.0 0.25
v 2.0 0.0 0.5
v 3.0 0.1 0.5
v 3.0 1.0 0.5
v 2.0 3.0 0.5
vt 0.0 0.0
vt 1.0 0.0
vt 1.0 1.0
vt 0.0 1.0
vt 0.25 0.75 0.0
vn 0.0 0.0 1.0
vn 0.0 0.6 0.8
vn 0.6 0.0 0.8
usemtl a
s 1
f 1/1/1 2/2/2 3/3/3
f 1/1/1 2/2/2 3/3/3 4/4/1
g second
f -4/5 -3/4 -2/3 -1/2
f -3 -2 -1 -4
f 1//2 3//3 4//1
f 2 3 4
f 2/1 3/2/3 4/3/2
f 1 2
l 1 2
"""


def ear_clip_here(face, V):
    """tinyobjloader 2.0's built-in ear clipping, in float32 scalars."""
    f32 = np.float32
    P = lambda i, a: f32(V[i[0]][a])
    n0 = len(face)
    axes = [1, 2]
    eps = f32(np.finfo(np.float32).eps)
    for k in range(n0):
        a, b, c = face[k % n0], face[(k + 1) % n0], face[(k + 2) % n0]
        e0 = [P(b, i) - P(a, i) for i in range(3)]
        e1 = [P(c, i) - P(b, i) for i in range(3)]
        cx = abs(e0[1] * e1[2] - e0[2] * e1[1])
        cy = abs(e0[2] * e1[0] - e0[0] * e1[2])
        cz = abs(e0[0] * e1[1] - e0[1] * e1[0])
        if cx > eps or cy > eps or cz > eps:
            if not (cx > cy and cx > cz):
                axes[0] = 0
                if cz > cx and cz > cy:
                    axes[1] = 1
            break
    rem, out = list(face), []
    guess, iters, prev = 0, len(rem), len(rem)
    while len(rem) > 3 and iters > 0:
        n = len(rem)
        if guess >= n:
            guess -= n
        if prev != n:
            prev, iters = n, n
        else:
            iters -= 1
        ind = [rem[(guess + k) % n] for k in range(3)]
        vx = [P(i, axes[0]) for i in ind]
        vy = [P(i, axes[1]) for i in ind]
        cross = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1])
        area = (vx[0] * vy[1] - vy[0] * vx[1]) * f32(0.5)
        if cross * area < 0:
            guess += 1
            continue
        inside = False
        for other in range(3, n):
            t = rem[(guess + other) % n]
            tx, ty = P(t, axes[0]), P(t, axes[1])
            c = False
            j = 2
            for i in range(3):
                if (vy[i] > ty) != (vy[j] > ty) and tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]:
                    c = not c
                j = i
            if c:
                inside = True
                break
        if inside:
            guess += 1
            continue
        out.append(ind)
        del rem[(guess + 1) % n]
    if len(rem) == 3:
        out.append(rem)
    return out


def read_obj_here(text, position, scale, textures):
    """The same reading, written independently of the library."""
    V, T, N, tris = [], [], [], []

    def fix(i, n):
        i = int(i)
        assert i != 0
        return i - 1 if i > 0 else n + i

    for raw in text.splitlines():
        tok = raw.strip().split()
        if not tok:
            continue
        if tok[0] == "v":
            V.append([np.float32(float(x)) for x in tok[1:4]])
        elif tok[0] == "vt":
            T.append([np.float32(float(x)) for x in (tok[1:3] + ["0", "0"])[:2]])
        elif tok[0] == "vn":
            N.append([np.float32(float(x)) for x in tok[1:4]])
        elif tok[0] == "f":
            face = []
            for t in tok[1:]:
                parts = t.split("/")
                v = fix(parts[0], len(V))
                vt = fix(parts[1], len(T)) if len(parts) > 1 and parts[1] else -1
                vn = fix(parts[2], len(N)) if len(parts) > 2 and parts[2] else -1
                face.append((v, vt, vn))
            if len(face) < 3:
                continue
            if len(face) == 3:
                tris.append(face)
                continue
            if len(face) > 4:
                tris.extend(ear_clip_here(face, V))
                continue
            P = [np.array(V[f[0]], np.float32) for f in face]
            e02, e13 = P[2] - P[0], P[3] - P[1]
            s02 = np.float32(np.float32(e02[0] * e02[0] + e02[1] * e02[1]) + e02[2] * e02[2])
            s13 = np.float32(np.float32(e13[0] * e13[0] + e13[1] * e13[1]) + e13[2] * e13[2])
            order = (0, 1, 2, 0, 2, 3) if s02 < s13 else (0, 1, 3, 1, 2, 3)
            tris.append([face[order[0]], face[order[1]], face[order[2]]])
            tris.append([face[order[3]], face[order[4]], face[order[5]]])
    out = []
    sc = np.float32(scale)
    for tri in tris:
        pos = [np.array(V[v], np.float32) * sc for v, _, _ in tri]
        nrm = [np.array(N[n], np.float32) if n >= 0 else np.array([0, 1, 0], np.float32) for _, _, n in tri]
        uv = [np.array([T[t][0], np.float32(1.0) - T[t][1]], np.float32) if t >= 0 else np.zeros(2, np.float32)
              for _, t, _ in tri]
        if tri[0][2] < 0:
            fn = np.cross((pos[1] - pos[0]).astype(np.float64), (pos[2] - pos[0]).astype(np.float64))
            nrm = [fn / np.linalg.norm(fn)] * 3
        out.append((pos, nrm, uv))
    return out


def write(tmp_path, text, name="mesh.obj", crlf=False):
    p = tmp_path / name
    p.write_bytes(text.replace("\n", "\r\n" if crlf else "\n").encode())
    return str(p)


@pytest.mark.parametrize("crlf", [False, True])
def test_obj_triangles_match_an_independent_reading(tmp_path, crlf):
    path = write(tmp_path, OBJ, crlf=crlf)
    pos, scale, tex = (0.5, -0.25, 2.0), 1.5, (3, 4, 5, 6)
    objs = ap.load_obj(path, pos, scale, tex)
    want = read_obj_here(OBJ, pos, scale, tex)
    assert len(objs) == len(want) == 10
    for o, (P, Nrm, UV) in zip(objs, want):
        assert o["type"] == ap.OBJ_TRIANGLE and o["shading_model"] == ap.MAT_GGX
        assert np.array_equal(o["center"], np.float32(pos)) and np.all(o["extent"] == np.float32(scale))
        assert np.array_equal(o["albedo"], [1, 1, 1]) and o["roughness"] == np.float32(0.4)
        assert o["metallic"] == 0 and o["specular"] == 0 and o["ior"] == np.float32(1.5)
        assert list(o["textures"]) == list(tex)
        assert np.array_equal(o["object_to_world"], np.eye(3)) and np.array_equal(o["world_to_object"], np.eye(3))
        for k in range(3):
            assert np.array_equal(o[f"v{k}_position"], P[k]), k
            assert np.array_equal(o[f"v{k}_uv"], UV[k]), k
            assert np.allclose(o[f"v{k}_normal"], Nrm[k], atol=1e-7), k
        w = [P[k] + np.float32(pos) for k in range(3)]
        assert np.array_equal(o["proxy_min"], np.minimum(np.minimum(w[0], w[1]), w[2]))
        assert np.array_equal(o["proxy_max"], np.maximum(np.maximum(w[0], w[1]), w[2]))
        assert np.array_equal(o["proxy_center"], ((w[0] + w[1]) + w[2]) / np.float32(3.0))
        assert o["light_index"] == -1 and o["object_index"] == -1 and o["visible"] == 1
    ref = ap.load_obj(path, pos, scale, tex, reference_proxy=True)   # UpdateProxy's own centre
    assert np.all(ref["proxy_center"] == np.float32(pos))
    same = [f for f in ap.SCENE_OBJECT_DTYPE.names if f != "proxy_center"]
    assert all(np.array_equal(ref[f], objs[f]) for f in same)
    # Quad 1 2 3 4: |v2 - v0|^2 = 2 < |v3 - v1|^2 = 2.0625, so (v0 v1 v2) (v0 v2 v3).
    assert np.array_equal(objs[1]["v2_position"], objs[2]["v1_position"])
    # Quad -3 -2 -1 -4 (vertices 6 7 8 5): 9.41 > 2, so (v0 v1 v3) (v1 v2 v3).
    assert np.array_equal(objs[5]["v1_position"], objs[6]["v0_position"])
    assert np.array_equal(objs[5]["v2_position"], objs[6]["v2_position"])


POLYGONS = """# a convex pentagon in the xy plane, a concave L hexagon in the xz plane,
# a heptagon with two reflex corners in the yz plane
v 0 0 0
v 2 0 0
v 2.5 1.5 0
v 1 2.5 0
v -0.5 1.5 0
v 0 1 0
v 3 1 0
v 3 1 1
v 1 1 1
v 1 1 3
v 0 1 3
v 5 0 0
v 5 2 0
v 5 2 1
v 5 1 1
v 5 1.5 2.5
v 5 0 2
v 5 0.5 1
f 1 2 3 4 5
f 6 7 8 9 10 11
f 12 13 14 15 16 17 18
"""


def shoelace(pts):
    x, y = pts[:, 0], pts[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_larger_faces_are_ear_clipped(tmp_path):
    objs = ap.load_obj(write(tmp_path, POLYGONS, "poly.obj"))
    want = read_obj_here(POLYGONS, (0, 0, 0), 1.0, (0, 0, 0, 0))
    assert len(objs) == len(want) == 3 + 4 + 5
    for o, (P, _, _) in zip(objs, want):
        for k in range(3):
            assert np.array_equal(o[f"v{k}_position"], P[k])
    # Each polygon's triangles tile it: their areas add up to the polygon's.
    V = np.array([[float(x) for x in l.split()[1:]] for l in POLYGONS.splitlines() if l.startswith("v ")])
    for (lo, hi), idx, drop in (((0, 3), [0, 1, 2, 3, 4], 2), ((3, 7), [5, 6, 7, 8, 9, 10], 1),
                                ((7, 12), [11, 12, 13, 14, 15, 16, 17], 0)):
        keep = [a for a in range(3) if a != drop]
        tri_area = 0.0
        for o in objs[lo:hi]:
            p = np.array([o[f"v{k}_position"] for k in range(3)], np.float64)[:, keep]
            tri_area += shoelace(p)
        assert abs(tri_area - shoelace(V[idx][:, keep])) < 1e-6


def test_obj_errors(tmp_path):
    with pytest.raises(ap.ApertureError) as e:
        ap.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 4\n", "range.obj"))
    assert e.value.code == ap.ERR_INVALID_ARG
    with pytest.raises(ap.ApertureError):
        ap.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 0 1 2\n", "zero.obj"))
    with pytest.raises(ap.ApertureError):
        ap.load_obj(str(tmp_path / "missing.obj"))
    assert len(ap.load_obj(write(tmp_path, "# nothing\n", "empty.obj"))) == 0


def checker(w, h, seed=0):
    g = np.random.default_rng(seed)
    return g.integers(0, 256, (h, w, 4), dtype=np.uint8)


def bilinear64(img, u, v, quantise=True):
    h, w = img.shape[:2]
    f = lambda t: t - np.floor(t)
    x = f(u) * w - 0.5
    y = f(v) * h - 0.5
    fx, fy = np.floor(x), np.floor(y)
    a, b = x - fx, y - fy
    if quantise:
        a, b = np.floor(a * 256 + 0.5) / 256, np.floor(b * 256 + 0.5) / 256
    i0, j0 = int(fx) % w, int(fy) % h
    i1, j1 = (i0 + 1) % w, (j0 + 1) % h
    T = img.astype(np.float64) / 255.0
    return (1 - a) * (1 - b) * T[j0, i0] + a * (1 - b) * T[j0, i1] + (1 - a) * b * T[j1, i0] + a * b * T[j1, i1]


def test_oracle_texture_sampler():
    imgs = [checker(1, 1, 1), checker(3, 5, 2), checker(64, 32, 3)]
    ol.texture_set(imgs)
    g = np.random.default_rng(4)
    for k, img in enumerate(imgs):
        h, w = img.shape[:2]
        # Texel centres return the texel.
        jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        uv = np.stack([(ii.ravel() + 0.5) / w, (jj.ravel() + 0.5) / h], 1).astype(np.float32)
        got = ol.texture_sample(k, uv)
        assert np.allclose(got, img.reshape(-1, 4) / 255.0, atol=1e-7)
        # FP64 filter with the same 8-bit weights; the unquantised filter within one weight quantum.
        uv = g.uniform(-3, 3, (500, 2)).astype(np.float32)
        got = ol.texture_sample(k, uv)
        for i in range(len(uv)):
            u, v = float(uv[i, 0]), float(uv[i, 1])
            assert np.allclose(got[i], bilinear64(img, u, v), atol=3e-6), (k, u, v)
            assert np.all(np.abs(got[i] - bilinear64(img, u, v, False)) <= 2.0 / 256 + 1e-6)
        # Wrap: whole periods do not move the sample (exact shifts of a float).
        assert np.array_equal(ol.texture_sample(k, uv[:50] * 0 + 0.3125), ol.texture_sample(k, uv[:50] * 0 + 1.3125))
    nonfinite = ol.texture_sample(2, np.array([[np.nan, 0.0], [np.inf, 0.0], [0.0, 0.0]], np.float32))
    assert np.array_equal(nonfinite[0], nonfinite[2]) and np.array_equal(nonfinite[1], nonfinite[2])
    ol.texture_set([])


def quad_scene(albedo=None, texture=0):
    box = ap.synth_box_scene(1, 2, 15.0)
    q = scenes.triangles([[-0.8, -0.9, -0.8], [-0.8, -0.9, -0.8]], [[0.8, -0.9, -0.8], [0.8, -0.9, 0.6]],
                         [[0.8, -0.9, 0.6], [-0.8, -0.9, 0.6]])
    q["v0_uv"] = [[0, 0], [0, 0]]
    q["v1_uv"] = [[1, 0], [1, 1]]
    q["v2_uv"] = [[1, 1], [0, 1]]
    if albedo is not None:
        q["albedo"] = albedo
    q["textures"][:, 0] = texture
    return scenes.concat([box, q])


def test_constant_texture_renders_as_its_albedo():
    # Channels of 0 and 255 blend exactly (the 8-bit weights sum to 1 in
    # floating point), so the textured albedo is the texel's bit for bit and
    # the images are identical; another constant agrees to rounding.
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 24, 20)
    prm = ap.RenderParams.defaults(spp=3, max_depth=10, seed=4)
    for c, exact in (([255, 0, 255, 255], True), ([200, 90, 30, 255], False)):
        c = np.array(c, np.uint8)
        ol.texture_set([np.broadcast_to(c, (4, 4, 4)).copy()])
        al = ((c[:3].astype(np.float32) / np.float32(255.0)) ** 2).astype(np.float32)
        a = ol.OracleBVH(quad_scene(albedo=[0.1, 0.9, 0.2], texture=1)).render(cam, prm)
        b = ol.OracleBVH(quad_scene(albedo=al)).render(cam, prm)
        c0 = ol.OracleBVH(quad_scene(albedo=[0.1, 0.9, 0.2])).render(cam, prm)
        if exact:
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        else:
            assert np.allclose(a[0].mean(axis=(0, 1)), b[0].mean(axis=(0, 1)), rtol=1e-3)
        assert not np.allclose(a[0], c0[0])   # the texture, not the material albedo, shades the quad
    ol.texture_set([])
