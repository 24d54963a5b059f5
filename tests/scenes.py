"""Small hand-built scenes for tests (numpy SceneObject arrays)."""
import numpy as np

import aperture as ap


def concat(parts):
    """np.concatenate that keeps the 432-byte SceneObject layout."""
    return np.concatenate([np.asarray(p, dtype=ap.SCENE_OBJECT_DTYPE) for p in parts]).astype(ap.SCENE_OBJECT_DTYPE)


def blank(n):
    """n default SceneObjects (Material/GBuffer defaults, identity transforms)."""
    o = np.zeros(n, dtype=ap.SCENE_OBJECT_DTYPE)
    o["shading_model"] = ap.MAT_DIFFUSE
    o["albedo"] = 1.0
    o["ior"] = 1.5
    o["sigma_a"] = 100.0
    o["normal"] = (0.0, 0.0, 1.0)
    o["abbe_number"] = 20.0
    o["world_to_object"] = np.eye(3, dtype=np.float32)
    o["object_to_world"] = np.eye(3, dtype=np.float32)
    o["cached_volume_index"] = -1
    o["object_index"] = -1
    o["light_index"] = -1
    o["visible"] = 1
    return o


def triangles(p0, p1, p2, center=None):
    """Triangle SceneObjects from world-space vertex arrays (N,3). The centre
    is the centroid unless given; local vertices are relative to it."""
    p0, p1, p2 = (np.asarray(x, dtype=np.float32).reshape(-1, 3) for x in (p0, p1, p2))
    n = len(p0)
    c = ((p0 + p1 + p2) / np.float32(3.0)).astype(np.float32) if center is None else np.broadcast_to(
        np.asarray(center, np.float32), (n, 3))
    o = blank(n)
    o["type"] = ap.OBJ_TRIANGLE
    o["center"] = c
    o["extent"] = 1.0
    o["v0_position"] = p0 - c
    o["v1_position"] = p1 - c
    o["v2_position"] = p2 - c
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    for k in ("v0_normal", "v1_normal", "v2_normal"):
        o[k] = nrm
    update_proxy(o)
    return o


def update_proxy(o):
    """SceneObject::UpdateProxy for identity-transform objects (object.cuh:309-342)."""
    c = o["center"].astype(np.float32)
    o["proxy_center"] = c
    tri = o["type"] == ap.OBJ_TRIANGLE
    if tri.any():
        ps = [o[k][tri] + c[tri] for k in ("v0_position", "v1_position", "v2_position")]
        o["proxy_min"][tri] = np.minimum(np.minimum(ps[0], ps[1]), ps[2])
        o["proxy_max"][tri] = np.maximum(np.maximum(ps[0], ps[1]), ps[2])
    cube = o["type"] == ap.OBJ_CUBE
    o["proxy_min"][cube] = c[cube] - o["extent"][cube]
    o["proxy_max"][cube] = c[cube] + o["extent"][cube]
    sph = o["type"] == ap.OBJ_SPHERE
    r = o["extent"][sph].max(axis=1, keepdims=True)
    o["proxy_min"][sph] = c[sph] - r
    o["proxy_max"][sph] = c[sph] + r
    return o


def cubes(centers, extents, emissive=None):
    centers = np.asarray(centers, np.float32).reshape(-1, 3)
    o = blank(len(centers))
    o["type"] = ap.OBJ_CUBE
    o["center"] = centers
    o["extent"] = np.abs(np.broadcast_to(np.asarray(extents, np.float32), centers.shape))
    if emissive is not None:
        o["shading_model"] = ap.MAT_LIGHT
        o["albedo"] = 0.0
        o["emissive"] = emissive
    return update_proxy(o)


def spheres(centers, radii, emissive=None):
    """Sphere SceneObjects; with `emissive`, lights (Material::CreateLight:
    MAT_LIGHT, albedo 0), sampled by NEESampleSphereCapArea."""
    centers = np.asarray(centers, np.float32).reshape(-1, 3)
    o = blank(len(centers))
    o["type"] = ap.OBJ_SPHERE
    o["center"] = centers
    o["extent"] = np.asarray(radii, np.float32).reshape(-1, 1)
    if emissive is not None:
        o["shading_model"] = ap.MAT_LIGHT
        o["albedo"] = 0.0
        o["emissive"] = emissive
    return update_proxy(o)


def sphere_lit_box(seed=1, subdiv=2, keep_cube_light=False, ggx=True):
    """The closed-box workload lit by two emissive spheres (one under the
    ceiling, one small and bright in a corner), optionally beside its ceiling
    cube light, with every third mesh GGX: NEE picks sphere caps (and box
    facets), BSDF rays that hit the spheres take the cap pdf for MIS."""
    objs = ap.synth_box_scene(seed, subdiv, 15.0)
    if not keep_cube_light:
        objs = objs[objs["shading_model"] != ap.MAT_LIGHT]
    if ggx:
        g = (np.arange(len(objs)) % 3 == 0) & (objs["shading_model"] == ap.MAT_DIFFUSE)
        objs["shading_model"][g] = ap.MAT_GGX
        objs["roughness"][g] = np.linspace(0.1, 0.9, int(g.sum()), dtype=np.float32)
        objs["metallic"][g] = 0.4
        objs["specular"][g] = 0.5
    lights = spheres([[0.0, 0.7, 0.0], [0.7, -0.75, -0.6]], [0.18, 0.06], emissive=[[12.0, 10.0, 8.0], [40.0, 60.0, 80.0]])
    return concat([objs, lights])


def sdfs(centers, extents, sdf_type=0, info=(0.375, 0.125, 0.0, 0.0), volume=-1, rotations=None):
    """SDF SceneObjects: torus (sdf_type 0, info.xy = R, r in unit space) or
    volume (sdf_type 1, volume = cachedVolumeIndex_), optionally rotated
    (rotations: object-to-world row matrices; world-to-object is the
    transpose). The proxy is the box of the 8 transformed corners."""
    centers = np.asarray(centers, np.float32).reshape(-1, 3)
    n = len(centers)
    o = blank(n)
    o["type"] = ap.OBJ_SDF
    o["center"] = centers
    o["extent"] = np.abs(np.broadcast_to(np.asarray(extents, np.float32), centers.shape))
    o["sdf_type"] = sdf_type
    o["sdf_info0"] = np.broadcast_to(np.asarray(info, np.float32), (n, 4))
    o["cached_volume_index"] = volume
    rot = np.broadcast_to(np.eye(3, dtype=np.float32), (n, 3, 3)) if rotations is None else \
        np.asarray(rotations, np.float32).reshape(n, 3, 3)
    o["object_to_world"] = rot
    o["world_to_object"] = np.transpose(rot, (0, 2, 1))
    signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    corners = centers[:, None, :] + np.einsum("nij,nkj->nki", rot, signs[None] * o["extent"][:, None, :])
    o["proxy_min"] = corners.min(1)
    o["proxy_max"] = corners.max(1)
    o["proxy_center"] = centers
    return o


def quad_grid(n, z=0.0, size=1.0):
    """n x n axis-aligned quads (2 triangles each) on the plane z, spanning
    [0,size]^2: every interior edge and vertex is shared."""
    xs = np.linspace(0.0, size, n + 1, dtype=np.float32)
    p0, p1, p2 = [], [], []
    for j in range(n):
        for i in range(n):
            a = (xs[i], xs[j], z)
            b = (xs[i + 1], xs[j], z)
            c = (xs[i + 1], xs[j + 1], z)
            d = (xs[i], xs[j + 1], z)
            p0 += [a, a]
            p1 += [b, c]
            p2 += [c, d]
    return triangles(p0, p1, p2)


def rays(o, d, tmin=0.0, tmax=np.inf):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = max(len(o), len(d))
    r = np.zeros(n, dtype=ap.RAY_DTYPE)
    r["o"] = np.broadcast_to(o, (n, 3))
    r["d"] = np.broadcast_to(d, (n, 3))
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def random_rays(n, seed, lo=-0.5, hi=1.5, normalize=True):
    """Rays with origins in [lo,hi]^3 and random directions, including
    axis-aligned and zero-component directions."""
    g = np.random.default_rng(seed)
    o = g.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = g.normal(size=(n, 3)).astype(np.float32)
    k = n // 8
    d[:k] = 0.0
    d[np.arange(k), g.integers(0, 3, k)] = g.choice([-1.0, 1.0], k)  # axis aligned
    d[k:2 * k, g.integers(0, 3)] = 0.0                                  # one zero component
    d[2 * k:3 * k, 0] = -0.0
    if normalize:
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays(o, d)
