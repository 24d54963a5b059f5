"""Benchmark of the Aperture MI355X core.

Workload (one GPU): BVH closest-hit over a 1,000,000-triangle soup (centroids
uniform in the unit cube, edges log-uniform in [1e-3, 1e-2], random
orientation, PCG32 seed 2) with the reference's binned-SAH BVH, tracing two
ray sets per step: 4096x4096 coherent primaries from a pinhole at
(0.5, 0.5, -1.5) looking +z with a 40 degree FOV, and one cosine-weighted
secondary per primary hit (origin offset 1e-4 along the geometric normal,
PCG32 seed 3). A step is one closest-hit pass over both sets with rays
already resident in HBM; value is rays per second over the timed steps.

The roofline object prices the traversal kernel by its algorithmic bytes:
per ray 32 (ray) + 8 (hit) + 32 per node record in the kernel's visit order
+ 36 per triangle test, counted by the kernel's own stats variant in an
untimed pass, divided by the kernel time measured with HIP events on the
launch stream. The CPU baseline runs the oracle's restatement of the
reference traversal on a bounded random sample of the same rays.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "aperturerendererv3-optix-geometry_amd"))

METRIC = "Mrays/s closest-hit; samples/s at 1080p; equal-time relMSE, NRC vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_profile():
    """Figures measured by rocprofv3 --pmc (profiles/trace_closest_c2_traffic.txt), if
    committed: HBM bytes per launch, VALU issue fraction and SIMD lane use."""
    path = os.path.join(REPO, "profiles", "trace_closest_c2_traffic.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            key, _, val = line.partition(":")
            if key in ("traffic_bytes_per_launch", "valu_issue_primary", "valu_issue_secondary", "lane_use_primary",
                       "lane_use_secondary"):
                out[key] = float(val)
    return out


def cpu_baseline(objs, rays_a, rays_b, seconds, threads):
    """The oracle's restatement of the reference TraverseBVH (its own slab
    test, OpenMP over rays) on a random sample of both ray sets, sized by a
    calibration pass to take about `seconds`."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol

    bvh = ol.OracleBVH(objs)
    g = np.random.default_rng(7)
    half = 1 << 16
    pool = np.concatenate([rays_a[g.choice(len(rays_a), min(half, len(rays_a)), replace=False)],
                           rays_b[g.choice(len(rays_b), min(half, len(rays_b)), replace=False)]])
    g.shuffle(pool)
    # Calibrate on growing batches (the first also warms the thread pool).
    n, dt = 4 * threads, 0.0
    while n < len(pool):
        t0 = time.perf_counter()
        bvh.trace(pool[:n], threads=threads)
        dt = time.perf_counter() - t0
        if dt > 1.0:
            break
        n = min(len(pool), n * 4)
    rate = n / max(dt, 1e-9)
    n = int(min(len(pool), max(n, rate * seconds)))
    sample = pool[:n]
    t0 = time.perf_counter()
    bvh.trace(sample, threads=threads)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    bvh.trace(sample, threads=threads, cull_margin=np.float32(2.0 ** -12))
    dt_cull = time.perf_counter() - t0
    return {
        "value": n / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} rays drawn at random from the primary and secondary sets (equal shares), oracle "
                  f"restatement of the reference TraverseBVH (its own slab test), OpenMP on {threads} threads "
                  f"(the host's CPU share per GPU; nproc {os.cpu_count()}), {dt:.1f} s; the same "
                  f"sample with the core's line culling ran at {n / dt_cull / 1e6:.4f} Mrays/s",
    }


def rel_mse(ap, img, ref, n, stream):
    """relMSE of two device XYZ images (mean over pixels and linear-sRGB
    channels of (I - I_ref)^2 / (I_ref^2 + 1e-2)), reduced in FP64 on the GPU
    by ap_relmse_f64."""
    return ap.relmse_f64(img, ref, n, eps=1e-2, stream=stream)


def c1_bench(ap, threads, cpu=True):
    """The closed-box workload as specified for the CPU baseline: 5 walls (10
    triangles), a 0.5x0.5 emissive ceiling quad (thin cube, radiance 15),
    two 1280-triangle icospheres, Lambert albedos from seed 1, 256x256
    pinhole (40 deg), 16 spp, max depth 8, one NEE sample at every vertex, no
    sky. The GPU renders it (HIP events, after one untimed frame); the
    oracle's restatement of the reference integrator renders the same 16
    frames with the reference TraverseBVH (its own slab test, no culling) on
    `threads` host threads. Both trace the same paths, so the CPU's rays are
    the GPU's ray count, and the two images and primary-hit ID buffers must
    be bit-identical."""
    objs = ap.synth_box_scene(1, 3, 15.0)
    W = H = 256
    cam = ap.Camera.make([0.0, 0.0, 3.4], [0.0, 0.0, -1.0], 40.0, W, H)
    prm = ap.RenderParams.defaults(spp=16, max_depth=8, nee_always=1, sky=0, seed=1)
    scene = ap.SceneRenderer(objs)
    stream = ap.Stream()
    n = W * H
    hdr, prim, counts = ap.DeviceBuffer(n * 12), ap.DeviceBuffer(n * 4), ap.DeviceBuffer(16)
    hdr.zero(stream)
    scene.render(cam, ap.RenderParams.defaults(spp=1, max_depth=8, nee_always=1, sky=0, seed=99), hdr,
                 stream=stream)
    hdr.zero(stream)
    counts.zero(stream)
    e0, e1 = ap.Event(), ap.Event()
    e0.record(stream)
    scene.render(cam, prm, hdr, stream=stream, primary_id=prim, ray_count=counts)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_ms(e1)
    rays = int(counts.to_numpy(np.uint64, 2).sum())
    out = {"workload": "closed box: 5 walls (10 triangles), 0.5x0.5 ceiling light (thin cube, radiance 15), two "
                       "1280-triangle icospheres, Lambert albedos from seed 1, 256x256 pinhole 40 deg, 16 spp, max "
                       "depth 8, NEE at every vertex, no sky",
           "rays": rays,
           "gpu": {"ms": round(ms, 3), "samples_per_s": round(n * 16 / (ms / 1e3), 1),
                   "mrays_per_s": round(rays / (ms / 1e3) / 1e6, 1)}}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as ol

        bvh = ol.OracleBVH(objs)
        t0 = time.perf_counter()
        o_hdr, _, o_prim = bvh.render(cam, prm, cull_margin=0.0, threads=threads)
        dt = time.perf_counter() - t0
        g_hdr = hdr.to_numpy(np.float32, n * 3).reshape(H, W, 3)
        g_prim = prim.to_numpy(np.int32, n).reshape(H, W)
        out["cpu"] = {"s": round(dt, 2), "samples_per_s": round(n * 16 / dt, 1), "mrays_per_s": round(rays / dt / 1e6, 4),
                      "threads": threads, "nproc": os.cpu_count(), "kind": "port",
                      "what": "oracle restatement of the reference integrator and its TraverseBVH (no culling), "
                              "OpenMP over pixels, the full 16 spp"}
        out["image_bit_identical"] = bool(np.array_equal(g_hdr, o_hdr))
        out["primary_ids_bit_identical"] = bool(np.array_equal(g_prim, o_prim))
        out["gpu_over_cpu"] = round(dt / (ms / 1e3), 1)
    scene.close()
    log(f"c1: {out}")
    return out


def reference_image(ap, scene, cam, stream, params, ref_spp):
    """A plain render of ref_spp frames (four wavelengths, seed 1000): the
    reference the equal-time comparisons measure relMSE against."""
    ref = ap.DeviceBuffer(cam.width * cam.height * 12)
    ref.zero(stream)
    t0 = time.perf_counter()
    scene.render(cam, ap.RenderParams.defaults(spp=ref_spp, seed=1000, wavelengths=4, **params), ref, stream=stream)
    stream.synchronize()
    log(f"nrc: reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    return ref


def nrc_bench(ap, scene, cam, stream, params, plain_spp, ref, ref_spp, warmup, train_passes, boost_frames=0,
              boost_steps=16, spread=0.01, width=64, precision="bf16", guided=False, lr=0.005):
    """Equal-time relMSE of the radiance-cache renderer against the plain
    wavefront tracer, both with four hero wavelengths per path and, with
    `guided`, both with directional guiding (probes reset before each timed
    run and rebuilt after every frame, inside the timed region). The plain
    tracer renders plain_spp frames (timed with HIP events); the cache
    renderer starts from freshly initialised weights (Adam learning rate
    `lr`) and renders frame after frame, training after each (and on
    train_passes training-only passes), for as many frames as its calibrated
    per-frame time fits in the same GPU time: its first `warmup` frames go to
    a scratch image, the rest are averaged; the first `boost_frames` averaged
    frames train with `boost_steps` Adam steps instead of 4, their extra cost
    calibrated and charged to the budget. Both images are compared with the
    plain reference image `ref`; relMSE is reduced in FP64 on the GPU."""
    W, H = cam.width, cam.height
    n = W * H
    g = 1 if guided else 0
    img = ap.DeviceBuffer(n * 12)
    img.zero(stream)
    if guided:
        scene.guiding_reset()
    e0, e1 = ap.Event(), ap.Event()
    e0.record(stream)
    scene.render(cam, ap.RenderParams.defaults(spp=plain_spp, seed=5, wavelengths=4, use_guiding=g, **params), img,
                 stream=stream)
    e1.record(stream)
    stream.synchronize()
    budget = e0.elapsed_ms(e1)
    plain = rel_mse(ap, img, ref, n, stream)
    cache = ap.RadianceCache(seed=1, width=width, precision=precision)
    nprm = ap.NrcRenderParams.defaults(train_paths=65536, train_passes=train_passes, spread=spread)
    nprm.learning_rate = lr
    scratch = ap.DeviceBuffer(n * 12)
    scratch.zero(stream)
    w0 = cache.params()

    boost = ap.NrcRenderParams.from_buffer_copy(nprm)
    boost.train_steps = boost_steps

    def nrc_frames(frames, buf, seed, first=0, p=nprm):
        if frames > 0:
            scene.render_nrc(cam, ap.RenderParams.defaults(spp=frames, frame_index=first, seed=seed, wavelengths=4,
                                                           use_guiding=g, **params), cache, p, buf, stream=stream)

    # Calibration (untimed; it also allocates the path, query and training
    # buffers): GPU time per frame of the cache renderer, frames submitted
    # back to back as in the timed run.
    nrc_frames(2, scratch, 3)
    stream.synchronize()
    e0.record(stream)
    nrc_frames(8, scratch, 4, 32)
    e1.record(stream)
    stream.synchronize()
    per_frame = e0.elapsed_ms(e1) / 8
    e0.record(stream)
    nrc_frames(4, scratch, 4, 40, boost)
    e1.record(stream)
    stream.synchronize()
    per_boost = e0.elapsed_ms(e1) / 4
    # Frames that fit the budget: warmup and boosted frames at their own cost.
    nb_fit = boost_frames
    rest = budget - warmup * per_frame - nb_fit * per_boost
    total = max(warmup + 1, warmup + nb_fit + int(max(rest, 0.0) / per_frame))
    # Timed run from the initial weights with a fresh optimiser (and empty
    # probes): `warmup` frames into a scratch image, the rest averaged into
    # the compared image.
    cache.set_params(w0)
    if guided:
        scene.guiding_reset()
    scratch.zero(stream)
    img.zero(stream)
    stream.synchronize()
    e0.record(stream)
    nrc_frames(warmup, scratch, 82)
    # The first boost_frames averaged frames train with boost_steps steps each.
    nb = min(boost_frames, total - warmup)
    nrc_frames(nb, img, 5, 0, boost)
    nrc_frames(total - warmup - nb, img, 5, nb)
    e1.record(stream)
    stream.synchronize()
    used = e0.elapsed_ms(e1)
    frames = total
    st, _ = scene.nrc_last_frame(samples=False)
    nrc = rel_mse(ap, img, ref, n, stream) if frames > warmup else float("nan")
    cache.close()
    if guided:
        scene.guiding_reset()
    out = {
        "guided": guided,
        "equal_time_ms": round(budget, 1),
        "plain_spp": plain_spp,
        "plain_relmse": round(plain, 5),
        "nrc_frames": frames - warmup,
        "nrc_warmup_frames": warmup,
        "nrc_training_passes_per_frame": train_passes,
        "nrc_boost": {"frames": boost_frames, "steps": boost_steps},
        "nrc_ms": round(used, 1),
        "nrc_ms_per_frame_calibrated": round(per_frame, 3),
        "nrc_ms_per_boosted_frame_calibrated": round(per_boost, 3),
        "nrc_relmse": round(nrc, 5),
        "relmse_ratio_plain_over_nrc": round(plain / nrc, 3) if nrc > 0 else None,
        "reference_spp": ref_spp,
        "reference_self_noise_estimate": round(plain * plain_spp / ref_spp, 6),
        "cache_queries_per_frame": st.queries,
        "training_paths_per_frame": st.train_paths,
        "training_samples_per_frame": st.samples,
        "cache": f"64-{width}x5-3 ReLU, {precision.upper()} MFMA / FP32 accumulate, relative L2, Adam lr {lr:g}, "
                 "4 steps x 16384 per frame from 65,536 training paths (the first boost frames take more steps, at "
                 f"their own cost), spread threshold {spread}",
        "spectral": "four hero wavelengths per path in the reference, plain and cache images",
    }
    log(f"nrc: {out}")
    return out


def materials_bench(ap, spp, stream, W=1920, H=1080):
    """Frame time at 1920x1080 of the C1 closed box holding glass and a
    textured mesh: a dispersive glass sphere (IOR 1.5, Abbe 20), a rotated
    IOR-2.4 glass box, a thin glass quad, and a 32x32-quad panel read from an
    OBJ file by ap_mesh_load_obj with a 1024^2 RGBA8 albedo texture; max
    depth 12, RR from depth 3, sky on; one and four wavelengths. The same box
    without them is timed alongside."""
    import tempfile

    n = 32
    lines = []
    for j in range(n + 1):
        for i in range(n + 1):
            lines.append(f"v {-0.9 + 1.8 * i / n:.6f} {-0.9 + 1.8 * j / n:.6f} "
                         f"{-0.93 + 0.03 * math.sin(i * 0.9 + j * 0.4):.6f}")
            lines.append(f"vt {2.0 * i / n:.6f} {2.0 * j / n:.6f}")
    lines.append("vn 0 0 1")
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i + 1
            lines.append(f"f {a}/{a}/1 {a + n + 1}/{a + n + 1}/1 {a + n + 2}/{a + n + 2}/1 {a + 1}/{a + 1}/1")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "panel.obj")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        yy, xx = np.mgrid[0:1024, 0:1024]
        img = np.stack([(xx // 64 + yy // 64) % 2 * 200 + 40, (xx * 255) // 1023, (yy * 255) // 1023,
                        np.full_like(xx, 255)], -1).astype(np.uint8)
        tex = ap.TextureCache()
        handle = tex.add_rgba8(img)
        mesh = ap.load_obj(path, (0.0, 0.0, 0.0), 1.0, (handle, 0, 0, 0))
    def glass(o, ior, sigma_a, albedo, abbe):
        o["shading_model"] = ap.MAT_GLASS
        o["ior"], o["sigma_a"], o["albedo"], o["abbe_number"] = ior, sigma_a, albedo, abbe
        return o
    box = ap.synth_box_scene(1, 3, 15.0)

    def blank():
        # A default SceneObject: GBuffer defaults, identity transforms, no light.
        o = np.zeros(1, ap.SCENE_OBJECT_DTYPE)
        o["albedo"], o["ior"], o["sigma_a"], o["normal"], o["abbe_number"] = 1.0, 1.5, 100.0, (0, 0, 1), 20.0
        o["world_to_object"] = o["object_to_world"] = np.eye(3, dtype=np.float32)
        o["cached_volume_index"] = o["object_index"] = o["light_index"] = -1
        o["visible"] = 1
        return o

    sph, cube, quad = blank(), blank(), blank()
    sph["type"], sph["center"], sph["extent"] = ap.OBJ_SPHERE, (-0.45, -0.55, 0.2), 0.3
    c, s_ = math.cos(0.6), math.sin(0.6)
    rot = np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]], np.float32)
    cube["type"], cube["center"], cube["extent"] = ap.OBJ_CUBE, (0.45, -0.6, 0.0), (0.2, 0.3, 0.2)
    cube["object_to_world"], cube["world_to_object"] = rot, rot.T
    quad["type"], quad["center"], quad["extent"] = ap.OBJ_TRIANGLE, (0.0, 0.3, 0.4), 1.0
    quad["v0_position"], quad["v1_position"], quad["v2_position"] = (-0.4, -0.3, 0), (0.4, -0.3, 0), (0.0, 0.35, 0)
    for k in ("v0_normal", "v1_normal", "v2_normal"):
        quad[k] = (0.0, 0.0, 1.0)
    glass(sph, 1.5, 1.0, (0.9, 0.95, 1.0), 20.0)
    glass(cube, 2.4, 2.0, (1.0, 0.8, 0.6), 0.0)
    glass(quad, 1.45, 1.0, (1.0, 1.0, 1.0), 30.0)
    # UpdateProxy: the triangle's vertex box; the rotated cube's box of its rotated extent; the sphere's radius.
    ps = np.stack([quad[k][0] + quad["center"][0] for k in ("v0_position", "v1_position", "v2_position")])
    quad["proxy_min"], quad["proxy_max"] = ps.min(0), ps.max(0)
    half = np.abs(rot) @ cube["extent"][0]
    cube["proxy_min"], cube["proxy_max"] = cube["center"][0] - half, cube["center"][0] + half
    sph["proxy_min"], sph["proxy_max"] = sph["center"][0] - 0.3, sph["center"][0] + 0.3
    for o in (sph, cube, quad):
        o["proxy_center"] = o["center"]
    objs = np.concatenate([box, sph, cube, quad, mesh]).astype(ap.SCENE_OBJECT_DTYPE)
    cam = ap.Camera.make([0.0, 0.0, 3.4], [0.0, 0.0, -1.0], 40.0, W, H)
    hdr = ap.DeviceBuffer(W * H * 12)
    params = dict(max_depth=12, rr_min_depth=3, sky=1)
    out = {"workload": f"C1 closed box + glass sphere (IOR 1.5, Abbe 20), rotated glass box (IOR 2.4), thin glass "
                       f"triangle, {len(mesh)}-triangle OBJ panel with a 1024^2 albedo texture, 1920x1080, max depth "
                       f"12, RR from depth 3, sky", "spp": spp}
    for label, o in (("with_glass_and_texture", objs), ("plain_box", box)):
        sc = ap.SceneRenderer(o)
        sc.set_texture_cache(tex)
        for wl in (1, 4):
            # An untimed call of the same shape (frames 1..spp), then frames spp+1..2spp timed.
            sc.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1, wavelengths=wl, **params), hdr,
                      stream=stream)
            stream.synchronize()
            e0, e1 = ap.Event(), ap.Event()
            e0.record(stream)
            sc.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1 + spp, wavelengths=wl, **params), hdr,
                      stream=stream)
            e1.record(stream)
            stream.synchronize()
            ms = e0.elapsed_ms(e1) / spp
            out[f"{label}_{wl}_wavelength{'s' if wl > 1 else ''}"] = {
                "ms_per_spp": round(ms, 3), "samples_per_s": round(W * H / (ms / 1e3), 1)}
        sc.close()
    tex.close()
    log(f"glass and textures: {out}")
    return out


COUNTS_FIXTURE = os.path.join(REPO, "tests", "golden", "integrator_counts_heightfield.json")


def integrator_roofline(res, key, samples):
    """Algorithmic bytes of one frame of the heightfield render, priced per
    path vertex: each closest-hit and shadow ray 32 (ray) + 8 (hit) + 32 per
    node record + 36 per triangle test; each shaded vertex 128 (material and
    type fetch) + 128 per spectral Interact (eight 16-byte LUT cells); each
    probe query 8 (index) + 64 when it finds a probe (four texels); each
    sample 24 (framebuffer read-modify-write). Per-ray node and test counts
    and per-vertex Interact and probe counts are the oracle's accounting of
    the same frame at 128x72 (the fixture, whose ray counts the GPU matches
    exactly at that size); ray counts are the GPU's own at 1920x1080. The
    result over the frame time on the render stream, against HBM peak."""
    with open(COUNTS_FIXTURE) as f:
        c = json.load(f)["views"]["view128x72"][key]
    ext, shadow = res["rays_per_spp"]
    per_ext = 40.0 + (32.0 * c["ext_nodes"] + 36.0 * c["ext_prims"]) / c["ext_rays"]
    per_shadow = 40.0 + (32.0 * c["shadow_nodes"] + 36.0 * c["shadow_prims"]) / c["shadow_rays"]
    hits = ext * c["hits"] / c["ext_rays"]
    interacts = hits * c["interacts"] / c["hits"]
    probes = hits * c["probe_queries"] / c["hits"]
    found = hits * c["probe_found"] / c["hits"]
    parts = {
        "closest_hit_traversal": ext * per_ext,
        "shadow_traversal": shadow * per_shadow,
        "material": 128.0 * hits,
        "spectral_interact": 128.0 * interacts,
        "probes": 8.0 * probes + 64.0 * found,
        "framebuffer": 24.0 * samples,
    }
    total = sum(parts.values())
    gbs = total / (res["ms_per_spp"] / 1e3) / 1e9
    return {
        "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_spp": round(total),
        "bytes_per_vertex": round(total / hits, 1),
        "share": {k: round(v / total, 4) for k, v in parts.items()},
        "nodes_per_ray": [round(c["ext_nodes"] / c["ext_rays"], 1), round(c["shadow_nodes"] / c["shadow_rays"], 1)],
        "interacts_per_vertex": round(c["interacts"] / c["hits"], 2),
        "counts": "oracle accounting of frame 49 at 128x72 (tests/golden/integrator_counts_heightfield.json)",
    }


def render_only(ap, mode, spp):
    """One mode of the 1080p heightfield render alone, for counter runs: a
    warm-up frame (for guiding, the frame that builds the first probes),
    then `spp` frames timed with HIP events. Every frame starts with one
    generate_kernel launch, which delimits frames in a kernel trace."""
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    scene = ap.SceneRenderer(objs)
    W, H = 1920, 1080
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, W, H)
    stream = ap.Stream()
    hdr = ap.DeviceBuffer(W * H * 12)
    hdr.zero(stream)
    counts = ap.DeviceBuffer(16)
    kw = dict(max_depth=16, rr_min_depth=4, sky=0, wavelengths=int(mode[-1]),
              use_guiding=1 if mode.startswith("guided") else 0)
    scene.render(cam, ap.RenderParams.defaults(spp=1, **kw), hdr, stream=stream)
    counts.zero(stream)
    e0, e1 = ap.Event(), ap.Event()
    e0.record(stream)
    scene.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1, **kw), hdr, stream=stream, ray_count=counts)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_ms(e1)
    rays = counts.to_numpy(np.uint64, 2)
    res = {"ms_per_spp": round(ms / spp, 3), "rays_per_spp": [int(rays[0]) // spp, int(rays[1]) // spp]}
    res["integrator_roofline"] = integrator_roofline(res, mode, W * H)
    scene.close()
    return {"render_only": mode, "workload": "1080p heightfield render, max depth 16, RR from depth 4, no sky",
            "frames": spp, **res}


def render_bench(ap, spp, cpu_seconds, threads, nrc_spp=0, nrc_ref_spp=2048, nrc_warmup=0, nrc_passes=0,
                 nrc_boost=(8, 16)):
    """Samples/s of the wavefront path tracer at 1920x1080 on the heightfield
    workload (1024^2-cell fBm heightfield, 2,097,152 triangles, plus 256
    320-triangle icospheres; Lambert/GGX 50/50; lit by 4 thin-cube lights, no
    sky), max depth 16, Russian roulette from depth 4. Measured with one
    wavelength per path (the reference's estimator), with four hero
    wavelengths per path, and with directional guiding (one and four
    wavelengths; probes rebuilt after every frame, the rebuild inside the
    timed region).
    Each mode renders `spp` untimed frames in one call, then `spp` more in
    one call timed with HIP events on the render stream (guided modes start
    from a probe reset and one untimed frame before both)."""
    t0 = time.perf_counter()
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    scene = ap.SceneRenderer(objs)
    log(f"render scene: {len(objs)} objects ({time.perf_counter() - t0:.1f} s)")
    W, H = 1920, 1080
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, W, H)
    stream = ap.Stream()
    hdr = ap.DeviceBuffer(W * H * 12)
    hdr.zero(stream)
    counts = ap.DeviceBuffer(16)
    params = dict(max_depth=16, rr_min_depth=4, sky=0)
    scene.render(cam, ap.RenderParams.defaults(spp=1, **params), hdr, stream=stream)
    stream.synchronize()

    def timed(**kw):
        # An untimed call of the same shape first (frames 1..spp): buffers
        # sized for a batch of spp frames, and for guided frames the second
        # frame set of the two in flight, are allocated outside the timed
        # call, which renders frames spp+1..2spp.
        scene.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1, **params, **kw), hdr, stream=stream)
        stream.synchronize()
        counts.zero(stream)
        e0, e1 = ap.Event(), ap.Event()
        t0 = time.perf_counter()
        e0.record(stream)
        scene.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1 + spp, **params, **kw), hdr, stream=stream,
                     ray_count=counts)
        e1.record(stream)
        stream.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
        ms = e0.elapsed_ms(e1)
        rays = counts.to_numpy(np.uint64, 2)
        return {
            "ms_per_spp": round(ms / spp, 3),
            "wall_ms_per_spp": round(wall / spp, 3),
            "samples_per_s": round(W * H * spp / (ms / 1e3), 1),
            "mrays_per_s": round(float(rays[0] + rays[1]) / (ms / 1e3) / 1e6, 1),
            "rays_per_sample": round(float(rays[0] + rays[1]) / (W * H * spp), 3),
            "rays_per_spp": [int(rays[0]) // spp, int(rays[1]) // spp],
        }

    one = timed(wavelengths=1)
    four = timed(wavelengths=4)
    scene.render(cam, ap.RenderParams.defaults(spp=1, use_guiding=1, **params), hdr, stream=stream)
    stream.synchronize()
    guided = timed(use_guiding=1)
    guided["probes"] = len(scene.guiding_export()[1])
    scene.guiding_reset()
    scene.render(cam, ap.RenderParams.defaults(spp=1, use_guiding=1, wavelengths=4, **params), hdr, stream=stream)
    stream.synchronize()
    guided4 = timed(use_guiding=1, wavelengths=4)
    guided4["probes"] = len(scene.guiding_export()[1])
    # Guiding in groups of 8 frames (guide_frames = 8): the group's frames
    # sample the probes built before it and are traced as one 16.6M-path
    # batch, then one update folds the group's samples in.
    scene.guiding_reset()
    scene.render(cam, ap.RenderParams.defaults(spp=1, use_guiding=1, wavelengths=4, **params), hdr, stream=stream)
    stream.synchronize()
    guided4_groups = timed(use_guiding=1, wavelengths=4, guide_frames=8)
    guided4_groups["probes"] = len(scene.guiding_export()[1])
    guided4_groups["guide_frames"] = 8
    # Frames with the screen G-buffer and statistics recorded (the extra
    # per-path first-hit record and the per-pixel updates in accumulate).
    gbuf = ap.DeviceBuffer(W * H * ap.SCREEN_GBUFFER_DTYPE.itemsize)
    stats = ap.DeviceBuffer(W * H * ap.SCREEN_STATISTICS_DTYPE.itemsize)
    gbuf.zero(stream)
    stats.zero(stream)
    scene.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1, wavelengths=4, **params), hdr, stream=stream,
                 gbuffer=gbuf, statistics=stats)
    stream.synchronize()
    e0, e1 = ap.Event(), ap.Event()
    e0.record(stream)
    scene.render(cam, ap.RenderParams.defaults(spp=spp, frame_index=1 + spp, wavelengths=4, **params), hdr,
                 stream=stream, gbuffer=gbuf, statistics=stats)
    e1.record(stream)
    stream.synchronize()
    with_records = {"ms_per_spp": round(e0.elapsed_ms(e1) / spp, 3)}
    # ap_tonemap at 1920x1080 (ACES): 12 B read + 4 B written per pixel.
    ldr = ap.DeviceBuffer(W * H * 4)
    ap.tonemap(hdr, W, H, tone=ap.TONE_ACES, ldr=ldr, stream=stream)
    reps = 50
    stream.synchronize()
    e0.record(stream)
    for _ in range(reps):
        ap.tonemap(hdr, W, H, tone=ap.TONE_ACES, ldr=ldr, stream=stream)
    e1.record(stream)
    stream.synchronize()
    tm_ms = e0.elapsed_ms(e1) / reps
    tonemap = {"ms": round(tm_ms, 4), "gb_per_s": round(W * H * 16 / (tm_ms / 1e3) / 1e9, 1),
               "curve": "ACES", "algorithmic_bytes": W * H * 16}
    roof = {k: integrator_roofline(r, key, W * H) for k, r, key in
            (("one_wavelength", one, "plain_1"), ("four_wavelengths", four, "plain_4"),
             ("guided_one_wavelength", guided, "guided_1"), ("guided_four_wavelengths", guided4, "guided_4"),
             ("guided_four_wavelengths_groups_of_8", guided4_groups, "guided_4"))}
    out = {
        "workload": "heightfield 1024^2 (2,097,152 tris) + 256 icospheres (81,920 tris) + 4 thin-cube lights, "
                    "no sky, 1920x1080, max depth 16, RR from depth 4",
        "spp": spp,
        "integrator_roofline": roof,
        "one_wavelength": one,
        "four_wavelengths": four,
        "guided_one_wavelength": guided,
        "guided_four_wavelengths": guided4,
        "guided_four_wavelengths_groups_of_8": guided4_groups,
        "four_wavelengths_with_gbuffer_and_statistics": with_records,
        "tonemap_1080p": tonemap,
        "glass_and_textures_1080p": materials_bench(ap, spp, stream, W, H),
    }
    log(f"render: {out}")
    if nrc_spp > 0:
        # The heightfield workload with the cache (guided on both sides, Adam
        # lr 1e-3), its learning-rate ablation, and the unguided comparison.
        ref = reference_image(ap, scene, cam, stream, params, nrc_ref_spp)
        out["nrc_vs_plain"] = nrc_bench(ap, scene, cam, stream, params, nrc_spp, ref, nrc_ref_spp, nrc_warmup,
                                        nrc_passes, *nrc_boost, guided=True, lr=1e-3)
        out["nrc_vs_plain_guided_lr_5e-3"] = nrc_bench(ap, scene, cam, stream, params, nrc_spp, ref, nrc_ref_spp,
                                                       nrc_warmup, nrc_passes, *nrc_boost, guided=True, lr=5e-3)
        out["nrc_vs_plain_unguided"] = nrc_bench(ap, scene, cam, stream, params, nrc_spp, ref, nrc_ref_spp,
                                                 nrc_warmup, nrc_passes, *nrc_boost, guided=False, lr=5e-3)
        ref.free()
        # The same comparison in a closed interior: the C1 box (coloured walls,
        # two spheres, one ceiling light) at 1920x1080, where long path tails
        # carry most of the light.
        box = ap.SceneRenderer(ap.synth_box_scene(1, 3, 15.0))
        box_cam = ap.Camera.make([0.0, 0.0, 3.4], [0.0, 0.0, -1.0], 40.0, W, H)
        box_params = dict(max_depth=8, rr_min_depth=3, sky=1)
        box.render(box_cam, ap.RenderParams.defaults(spp=1, wavelengths=4, **box_params), hdr, stream=stream)
        stream.synchronize()
        box_ref = reference_image(ap, box, box_cam, stream, box_params, nrc_ref_spp)
        res = nrc_bench(ap, box, box_cam, stream, box_params, nrc_spp, box_ref, nrc_ref_spp, nrc_warmup, nrc_passes,
                        *nrc_boost, guided=True, lr=1e-3)
        box_ref.free()
        res["workload"] = ("closed box of the C1 workload (5 walls, two 1280-triangle spheres, 0.5x0.5 ceiling "
                           "light) at 1920x1080, max depth 8, RR from depth 3")
        out["nrc_vs_plain_closed_box"] = res
        box.close()
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as ol

        bvh = ol.OracleBVH(objs)
        margin = np.float32(scene.info().line_margin)
        crop = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0 * 48 / 1080, 48, 48)
        t0 = time.perf_counter()
        bvh.render(crop, ap.RenderParams.defaults(spp=1, **params), cull_margin=margin, threads=threads)
        dt = time.perf_counter() - t0
        spp_cpu = max(1, int(cpu_seconds / max(dt, 1e-3)))
        t0 = time.perf_counter()
        bvh.render(crop, ap.RenderParams.defaults(spp=spp_cpu, frame_index=1, **params), cull_margin=margin,
                   threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_samples_per_s"] = round(48 * 48 * spp_cpu / dt, 1)
        out["cpu_sample"] = (f"oracle restatement of the reference integrator with the core's culled traversal, "
                             f"{threads} threads, 48x48 centre crop x {spp_cpu} spp, {dt:.1f} s")
    scene.close()
    return out


def batch_bench(ap, spp, depth, n_scenes=48, seed=6):
    """Batched heterogeneous rendering: n_scenes heightfield + icosphere
    scenes (triangles log-uniform in [1e4, 5e5], 4% of them in spheres;
    resolutions uniform in {128^2, 256^2, 512^2, 1024^2}; 4 thin-cube lights,
    no sky) rendered together by one batch: shared queues, one traversal
    launch per bounce over the rays of all scenes, each frame a replayed HIP
    graph, launch shape autotuned. Timed against the same frames rendered
    scene by scene with ap_render, and against the batch without the graph."""
    t0 = time.perf_counter()
    objs, cams = ap.synth_batch_scenes(n_scenes, seed)
    scenes = [ap.SceneRenderer(o) for o in objs]
    log(f"batch: {n_scenes} scenes, {sum(len(o) for o in objs)} objects, "
        f"{sum(c.width * c.height for c in cams)} pixels per frame ({time.perf_counter() - t0:.1f} s)")
    stream = ap.Stream()
    params = dict(max_depth=depth, rr_min_depth=4, sky=0, wavelengths=4)
    hdrs = []
    for c in cams:
        h = ap.DeviceBuffer(c.width * c.height * 12)
        h.zero(stream)
        hdrs.append(h)
    batch = ap.Batch(scenes, cams)
    w, r, tune_ms = batch.autotune(ap.RenderParams.defaults(spp=1, **params), frames=2, stream=stream)
    log(f"batch: autotune -> {w} rays per wave, refill {r} ({tune_ms:.2f} ms per frame)")
    pix = sum(c.width * c.height for c in cams)
    e0, e1 = ap.Event(), ap.Event()

    def timed(fn, **extra):
        fn(ap.RenderParams.defaults(spp=1, frame_index=0, **params, **extra))  # untimed (graph capture, allocation)
        stream.synchronize()
        e0.record(stream)
        out = fn(ap.RenderParams.defaults(spp=spp, frame_index=1, **params, **extra))
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1), out

    ms_graph, rays = timed(lambda p: batch.render(p, hdrs, stream=stream))
    batch.set_graph(False)
    ms_direct, _ = timed(lambda p: batch.render(p, hdrs, stream=stream))

    def one_by_one(p):
        for sc, c, h in zip(scenes, cams, hdrs):
            sc.render(c, p, h, stream=stream)

    ms_seq, _ = timed(one_by_one)
    # Guided: every scene sampling and rebuilding its own probes each frame
    # (one probe build per scene and frame, the builds spread over eight
    # streams); a frame and its 48 builds replayed from one captured graph,
    # and launched directly.
    ms_guided_direct, _ = timed(lambda p: batch.render(p, hdrs, stream=stream), use_guiding=1)
    batch.set_graph(True)
    ms_guided, _ = timed(lambda p: batch.render(p, hdrs, stream=stream), use_guiding=1)
    batch.close()
    for sc in scenes:
        sc.close()
    out = {
        "workload": f"{n_scenes} heightfield+icosphere scenes, triangles log-uniform in [1e4, 5e5], resolutions "
                    f"uniform in {{128^2, 256^2, 512^2, 1024^2}}, {spp} spp, max depth {depth}, RR from depth 4, "
                    "four hero wavelengths, 4 lights, no sky",
        "scenes": n_scenes,
        "triangles": int(sum(len(o) for o in objs)),
        "pixels_per_frame": int(pix),
        "launch": {"wave_rays": w, "refill": r},
        "graph": {"ms": round(ms_graph, 2), "samples_per_s": round(pix * spp / (ms_graph / 1e3), 1),
                  "mrays_per_s": round((rays[0] + rays[1]) / (ms_graph / 1e3) / 1e6, 1)},
        "direct_launches": {"ms": round(ms_direct, 2), "samples_per_s": round(pix * spp / (ms_direct / 1e3), 1)},
        "scene_by_scene": {"ms": round(ms_seq, 2), "samples_per_s": round(pix * spp / (ms_seq / 1e3), 1)},
        "speedup_over_scene_by_scene": round(ms_seq / ms_graph, 3),
        "guided": {"ms": round(ms_guided, 2), "samples_per_s": round(pix * spp / (ms_guided / 1e3), 1),
                   "direct_launches_ms": round(ms_guided_direct, 2)},
    }
    log(f"batch: {out}")
    return out


def batch_nrc_bench(ap, n_scenes, plain_spp, ref_spp, depth=12, seed=9, boost_frames=8, boost_steps=16,
                    c5=False):
    """Equal-time relMSE of batched rendering with per-scene learned radiance
    caches (one width-32 FP16 network per scene in a grouped cache: the
    scenes' cache queries answered by one grouped inference per frame, every
    network trained on its own scene's samples in the same launches) against
    the plain batch, both with four hero wavelengths per path. n_scenes
    scenes, alternately a heightfield + icosphere scene (triangles
    log-uniform in [1e4, 2e5], open sky side, no sky light) and a closed box
    (coloured walls, one ceiling light), resolutions uniform in {256^2,
    512^2}. The plain batch renders plain_spp frames (HIP events); the cache
    batch starts from fresh networks and renders as many frames as its
    calibrated time per frame fits in the same GPU time, its first
    boost_frames frames training with boost_steps Adam steps (their cost
    calibrated and charged). relMSE per scene against a plain batch of
    ref_spp frames on another seed; reported as the mean over scenes.
    With c5, the scenes are the batched workload's own (synth_batch_scenes:
    48 heightfield scenes, triangles log-uniform in [1e4, 5e5], resolutions
    in {128^2 .. 1024^2}, seed 6)."""
    g = np.random.default_rng(seed)
    objs, cams = [], []
    if c5:
        objs, cams = ap.synth_batch_scenes(n_scenes, seed=6)
    for k in range(0 if c5 else n_scenes):
        res = int(g.choice([256, 512]))
        if k % 2 == 0:
            tris = float(np.exp(g.uniform(np.log(1e4), np.log(2e5))))
            spheres = max(1, int(tris * 0.04 / 320))
            grid = max(8, int(round(np.sqrt(max(tris - 320 * spheres, 128) / 2))))
            objs.append(ap.synth_heightfield_scene(grid, spheres, 300 + k, 20.0))
            cams.append(ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, res, res))
        else:
            objs.append(ap.synth_box_scene(1 + k, 2, 15.0))
            cams.append(ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, res, res))
    scenes = [ap.SceneRenderer(o) for o in objs]
    pix = sum(c.width * c.height for c in cams)
    stream = ap.Stream()
    params = dict(max_depth=depth, rr_min_depth=4, sky=0, wavelengths=4)
    batch = ap.Batch(scenes, cams)
    batch.autotune(ap.RenderParams.defaults(spp=1, **params), frames=2, stream=stream)

    def buffers():
        out = []
        for c in cams:
            h = ap.DeviceBuffer(c.width * c.height * 12)
            h.zero(stream)
            out.append(h)
        return out

    def rel(bufs, refs):
        return [rel_mse(ap, b, r, c.width * c.height, stream) for b, r, c in zip(bufs, refs, cams)]

    t0 = time.perf_counter()
    ref = buffers()
    batch.render(ap.RenderParams.defaults(spp=ref_spp, seed=1000, **params), ref, stream=stream)
    stream.synchronize()
    log(f"batch nrc: {n_scenes} scenes, {pix} pixels per frame, reference {ref_spp} spp in "
        f"{time.perf_counter() - t0:.1f} s")
    e0, e1 = ap.Event(), ap.Event()
    img = buffers()
    batch.render(ap.RenderParams.defaults(spp=1, seed=4, **params), img, stream=stream)  # untimed graph capture
    img = buffers()
    stream.synchronize()
    e0.record(stream)
    batch.render(ap.RenderParams.defaults(spp=plain_spp, seed=5, **params), img, stream=stream)
    e1.record(stream)
    stream.synchronize()
    budget = e0.elapsed_ms(e1)
    plain = rel(img, ref)
    cache = ap.RadianceCache(seed=1, groups=n_scenes, width=32, precision="fp16")
    w0 = cache.params()
    nprm = ap.NrcRenderParams.defaults(train_paths=8192, batch=4096)
    boost = ap.NrcRenderParams.from_buffer_copy(nprm)
    boost.train_steps = boost_steps

    def nrc_frames(frames, bufs, seed, first=0, p=nprm):
        if frames > 0:
            batch.render_nrc(ap.RenderParams.defaults(spp=frames, frame_index=first, seed=seed, **params), cache, p,
                             bufs, stream=stream)

    scratch = buffers()
    nrc_frames(2, scratch, 3)
    stream.synchronize()
    e0.record(stream)
    nrc_frames(6, scratch, 4)
    e1.record(stream)
    stream.synchronize()
    per_frame = e0.elapsed_ms(e1) / 6
    e0.record(stream)
    nrc_frames(3, scratch, 4, 0, boost)
    e1.record(stream)
    stream.synchronize()
    per_boost = e0.elapsed_ms(e1) / 3
    total = max(1, boost_frames + int(max(budget - boost_frames * per_boost, 0.0) / per_frame))
    cache.set_params(w0)
    img = buffers()
    stream.synchronize()
    e0.record(stream)
    nb = min(boost_frames, total)
    nrc_frames(nb, img, 5, 0, boost)
    nrc_frames(total - nb, img, 5, nb)
    e1.record(stream)
    stream.synchronize()
    used = e0.elapsed_ms(e1)
    st = batch.nrc_last_frame()
    nrc = rel(img, ref)
    cache.close()
    batch.close()
    for sc in scenes:
        sc.close()
    kinds = ["heightfield" if (c5 or k % 2 == 0) else "box" for k in range(n_scenes)]
    mean = lambda v, kind=None: float(np.mean([x for x, t in zip(v, kinds) if kind in (None, t)]))
    workload = (f"{n_scenes} heightfield+icosphere scenes of the batched workload (triangles log-uniform in "
                "[1e4, 5e5], resolutions uniform in {128^2, 256^2, 512^2, 1024^2}, seed 6)" if c5 else
                f"{n_scenes} scenes (heightfield+icosphere and closed boxes, alternating), resolutions uniform in "
                "{256^2, 512^2}")
    out = {
        "workload": f"{workload}, max depth {depth}, RR from depth 4, four hero wavelengths, no sky; one "
                    "width-32 FP16 radiance-cache network per scene",
        "pixels_per_frame": int(pix),
        "equal_time_ms": round(budget, 1),
        "plain_spp": plain_spp,
        "plain_ms_per_frame": round(budget / plain_spp, 3),
        "nrc_frames": total,
        "nrc_ms": round(used, 1),
        "nrc_ms_per_frame_calibrated": round(per_frame, 3),
        "nrc_ms_per_boosted_frame_calibrated": round(per_boost, 3),
        "nrc_boost": {"frames": boost_frames, "steps": boost_steps},
        "queries_last_frame": int(sum(x.queries for x in st)),
        "samples_last_frame": int(sum(x.samples for x in st)),
        "relmse_plain_mean": round(mean(plain), 5),
        "relmse_nrc_mean": round(mean(nrc), 5),
        "relmse_ratio_plain_over_nrc": round(mean(plain) / mean(nrc), 3),
        "ratio_heightfield_scenes": round(mean(plain, "heightfield") / mean(nrc, "heightfield"), 3),
        "ratio_box_scenes": None if c5 else round(mean(plain, "box") / mean(nrc, "box"), 3),
        "per_scene_ratio_min_max": [round(min(a / b for a, b in zip(plain, nrc)), 3),
                                    round(max(a / b for a, b in zip(plain, nrc)), 3)],
    }
    log(f"batch nrc: {out}")
    return out


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=20)
    ap_.add_argument("--warmup", type=int, default=3)
    ap_.add_argument("--triangles", type=int, default=1_000_000)
    ap_.add_argument("--res", type=int, default=4096)
    ap_.add_argument("--cpu-seconds", type=float, default=20.0)
    ap_.add_argument("--no-cpu", action="store_true")
    ap_.add_argument("--render-spp", type=int, default=8, help="spp of the 1080p render measurement (0 skips it)")
    ap_.add_argument("--nrc-spp", type=int, default=64,
                     help="plain spp that sets the equal-time budget of the radiance-cache comparison (0 skips it)")
    ap_.add_argument("--nrc-ref-spp", type=int, default=2048, help="spp of the plain reference image")
    ap_.add_argument("--nrc-warmup", type=int, default=0, help="cache frames rendered before the averaged ones")
    ap_.add_argument("--nrc-passes", type=int, default=0, help="training-only passes of the cache after each frame")
    ap_.add_argument("--batch-spp", type=int, default=16, help="spp of the 48-scene batch measurement (0 skips it)")
    ap_.add_argument("--batch-nrc-scenes", type=int, default=16,
                     help="scenes of the per-scene radiance-cache batch comparison (0 skips it)")
    ap_.add_argument("--batch-nrc-spp", type=int, default=64,
                     help="plain spp that sets the equal-time budget of the per-scene cache batch comparison")
    ap_.add_argument("--batch-nrc-ref-spp", type=int, default=1024, help="spp of its plain reference images")
    ap_.add_argument("--batch-nrc-c5-spp", type=int, default=16,
                     help="plain spp of the per-scene cache comparison on the 48-scene batched workload (0 skips it)")
    ap_.add_argument("--batch-nrc-c5-ref-spp", type=int, default=512, help="spp of its plain reference images")
    ap_.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no autotuning)")
    ap_.add_argument("--c1", type=int, default=1, help="run the 256x256 closed-box workload on GPU and CPU (0 skips it)")
    ap_.add_argument("--sweep", action="store_true",
                     help="time each traversal launch shape (rays per wave x refill) and print the table")
    ap_.add_argument("--copy-probe", action="store_true",
                     help="also run the library's streaming copy kernel over 1 GiB (HBM counter scale reference)")
    ap_.add_argument("--render-only", choices=["plain_1", "plain_4", "guided_1", "guided_4"], default=None,
                     help="profiling workload: only the 1080p heightfield render in one mode (one warm-up frame, "
                          "then --render-spp frames), one JSON line of its own, no headline metric")
    args = ap_.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the traversal core has no multi-GPU path; "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import aperture as ap

    if ap.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    ap.check(ap.lib().ap_device_set(0))
    if args.render_only:
        print(json.dumps(render_only(ap, args.render_only, max(1, args.render_spp))), flush=True)
        return
    stream = ap.Stream()

    t0 = time.perf_counter()
    objs = ap.synth_triangle_soup(args.triangles, seed=2, edge_min=1e-3, edge_max=1e-2)
    scene = ap.SceneRenderer(objs)
    info = scene.info()
    log(f"scene: {args.triangles} triangles, {info.node_count} nodes, depth {info.max_depth}, "
        f"stack {info.stack_capacity}, {info.device_bytes / 2**20:.1f} MiB on device ({time.perf_counter() - t0:.1f} s)")

    dev = ap.DeviceBuffer.from_numpy

    prim = ap.synth_camera_rays(args.res, args.res, [0.5, 0.5, -1.5], [0, 0, 1], 40.0)
    n_prim = len(prim)
    d_prim = dev(prim)
    d_phit = ap.DeviceBuffer(n_prim * 8)
    scene.trace_closest(d_prim, d_phit, n=n_prim, stream=stream)
    stream.synchronize()
    phit = d_phit.to_numpy(ap.HIT_DTYPE, n_prim)
    sec = scene.secondary_rays(prim, phit, seed=3, offset=1e-4)
    n_sec = len(sec)
    d_sec = dev(sec)
    d_shit = ap.DeviceBuffer(max(n_sec, 1) * 8)
    log(f"rays: {n_prim} primaries ({(phit['id'] >= 0).mean():.3f} hit), {n_sec} secondaries")

    # Algorithmic bytes of one step, from the kernel's own visit counts.
    d_stats = ap.DeviceBuffer(max(n_prim, n_sec) * 8)
    algo = {}
    for name, d_r, d_h, n in (("primary", d_prim, d_phit, n_prim), ("secondary", d_sec, d_shit, n_sec)):
        scene.trace_closest(d_r, d_h, n=n, stream=stream, stats=d_stats)
        stream.synchronize()
        st = d_stats.to_numpy(ap.STATS_DTYPE, n)
        nodes = int(st["nodes"].astype(np.int64).sum())
        prims = int(st["prims"].astype(np.int64).sum())
        algo[name] = {"bytes": 40 * n + 32 * nodes + 36 * prims, "nodes_per_ray": nodes / n,
                      "tests_per_ray": prims / n}
    log(f"visit counts: {algo}")

    # Launch shape per ray set, picked by the library's autotuner (untimed).
    shapes = {}
    if args.no_tune:
        shapes = {"primary": scene.get_launch(), "secondary": scene.get_launch()}
    else:
        for name, d_r, d_h, n in (("primary", d_prim, d_phit, n_prim), ("secondary", d_sec, d_shit, n_sec)):
            wr, rf, ms = scene.autotune(d_r, d_h, n=n, stream=stream)
            shapes[name] = (wr, rf)
            log(f"autotune {name}: {wr} rays per wave, refill at {rf} idle lanes ({ms:.3f} ms)")

    def step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        scene.set_launch(*shapes["primary"])
        scene.trace_closest(d_prim, d_phit, n=n_prim, stream=stream)
        if evs is not None:
            evs[1].record(stream)
        scene.set_launch(*shapes["secondary"])
        scene.trace_closest(d_sec, d_shit, n=n_sec, stream=stream)
        if evs is not None:
            evs[2].record(stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()

    if args.sweep:
        ev0, ev1 = ap.Event(), ap.Event()
        for wr in (64, 128, 256, 512, 1024, 2048, 4096):
            for rf in (1, 4, 8, 16, 32, 64):
                scene.set_launch(wr, rf)
                res = []
                for d_r, d_h, n in ((d_prim, d_phit, n_prim), (d_sec, d_shit, n_sec)):
                    scene.trace_closest(d_r, d_h, n=n, stream=stream)
                    ev0.record(stream)
                    for _ in range(3):
                        scene.trace_closest(d_r, d_h, n=n, stream=stream)
                    ev1.record(stream)
                    stream.synchronize()
                    res.append(n * 3 / ev0.elapsed_ms(ev1) / 1e3)
                log(f"sweep wave_rays {wr:5d} refill {rf:2d}: primary {res[0]:7.1f} secondary {res[1]:7.1f} Mrays/s")
        scene.set_launch(512, 16)

    if args.copy_probe:
        nbytes = 1 << 30
        src, dst = ap.DeviceBuffer(nbytes), ap.DeviceBuffer(nbytes)
        src.zero(stream)
        for _ in range(2):
            dst.copy_from(src, nbytes, stream)
        stream.synchronize()
        log(f"copy probe: 2 x {nbytes} bytes through ap_device_copy")
        src.free()
        dst.free()

    evs = [[ap.Event() for _ in range(3)] for _ in range(args.steps)]
    start, end = ap.Event(), ap.Event()
    ap.synchronize()
    start.record(stream)
    for k in range(args.steps):
        step(evs[k])
    end.record(stream)
    stream.synchronize()
    ap.synchronize()
    total_ms = start.elapsed_ms(end)
    t_prim = sum(e[0].elapsed_ms(e[1]) for e in evs) / args.steps
    t_sec = sum(e[1].elapsed_ms(e[2]) for e in evs) / args.steps
    ms_per_step = total_ms / args.steps
    rays_per_step = n_prim + n_sec
    value = rays_per_step / (ms_per_step / 1e3) / 1e6
    bytes_step = algo["primary"]["bytes"] + algo["secondary"]["bytes"]
    kernel_ms = t_prim + t_sec
    achieved = bytes_step / (kernel_ms / 1e3) / 1e9
    prof = read_profile()
    traffic = prof.get("traffic_bytes_per_launch")
    log(f"primary {t_prim:.3f} ms ({n_prim / t_prim / 1e3:.0f} Mrays/s), secondary {t_sec:.3f} ms "
        f"({n_sec / t_sec / 1e3:.0f} Mrays/s)")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"bvh_closest_hit: {args.triangles}-triangle soup, binned-SAH BVH (leaf 8, 12 bins), "
                        f"{args.res}x{args.res} primaries + {n_sec} cosine secondaries per step",
            "triangles": args.triangles,
            "rays_per_step": rays_per_step,
            "primary_mrays_s": round(n_prim / t_prim / 1e3, 1),
            "secondary_mrays_s": round(n_sec / t_sec / 1e3, 1),
            "traversal_mode": "culled",
            "launch": {k: {"wave_rays": v[0], "refill": v[1]} for k, v in shapes.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_step / 2,
            "avg_launch_ms": round(kernel_ms / 2, 4),
            "nodes_per_ray": {k: round(v["nodes_per_ray"], 2) for k, v in algo.items()},
            "tests_per_ray": {k: round(v["tests_per_ray"], 2) for k, v in algo.items()},
            "note": "algorithmic bytes over HBM peak; the records (69 MB) are served mostly from L2 and MALL, so "
                    "frac can exceed 1: the kernel is bound by VALU issue (valu_issue, from the committed counters)",
            "valu_issue": {k: prof.get(f"valu_issue_{k}") for k in ("primary", "secondary")},
            "lane_use": {k: prof.get(f"lane_use_{k}") for k in ("primary", "secondary")},
        },
        "cpu_baseline": None,
    }
    threads = min(os.cpu_count() or 1, 16)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(objs, prim, sec, args.cpu_seconds, threads)
    scene.close()
    del objs, prim, sec
    if args.c1:
        line["c1"] = c1_bench(ap, threads, cpu=not args.no_cpu)
    if args.render_spp > 0:
        line["render"] = render_bench(ap, args.render_spp, 0.0 if args.no_cpu else 10.0, threads,
                                      nrc_spp=args.nrc_spp, nrc_ref_spp=args.nrc_ref_spp,
                                      nrc_warmup=args.nrc_warmup, nrc_passes=args.nrc_passes)
    if args.batch_spp > 0:
        line["batch"] = batch_bench(ap, args.batch_spp, 12)
    if args.batch_nrc_scenes > 0:
        line["batch_nrc"] = batch_nrc_bench(ap, args.batch_nrc_scenes, args.batch_nrc_spp, args.batch_nrc_ref_spp)
    if args.batch_nrc_c5_spp > 0:
        line["batch_nrc_c5"] = batch_nrc_bench(ap, 48, args.batch_nrc_c5_spp, args.batch_nrc_c5_ref_spp, c5=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
