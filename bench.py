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


def read_traffic():
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/), if committed."""
    path = os.path.join(REPO, "profiles", "trace_closest_c2_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(objs, rays_a, rays_b, seconds, threads):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol

    bvh = ol.OracleBVH(objs)
    g = np.random.default_rng(7)
    pool = np.concatenate([rays_a[g.choice(len(rays_a), 4096, replace=False)],
                           rays_b[g.choice(len(rays_b), 4096, replace=False)]])
    g.shuffle(pool)
    probe = pool[: max(threads, 16)]
    t0 = time.perf_counter()
    bvh.trace(probe, threads=threads)
    dt = time.perf_counter() - t0
    n = int(min(len(pool), max(len(probe), len(probe) * seconds / max(dt, 1e-6))))
    sample = pool[:n]
    t0 = time.perf_counter()
    bvh.trace(sample, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": n / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} rays drawn evenly from the primary and secondary sets, oracle restatement of the "
                  f"reference TraverseBVH (its own slab test, no line culling), {dt:.1f} s",
    }


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=20)
    ap_.add_argument("--warmup", type=int, default=3)
    ap_.add_argument("--triangles", type=int, default=1_000_000)
    ap_.add_argument("--res", type=int, default=4096)
    ap_.add_argument("--cpu-seconds", type=float, default=20.0)
    ap_.add_argument("--no-cpu", action="store_true")
    args = ap_.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the traversal core has no multi-GPU path in this build "
              "(see DESIGN.md for how it would shard); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import aperture as ap

    if ap.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    ap.check(ap.lib().ap_device_set(0))
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

    def step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        scene.trace_closest(d_prim, d_phit, n=n_prim, stream=stream)
        if evs is not None:
            evs[1].record(stream)
        scene.trace_closest(d_sec, d_shit, n=n_sec, stream=stream)
        if evs is not None:
            evs[2].record(stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()

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
    traffic = read_traffic()
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
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        line["cpu_baseline"] = cpu_baseline(objs, prim, sec, args.cpu_seconds, threads)
    print(json.dumps(line), flush=True)
    scene.close()


if __name__ == "__main__":
    main()
