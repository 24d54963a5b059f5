"""Writes integrator_counts_heightfield.json: the oracle's visit accounting
(oracle/or_count.h) of one frame of the 1080p heightfield render workload
(1024^2-cell heightfield, 2,097,152 triangles, 256 icospheres, 4 thin-cube
lights, max depth 16, roulette from depth 4, no sky), traversed in the
core's CULLED order, in two views of its camera:

- "crop64": the 64x64 centre crop (vertical FOV scaled by 64/1080), and
- "view128x72": the whole field of view at 128x72 pixels, whose per-ray and
  per-vertex averages stand for the 1920x1080 frame.

Each view is counted at one and four wavelengths, plain and guided. Guided
counts are of frame 49, rendered after frame 48 has built the probes from
the same view; plain counts are of frame 49 alone. The bench prices a frame
with these averages (node records, triangle tests, material and LUT
fetches, probe texels) and the GPU's own ray counts.

Run from the repository root with the library and the oracle built:
    python3 tests/golden/make_integrator_counts.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "aperturerendererv3-optix-geometry_amd"))

import aperture as ap  # noqa: E402
import oracle_lib as ol  # noqa: E402

ORIGIN, DIRECTION, FOV = [0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0
VIEWS = {
    "crop64": (FOV * 64 / 1080, 64, 64),
    "view128x72": (FOV, 128, 72),
}


def count_frame(bvh, margin, cam, guided, wavelengths, threads):
    kw = dict(spp=1, max_depth=16, rr_min_depth=4, wavelengths=wavelengths, use_guiding=1 if guided else 0)
    guide = ol.Guide() if guided else None
    if guided:
        bvh.render(cam, ap.RenderParams.defaults(frame_index=48, **kw), cull_margin=margin, threads=threads,
                   guide=guide)
    with ol.counting() as c:
        bvh.render(cam, ap.RenderParams.defaults(frame_index=49, **kw), cull_margin=margin, threads=threads,
                   guide=guide)
    return c.totals


def main():
    threads = os.cpu_count() or 1
    t0 = time.time()
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    bvh = ol.OracleBVH(objs)
    margin = bvh.line_margin()
    print(f"{len(objs)} objects, BVH in {time.time() - t0:.1f} s, line margin {float(margin):.9g}", flush=True)
    out = {
        "workload": "heightfield 1024^2 (2,097,152 tris) + 256 icospheres (81,920 tris) + 4 thin-cube lights, "
                    "no sky, max depth 16, roulette from depth 4, CULLED traversal",
        "camera": {"origin": ORIGIN, "direction": DIRECTION, "fov_deg_1080p": FOV},
        "frame": 49,
        "counts": list(ol.COUNT_NAMES),
        "views": {},
    }
    for name, (fov, w, h) in VIEWS.items():
        cam = ap.Camera.make(ORIGIN, DIRECTION, fov, w, h)
        view = {"width": w, "height": h, "fov_deg": fov}
        for guided in (False, True):
            for nl in (1, 4):
                key = f"{'guided' if guided else 'plain'}_{nl}"
                view[key] = count_frame(bvh, margin, cam, guided, nl, threads)
                print(name, key, view[key], flush=True)
        out["views"][name] = view
    with open(os.path.join(HERE, "integrator_counts_heightfield.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
