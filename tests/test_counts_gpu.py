"""The integrator's ray counts on the GPU equal the oracle's visit accounting
of the same frame (tests/golden/integrator_counts_heightfield.json, made by
tests/golden/make_integrator_counts.py): the 2.18 M-triangle heightfield seen
whole at 128x72, frame 49, plain and guided (after frame 48 has built the
probes), one and four wavelengths. The closest-hit and shadow rays the
core's launches trace are exactly the oracle's closest-hit and any-hit
traversals, so the fixture's per-ray node and triangle averages describe the
rays the GPU traces when the bench prices a frame in algorithmic bytes.
"""
import json
import os

import numpy as np
import pytest

import aperture as ap

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_ray_counts_equal_the_oracle_accounting(gpu):
    with open(os.path.join(HERE, "golden", "integrator_counts_heightfield.json")) as f:
        fx = json.load(f)
    view = fx["views"]["view128x72"]
    objs = ap.synth_heightfield_scene(1024, 256, 4, 20.0)
    scene = ap.SceneRenderer(objs)
    cam = ap.Camera.make(fx["camera"]["origin"], fx["camera"]["direction"], view["fov_deg"], view["width"],
                         view["height"])
    n = cam.width * cam.height
    for guided in (False, True):
        for nl in (1, 4):
            key = f"{'guided' if guided else 'plain'}_{nl}"
            print(f"[counts] {key}", flush=True)
            kw = dict(spp=1, max_depth=16, rr_min_depth=4, wavelengths=nl, use_guiding=1 if guided else 0)
            hdr = ap.DeviceBuffer(n * 12)
            hdr.zero()
            counts = ap.DeviceBuffer(16)
            if guided:
                scene.guiding_reset()
                scene.render(cam, ap.RenderParams.defaults(frame_index=48, **kw), hdr)
            counts.zero()
            scene.render(cam, ap.RenderParams.defaults(frame_index=49, **kw), hdr, ray_count=counts)
            ap.synchronize()
            ext, shadow = (int(x) for x in counts.to_numpy(np.uint64, 2))
            want = view[key]
            print(f"[counts] {key}: closest-hit {ext} (oracle {want['ext_rays']}), shadow {shadow} "
                  f"(oracle {want['shadow_rays']})", flush=True)
            assert ext == want["ext_rays"]
            assert shadow == want["shadow_rays"]
    scene.close()
