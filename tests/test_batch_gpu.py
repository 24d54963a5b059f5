"""GPU checks of batched multi-scene rendering (ap_batch_*).

A batch renders every scene's pixels through shared queues and one
traversal launch per bounce over the rays of all scenes, with a frame
captured as a HIP graph and replayed. Per-pixel random streams, traversal and
shading are those of ap_render, so each scene's running mean and FP64 sums
must equal, bit for bit, what ap_render gives for that scene alone - with the
graph replayed over several frames (each replay checked through the final
image and the sums of every frame), without the graph, with four wavelengths
per path, and after the autotuner has picked a launch shape. Calls issued
back to back on a non-blocking stream, with no host wait between them
(progressive frames, new framebuffers, an autotune), keep their order.
The batch is also compared directly with the oracle's CalculateRadiance per
scene (bit-identical images and FP64 sums); a guided batch equals each
scene's guided render alone (images, sums and probes); and one frame of the full
48-scene workload (ap.synth_batch_scenes: 7.4 M triangles, 13.7 M pixels)
equals the scene-by-scene renders for every scene, its smallest scene the
oracle too.
"""
import oracle_lib as ol
import numpy as np
import pytest

import aperture as ap
import test_render_gpu as rg

pytestmark = pytest.mark.gpu


def scenes_and_cameras():
    objs = [ap.synth_box_scene(1, 2, 15.0), ap.synth_heightfield_scene(48, 6, 4, 20.0),
            ap.synth_box_scene(4, 1, 10.0), ap.synth_heightfield_scene(20, 2, 9, 20.0)]
    cams = [ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 40, 32),
            ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 36),
            ap.Camera.make([0.2, 0.1, 3.0], [0, 0, -1], 50.0, 16, 16),
            ap.Camera.make([0.0, 0.6, 1.2], [0.0, -0.8, -1.4], 60.0, 33, 17)]
    return objs, cams


def batch_render(batch, cams, prm, stream):
    hdrs, sums = [], []
    for c in cams:
        n = c.width * c.height
        h = ap.DeviceBuffer(n * 12)
        h.zero()
        hdrs.append(h)
        sums.append(ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64)))
    rays = batch.render(prm, hdrs, sums, stream=stream)
    stream.synchronize()
    out = []
    for c, h, s in zip(cams, hdrs, sums):
        n = c.width * c.height
        out.append((h.to_numpy(np.float32, n * 3).reshape(c.height, c.width, 3),
                    s.to_numpy(np.float64, n * 3).reshape(c.height, c.width, 3)))
    return out, rays


def check_against_single(scenes, cams, prm, got, label):
    for k, (sc, cam) in enumerate(zip(scenes, cams)):
        ref = rg.render_gpu(sc, cam, prm)
        ident = float(np.all(got[k][0] == ref["hdr"], axis=-1).mean())
        print(f"[batch] {label}: scene {k} {cam.width}x{cam.height}: pixels bit-identical {ident:.4f}", flush=True)
        assert np.array_equal(got[k][0], ref["hdr"]), (label, k)
        assert np.array_equal(got[k][1], ref["sums"]), (label, k)


def test_batch_equals_single_scene_renders(gpu):
    objs, cams = scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    prm = ap.RenderParams.defaults(spp=3, max_depth=12, rr_min_depth=4, seed=6)
    got, rays = batch_render(batch, cams, prm, stream)
    print(f"[batch] graph: rays {rays}", flush=True)
    assert rays[0] > 0 and rays[1] > 0
    check_against_single(scenes, cams, prm, got, "graph, 3 frames")
    batch.set_graph(False)
    got, _ = batch_render(batch, cams, prm, stream)
    check_against_single(scenes, cams, prm, got, "direct launches")
    batch.set_graph(True)
    prm4 = ap.RenderParams.defaults(spp=2, frame_index=5, max_depth=12, rr_min_depth=4, seed=7, wavelengths=4)
    got, _ = batch_render(batch, cams, prm4, stream)
    check_against_single(scenes, cams, prm4, got, "four wavelengths, from frame 5")
    batch.close()


def test_autotuner_keeps_images(gpu):
    objs, cams = scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    prm = ap.RenderParams.defaults(spp=2, max_depth=12, rr_min_depth=4, seed=8)
    w, r, ms = batch.autotune(prm, frames=2, stream=stream)
    print(f"[batch] autotune: {w} rays per wave, refill {r}, {ms:.3f} ms per frame", flush=True)
    assert (w, r) == batch.get_launch() and w in (128, 256, 512) and r in (8, 16) and ms > 0.0
    got, _ = batch_render(batch, cams, prm, stream)
    check_against_single(scenes, cams, prm, got, "autotuned")
    with pytest.raises(ap.ApertureError):
        batch.set_launch(100, 8)
    # Guiding in groups of frames is refused in a batch (a guided batch
    # updates every scene's probes after every frame, tested below).
    with pytest.raises(ap.ApertureError) as e:
        batch.render(ap.RenderParams.defaults(use_guiding=1, guide_frames=2), [None] * len(cams), stream=stream)
    assert e.value.code == ap.ERR_UNSUPPORTED
    batch.close()


def _buffers(cams):
    hdrs, sums = [], []
    for c in cams:
        n = c.width * c.height
        h = ap.DeviceBuffer(n * 12)
        h.zero()
        hdrs.append(h)
        sums.append(ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64)))
    return hdrs, sums


def _read(cams, hdrs, sums):
    out = []
    for c, h, s in zip(cams, hdrs, sums):
        n = c.width * c.height
        out.append((h.to_numpy(np.float32, n * 3).reshape(c.height, c.width, 3),
                    s.to_numpy(np.float64, n * 3).reshape(c.height, c.width, 3)))
    return out


def test_back_to_back_calls_without_waits(gpu):
    """The frame counter is set on the caller's stream and the scene records
    are rewritten only after the previous call's frames are done: two
    progressive calls, a call into new framebuffers, an autotune over every
    launch shape (one graph each, none destroyed) and a last call, all queued
    with no host wait, equal the single-scene renders."""
    objs, cams = scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    kw = dict(max_depth=12, rr_min_depth=4, seed=11)
    a_h, a_s = _buffers(cams)
    b_h, b_s = _buffers(cams)
    c_h, c_s = _buffers(cams)
    ap.synchronize()
    batch.render(ap.RenderParams.defaults(spp=2, frame_index=0, **kw), a_h, a_s, stream=stream, count_rays=False)
    batch.render(ap.RenderParams.defaults(spp=2, frame_index=2, **kw), a_h, a_s, stream=stream, count_rays=False)
    batch.render(ap.RenderParams.defaults(spp=4, frame_index=0, **kw), b_h, b_s, stream=stream, count_rays=False)
    batch.autotune(ap.RenderParams.defaults(spp=1, **kw), frames=1, stream=stream)
    batch.render(ap.RenderParams.defaults(spp=4, frame_index=0, **kw), c_h, c_s, stream=stream, count_rays=False)
    stream.synchronize()
    prm = ap.RenderParams.defaults(spp=4, **kw)
    for label, (h, s) in (("two progressive calls", (a_h, a_s)), ("new framebuffers", (b_h, b_s)),
                          ("after an autotune", (c_h, c_s))):
        check_against_single(scenes, cams, prm, _read(cams, h, s), label)
    batch.close()


def oracle_render(objs, scene, cam, prm):
    margin = np.float32(scene.info().line_margin)
    hdr, sums, _ = ol.OracleBVH(objs).render(cam, prm, cull_margin=margin)
    return hdr, sums


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_batch_matches_the_oracle(gpu, wavelengths):
    objs, cams = scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    prm = ap.RenderParams.defaults(spp=2, max_depth=12, rr_min_depth=4, seed=13, wavelengths=wavelengths)
    got, _ = batch_render(batch, cams, prm, stream)
    for k, (o, sc, cam) in enumerate(zip(objs, scenes, cams)):
        hdr, sums = oracle_render(o, sc, cam, prm)
        ident = float(np.all(got[k][0] == hdr, axis=-1).mean())
        print(f"[batch] vs oracle, {wavelengths} wavelengths: scene {k} {cam.width}x{cam.height}: "
              f"pixels bit-identical {ident:.4f}", flush=True)
        assert np.array_equal(got[k][0], hdr) and np.array_equal(got[k][1], sums), k
    batch.close()


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_guided_batch_equals_guided_scenes(gpu, wavelengths):
    """Guiding in a batch: every scene's paths sample its own probes, each
    frame's samples go to its own slots and rebuild its own probes. Three
    calls of one frame each (the first from empty probes) must leave every
    scene's image, FP64 sums, probes, probe counts and index volume equal,
    bit for bit, to ap_render of the scene alone over the same frames (which
    tests/test_guiding_gpu.py ties to the oracle)."""
    guided_batch_vs_alone(wavelengths, [(48, 1), (49, 1), (50, 1)])


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_guided_batch_replayed_frames_equal_guided_scenes(gpu, wavelengths):
    """Calls of several guided frames: a call's first frame runs directly,
    the others replay one captured frame (render launches and every scene's
    probe build), including early frames whose builds must grow probe or
    volume storage (finished on the host after the replay, the captured
    frame then captured again); equal to each scene alone, probes included."""
    guided_batch_vs_alone(wavelengths, [(48, 3), (51, 4)])


def guided_batch_vs_alone(wavelengths, calls):
    objs, cams = scenes_and_cameras()
    stream = ap.Stream()
    kw = dict(max_depth=10, rr_min_depth=3, seed=12, wavelengths=wavelengths, use_guiding=1)
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    hdrs, sums = [], []
    for c in cams:
        n = c.width * c.height
        h = ap.DeviceBuffer(n * 12)
        h.zero()
        hdrs.append(h)
        sums.append(ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64)))
    for f0, spp in calls:
        batch.render(ap.RenderParams.defaults(frame_index=f0, spp=spp, **kw), hdrs, sums, stream=stream)
    stream.synchronize()
    first, last = calls[0][0], calls[-1][0] + calls[-1][1]
    for k, (o, c) in enumerate(zip(objs, cams)):
        n = c.width * c.height
        alone = ap.SceneRenderer(o)
        h = ap.DeviceBuffer(n * 12)
        h.zero()
        sm = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
        alone.render(c, ap.RenderParams.defaults(frame_index=first, spp=last - first, **kw), h, sum_xyz=sm)
        ap.synchronize()
        got = hdrs[k].to_numpy(np.float32, n * 3), sums[k].to_numpy(np.float64, n * 3)
        want = h.to_numpy(np.float32, n * 3), sm.to_numpy(np.float64, n * 3)
        gdata, gcount, gvol, gbounds = scenes[k].guiding_export()
        odata, ocount, ovol, obounds = alone.guiding_export()
        print(f"[batch] guided scene {k} {c.width}x{c.height}: {len(gcount)} probes (alone {len(ocount)}), "
              f"pixels bit-identical {float(np.mean(got[0] == want[0])):.4f}", flush=True)
        assert len(gcount) > 0
        assert np.array_equal(gbounds, obounds) and np.array_equal(gvol, ovol) and np.array_equal(gcount, ocount)
        assert np.array_equal(gdata.view(np.uint32), odata.view(np.uint32))
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        alone.close()
    batch.close()
    for sc in scenes:
        sc.close()


def test_full_size_batch_workload_one_frame(gpu):
    """One frame of the 48-scene workload as the bench renders it (depth 12,
    roulette from depth 4, four wavelengths, no sky), autotuned, through the
    captured graph."""
    objs, cams = ap.synth_batch_scenes(48, 6)
    print(f"[batch] full workload: {len(objs)} scenes, {sum(len(o) for o in objs)} objects, "
          f"{sum(c.width * c.height for c in cams)} pixels", flush=True)
    scenes = [ap.SceneRenderer(o) for o in objs]
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    P = dict(max_depth=12, rr_min_depth=4, sky=0, wavelengths=4, seed=6)
    w, r, ms = batch.autotune(ap.RenderParams.defaults(spp=1, **P), frames=1, stream=stream)
    print(f"[batch] full workload: autotuned {w} rays per wave, refill {r}, {ms:.1f} ms per frame", flush=True)
    prm = ap.RenderParams.defaults(spp=1, **P)
    got, rays = batch_render(batch, cams, prm, stream)
    print(f"[batch] full workload: rays {rays}", flush=True)
    check_against_single(scenes, cams, prm, got, "48-scene workload")
    k = int(np.argmin([c.width * c.height for c in cams]))
    hdr, sums = oracle_render(objs[k], scenes[k], cams[k], prm)
    print(f"[batch] full workload: scene {k} ({cams[k].width}x{cams[k].height}) against the oracle", flush=True)
    assert np.array_equal(got[k][0], hdr) and np.array_equal(got[k][1], sums)
    batch.close()
    for sc in scenes:
        sc.close()
