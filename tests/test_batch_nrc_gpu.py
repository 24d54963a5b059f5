"""GPU checks of batched rendering with per-scene radiance caches
(ap_batch_render_nrc): one network per scene in a grouped cache, every
scene's cache queries answered by one grouped inference and every network
trained in the same launches.

Bars:
  * with width-64 BF16 networks the batch is ap_render_nrc scene by scene:
    each scene's image and FP64 sums and each network's weights after
    training equal, bit for bit, what ap_render_nrc gives for the scene
    alone with a one-network cache starting from the same weights (which
    tests/test_nrc_render_gpu.py ties to the oracle), and the per-scene
    counters match;
  * width-32 FP16 networks: the same paths end into the caches (query and
    sample counts equal the BF16 run's), training is deterministic and moves
    every network;
  * trained per-scene FP16 caches lower each scene's relMSE against a
    4096-spp plain reference well below the plain batch's at equal spp;
  * frozen width-32 FP16 caches against the oracle's cache renderer with the
    same network (oracle_lib render_nrc width 32, FP16 rounding points), scene
    by scene: equal cache-query counts, and pixels within the frozen-cache bar
    of tests/test_nrc_render_gpu.py (1e-4 relative for 99.5%, 2e-2 for all:
    the MFMA and the FP64 network round at the same points but sum in other
    orders).
"""
import oracle_lib as ol
import numpy as np
import pytest

import aperture as ap
import test_batch_gpu as bg
import test_nrc_cpu as ncpu
import test_nrc_render_gpu as nr

pytestmark = pytest.mark.gpu


def batch_render_nrc(batch, cams, prm, cache, nprm, stream):
    hdrs, sums = [], []
    for c in cams:
        n = c.width * c.height
        h = ap.DeviceBuffer(n * 12)
        h.zero()
        hdrs.append(h)
        sums.append(ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64)))
    batch.render_nrc(prm, cache, nprm, hdrs, sums, stream=stream)
    stream.synchronize()
    out = []
    for c, h, s in zip(cams, hdrs, sums):
        n = c.width * c.height
        out.append((h.to_numpy(np.float32, n * 3).reshape(c.height, c.width, 3),
                    s.to_numpy(np.float64, n * 3).reshape(c.height, c.width, 3)))
    return out


def weights(n):
    return np.stack([ncpu.init_params(60 + k, scale=0.7) for k in range(n)])


@pytest.mark.parametrize("steps", [0, 2])
def test_batch_caches_equal_single_scene_caches(gpu, steps):
    print(f"[batch-nrc] width 64 BF16, train_steps {steps}", flush=True)
    objs, cams = bg.scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    prm = ap.RenderParams.defaults(spp=3, max_depth=10, rr_min_depth=3, seed=21)
    nprm = ap.NrcRenderParams.defaults(train_paths=256, batch=512, train_steps=steps, unbiased_every=4)
    w = weights(len(scenes))
    cache = ap.RadianceCache(seed=1, groups=len(scenes), width=64, precision="bf16")
    cache.set_params(w)
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    got = batch_render_nrc(batch, cams, prm, cache, nprm, stream)
    stats = batch.nrc_last_frame()
    after = cache.params()
    for k, (sc, cam) in enumerate(zip(scenes, cams)):
        single = ap.RadianceCache(seed=1)
        single.set_params(w[k])
        hdr, sums = nr.render_nrc(sc, cam, prm, single, nprm)
        st, _ = sc.nrc_last_frame(samples=False)
        ident = float(np.all(got[k][0] == hdr, axis=-1).mean())
        print(f"[batch-nrc]   scene {k} {cam.width}x{cam.height}: queries {stats[k].queries} vs {st.queries}, "
              f"samples {stats[k].samples} vs {st.samples}, pixels bit-identical {ident:.4f}", flush=True)
        assert (stats[k].queries, stats[k].train_paths, stats[k].samples) == (st.queries, st.train_paths, st.samples)
        assert np.array_equal(got[k][0], hdr) and np.array_equal(got[k][1], sums)
        assert np.array_equal(after[k], single.params())
        single.close()
    assert sum(s.queries for s in stats) > 0 and sum(s.samples for s in stats) > 0
    if steps:
        assert all(np.abs(after[k] - w[k]).max() > 0.0 for k in range(len(scenes)) if stats[k].samples > 0)
    batch.close()
    cache.close()


@pytest.mark.parametrize("nl", [1, 4])
def test_guided_batch_caches_equal_single_scene_caches(gpu, nl):
    """Guided batches with per-scene caches: each scene samples its own
    probes, a path that ends into its scene's cache gives the probes the
    prediction as its endpoint radiance, and each scene's probes are rebuilt
    from its own samples after every frame. Images, FP64 sums, trained
    weights, per-scene counters and probes equal ap_render_nrc of each scene
    alone, guided, with its own one-network cache."""
    print(f"[batch-nrc] guided, width 64 BF16, {nl} wavelength(s)", flush=True)
    objs, cams = bg.scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    prm = ap.RenderParams.defaults(spp=3, max_depth=10, rr_min_depth=3, seed=23, use_guiding=1, wavelengths=nl)
    nprm = ap.NrcRenderParams.defaults(train_paths=256, batch=512, train_steps=2, unbiased_every=4)
    w = weights(len(scenes))
    cache = ap.RadianceCache(seed=1, groups=len(scenes), width=64, precision="bf16")
    cache.set_params(w)
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    got = batch_render_nrc(batch, cams, prm, cache, nprm, stream)
    stats = batch.nrc_last_frame()
    after = cache.params()
    probes = [sc.guiding_export() for sc in scenes]
    for sc in scenes:
        sc.guiding_reset()
    for k, (sc, cam) in enumerate(zip(scenes, cams)):
        single = ap.RadianceCache(seed=1)
        single.set_params(w[k])
        hdr, sums = nr.render_nrc(sc, cam, prm, single, nprm)
        st, _ = sc.nrc_last_frame(samples=False)
        gdata, gcount, gvol, gbounds = sc.guiding_export()
        bdata, bcount, bvol, bbounds = probes[k]
        print(f"[batch-nrc]   scene {k} {cam.width}x{cam.height}: queries {stats[k].queries} vs {st.queries}, "
              f"probes {len(bcount)} vs {len(gcount)}", flush=True)
        assert (stats[k].queries, stats[k].train_paths, stats[k].samples) == (st.queries, st.train_paths, st.samples)
        assert np.array_equal(got[k][0], hdr) and np.array_equal(got[k][1], sums)
        assert np.array_equal(after[k], single.params())
        assert np.array_equal(bbounds, gbounds) and np.array_equal(bvol, gvol) and np.array_equal(bcount, gcount)
        assert np.array_equal(bdata.view(np.uint32), gdata.view(np.uint32))
        single.close()
    assert sum(s.queries for s in stats) > 0 and sum(len(p[1]) for p in probes) > 0
    batch.close()
    cache.close()


def test_fp16_width32_caches_train_deterministically(gpu):
    objs, cams = bg.scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    prm = ap.RenderParams.defaults(spp=4, max_depth=10, rr_min_depth=3, seed=22, wavelengths=4)
    nprm = ap.NrcRenderParams.defaults(train_paths=512, batch=512, train_steps=2)
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    runs = []
    for width, precision in ((32, "fp16"), (32, "fp16"), (64, "bf16")):
        cache = ap.RadianceCache(seed=3, groups=len(scenes), width=width, precision=precision)
        w0 = cache.params()
        out = batch_render_nrc(batch, cams, prm, cache, nprm, stream)
        runs.append((out, cache.params(), w0, [(s.queries, s.samples) for s in batch.nrc_last_frame()]))
        cache.close()
    (o1, w1, w0, s1), (o2, w2, _, s2), (_, _, _, s3) = runs
    print(f"[batch-nrc] fp16 width 32: per-scene (queries, samples) {s1}; "
          f"weight change {[float(np.abs(w1[k] - w0[k]).max()) for k in range(len(scenes))]}", flush=True)
    assert s1 == s2 == s3
    for k in range(len(scenes)):
        assert np.array_equal(o1[k][0], o2[k][0]) and np.isfinite(o1[k][0]).all()
        if s1[k][1] > 0:
            assert np.abs(w1[k] - w0[k]).max() > 0.0
    assert np.array_equal(w1, w2)
    batch.close()


def test_replayed_cache_frames_equal_direct_frames(gpu):
    """A call of several frames runs its first frame directly and replays a
    captured frame for the rest (frame hash, training draws and Adam's bias
    corrections read on the device). Four frames in one call must equal four
    calls of one frame each (all direct), bit for bit: images, FP64 sums,
    trained weights; and a second cache created after the first is destroyed
    (possibly at its address) must not replay the first one's frame."""
    objs, cams = bg.scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    nprm = ap.NrcRenderParams.defaults(train_paths=512, batch=512, train_steps=3)
    stream = ap.Stream()
    outs = []
    batch = ap.Batch(scenes, cams)
    for calls in ((4,), (1, 1, 1, 1), (4,)):
        cache = ap.RadianceCache(seed=5, groups=len(scenes), width=32, precision="fp16")
        hdrs, sums = [], []
        for c in cams:
            n = c.width * c.height
            h = ap.DeviceBuffer(n * 12)
            h.zero()
            hdrs.append(h)
            sums.append(ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64)))
        frame = 0
        for spp in calls:
            prm = ap.RenderParams.defaults(spp=spp, frame_index=frame, max_depth=10, rr_min_depth=3, seed=31,
                                           wavelengths=4)
            batch.render_nrc(prm, cache, nprm, hdrs, sums, stream=stream)
            frame += spp
        stream.synchronize()
        imgs = [(h.to_numpy(np.float32, c.width * c.height * 3), s.to_numpy(np.float64, c.width * c.height * 3))
                for c, h, s in zip(cams, hdrs, sums)]
        outs.append((imgs, cache.params(), [(x.queries, x.samples) for x in batch.nrc_last_frame()]))
        print(f"[batch-nrc] calls {calls}: per-scene (queries, samples) {outs[-1][2]}", flush=True)
        cache.close()
    batch.close()
    for imgs, w, st in outs[1:]:
        assert st == outs[0][2]
        assert np.array_equal(w, outs[0][1])
        for (h0, s0), (h1, s1) in zip(outs[0][0], imgs):
            assert np.array_equal(h0, h1) and np.array_equal(s0, s1)
    assert np.isfinite(outs[0][1]).all()


def test_per_scene_fp16_caches_gain_over_the_plain_batch(gpu):
    objs = [ap.synth_box_scene(1, 2, 15.0), ap.synth_box_scene(4, 1, 10.0)]
    cams = [ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 96, 96), ap.Camera.make([0.2, 0.1, 3.0], [0, 0, -1], 50.0,
                                                                                   64, 64)]
    scenes = [ap.SceneRenderer(o) for o in objs]
    P = dict(max_depth=8, rr_min_depth=3, wavelengths=4)
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    refs = [x[0] for x in bg.batch_render(batch, cams, ap.RenderParams.defaults(spp=4096, seed=1000, **P),
                                           stream)[0]]
    plain = [x[0] for x in bg.batch_render(batch, cams, ap.RenderParams.defaults(spp=32, seed=5, **P), stream)[0]]
    cache = ap.RadianceCache(seed=1, groups=2, width=32, precision="fp16")
    batch_render_nrc(batch, cams, ap.RenderParams.defaults(spp=128, seed=77, **P), cache,
                     ap.NrcRenderParams.defaults(train_paths=2048, batch=4096), stream)
    nrc = [x[0] for x in batch_render_nrc(batch, cams, ap.RenderParams.defaults(spp=32, seed=5, **P), cache,
                                          ap.NrcRenderParams.defaults(train_paths=0, train_steps=0), stream)]
    for k in range(2):
        e_plain, e_nrc = nr.srgb_rel_mse(plain[k], refs[k]), nr.srgb_rel_mse(nrc[k], refs[k])
        print(f"[batch-nrc] scene {k}: relMSE plain {e_plain:.4g}, per-scene FP16 cache {e_nrc:.4g} "
              f"({e_plain / e_nrc:.2f}x lower)", flush=True)
        assert e_nrc < 0.5 * e_plain
    batch.close()
    cache.close()


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_frozen_fp16_caches_match_the_oracle(gpu, wavelengths):
    objs, cams = bg.scenes_and_cameras()
    scenes = [ap.SceneRenderer(o) for o in objs]
    prm = ap.RenderParams.defaults(spp=2, max_depth=10, rr_min_depth=3, seed=23, wavelengths=wavelengths)
    nprm = ap.NrcRenderParams.defaults(train_paths=256, batch=512, train_steps=0, unbiased_every=4)
    w = np.stack([ncpu.init_params_w(32, 70 + k) for k in range(len(scenes))])
    cache = ap.RadianceCache(seed=1, groups=len(scenes), width=32, precision="fp16")
    cache.set_params(w)
    batch = ap.Batch(scenes, cams)
    stream = ap.Stream()
    got = batch_render_nrc(batch, cams, prm, cache, nprm, stream)
    stats = batch.nrc_last_frame()
    for k, (o, sc, cam) in enumerate(zip(objs, scenes, cams)):
        # The scene's cache bounds, as ap_render_nrc and the batch set them.
        single = ap.RadianceCache(seed=1, width=32, precision="fp16")
        single.set_params(w[k])
        nr.render_nrc(sc, cam, ap.RenderParams.defaults(spp=1, max_depth=1), single, nprm)
        lo, hi = single.bounds()
        single.close()
        margin = np.float32(sc.info().line_margin)
        o_hdr, o_sums, _, o_nq = ol.OracleBVH(o).render_nrc(cam, prm, w[k], lo, hi, spread=nprm.spread,
                                                            train_paths=nprm.train_paths,
                                                            unbiased_every=nprm.unbiased_every, cull_margin=margin,
                                                            width=32, mode=2)
        rel = np.abs(got[k][1] - o_sums) / (np.abs(o_sums) + 1e-6)
        ident = float(np.all(got[k][1] == o_sums, axis=-1).mean())
        print(f"[batch-nrc] vs oracle, FP16 width 32, {wavelengths} wavelengths: scene {k}: queries "
              f"{stats[k].queries} vs {o_nq}, pixels bit-identical {ident:.4f}, rel p99.5 "
              f"{np.quantile(rel, 0.995):.3g} max {rel.max():.3g}", flush=True)
        assert stats[k].queries == o_nq
        assert np.quantile(rel, 0.995) < 1e-4 and rel.max() < 2e-2
    assert sum(s.queries for s in stats) > 0
    batch.close()
    cache.close()
