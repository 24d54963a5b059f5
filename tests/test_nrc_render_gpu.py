"""GPU checks of rendering with the learned radiance cache (ap_render_nrc)
against the oracle's restatement (CalculateRadianceNRC, oracle/or_render.cpp)
with the cache frozen (train_steps = 0).

Bars:
  * paths, termination decisions and query records follow the same FP32
    operations on both sides: the number of paths ending into the cache, the
    training-sample keys (pixel, vertex) and the 13 query floats of every
    training sample are bit-identical, and so is every pixel whose path never
    reached the cache;
  * the cache's prediction is the MFMA network (FP32 accumulation) on the GPU
    and the FP64 network with the same BF16 rounding points in the oracle, so
    pixels and targets that depend on it agree within 1e-4 relative for
    99.5% and 2e-2 for all (a BF16 rounding boundary can fall between the
    two accumulations);
  * a cache that is never reached leaves ap_render's image bit for bit;
  * training is deterministic: two runs from the same weights give the same
    image and the same weights, and the weights move;
  * with directional guiding (probes built by a guided frame, bit-identical on
    both sides), a frozen-cache guided frame follows the same paths: query
    counts, training keys and query records bit-exact, pixels within the
    bar above; the probes rebuilt from its guiding samples (which take the
    cache prediction as a path's endpoint radiance) have the same count and
    index volume, and texels within 1e-3.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import test_nrc_cpu as ncpu
import test_render_gpu as rg

pytestmark = pytest.mark.gpu


def render_nrc(scene, cam, prm, cache, nprm):
    n = cam.width * cam.height
    d_hdr = ap.DeviceBuffer(n * 12)
    d_hdr.zero()
    d_sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
    scene.render_nrc(cam, prm, cache, nprm, d_hdr, sum_xyz=d_sums)
    ap.synchronize()
    shp = (cam.height, cam.width, 3)
    return d_hdr.to_numpy(np.float32, n * 3).reshape(shp), d_sums.to_numpy(np.float64, n * 3).reshape(shp)


def frozen_parity(objs, cam, prm, label, train_paths=256, unbiased_every=4, spread=0.01, weights_seed=41):
    print(f"[nrc-render] {label}: {len(objs)} objects, {cam.width}x{cam.height}, spp {prm.spp}", flush=True)
    scene = ap.SceneRenderer(objs)
    margin = np.float32(scene.info().line_margin)
    cache = ap.RadianceCache(seed=1)
    w = ncpu.init_params(weights_seed, scale=0.7)
    cache.set_params(w)
    nprm = ap.NrcRenderParams.defaults(train_paths=train_paths, train_steps=0, unbiased_every=unbiased_every,
                                       spread=spread)
    g_hdr, g_sums = render_nrc(scene, cam, prm, cache, nprm)
    st, g_s = scene.nrc_last_frame()
    lo, hi = cache.bounds()
    o_hdr, o_sums, o_s, o_nq = ol.OracleBVH(objs).render_nrc(cam, prm, w, lo, hi, spread=spread,
                                                             train_paths=train_paths, unbiased_every=unbiased_every,
                                                             cull_margin=margin)
    rel = np.abs(g_sums - o_sums) / (np.abs(o_sums) + 1e-6)
    ident = float(np.all(g_sums == o_sums, axis=-1).mean())
    print(f"[nrc-render] {label}: queries {st.queries} vs {o_nq}, samples {st.samples} vs {len(o_s)}, "
          f"pixels bit-identical {ident:.4f}, rel err p99.5 {np.quantile(rel, 0.995):.3g} max {rel.max():.3g}",
          flush=True)
    assert st.queries == o_nq and st.queries > 0
    assert st.samples == len(o_s) and st.samples > 0
    assert np.array_equal(g_s[:, 19].view(np.uint32), o_s[:, 19].view(np.uint32))
    assert np.array_equal(g_s[:, :16], o_s[:, :16])
    assert np.quantile(rel, 0.995) < 1e-4 and rel.max() < 2e-2
    t = np.abs(g_s[:, 16:19] - o_s[:, 16:19]) / (np.abs(o_s[:, 16:19]) + 1e-4)
    assert np.quantile(t, 0.995) < 1e-4 and t.max() < 2e-2
    return scene, cache


def test_frozen_cache_box_matches_oracle(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 48, 48)
    prm = ap.RenderParams.defaults(spp=2, max_depth=12, seed=7)
    frozen_parity(objs, cam, prm, "box")


def test_frozen_cache_heightfield_matches_oracle(gpu):
    objs = ap.synth_heightfield_scene(24, 3, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 48, 32)
    prm = ap.RenderParams.defaults(spp=1, max_depth=16, rr_min_depth=4, seed=11)
    frozen_parity(objs, cam, prm, "heightfield", train_paths=512, unbiased_every=16)


def test_unreached_cache_is_ap_render(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 32, 32)
    prm = ap.RenderParams.defaults(spp=3, max_depth=8, seed=2)
    scene = ap.SceneRenderer(objs)
    ref = rg.render_gpu(scene, cam, prm)
    cache = ap.RadianceCache(seed=1)
    hdr, sums = render_nrc(scene, cam, prm, cache, ap.NrcRenderParams.defaults(spread=1e30, train_paths=64,
                                                                                 train_steps=0))
    assert scene.nrc_last_frame(samples=False)[0].queries == 0
    assert np.array_equal(hdr, ref["hdr"]) and np.array_equal(sums, ref["sums"])


def test_training_is_deterministic_and_moves_the_weights(gpu):
    objs = ap.synth_heightfield_scene(32, 3, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 64)
    prm = ap.RenderParams.defaults(spp=6, max_depth=16, rr_min_depth=4, seed=3)
    nprm = ap.NrcRenderParams.defaults(train_paths=1024, batch=1024, train_steps=2)
    out = []
    for _ in range(2):
        scene = ap.SceneRenderer(objs)
        cache = ap.RadianceCache(seed=9)
        w0 = cache.params()
        hdr, _ = render_nrc(scene, cam, prm, cache, nprm)
        st, s = scene.nrc_last_frame()
        out.append((hdr, cache.params(), w0, st.samples))
    (h1, w1, w0, n1), (h2, w2, _, n2) = out
    print(f"[nrc-render] training: samples/frame {n1}, weight change {np.abs(w1 - w0).max():.3g}", flush=True)
    assert n1 == n2 and n1 > 0
    assert np.array_equal(h1, h2) and np.array_equal(w1, w2)
    assert np.abs(w1 - w0).max() > 0.0
    assert np.isfinite(h1).all() and np.isfinite(w1).all()


def test_frozen_cache_four_wavelengths_matches_oracle(gpu):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 40, 40)
    prm = ap.RenderParams.defaults(spp=2, max_depth=12, seed=8, wavelengths=4)
    frozen_parity(objs, cam, prm, "box, four wavelengths per path")


def test_training_passes_touch_only_the_cache(gpu):
    """Training-only passes trace and train but add nothing to the image:
    with the cache frozen they leave the image bit for bit, and with
    training on they move the weights further than the frames alone."""
    objs = ap.synth_heightfield_scene(32, 3, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 48)
    prm = ap.RenderParams.defaults(spp=3, max_depth=16, rr_min_depth=4, seed=4, wavelengths=4)
    images, moves = [], []
    for passes, steps in ((0, 0), (3, 0), (0, 2), (3, 2)):
        scene = ap.SceneRenderer(objs)
        cache = ap.RadianceCache(seed=2)
        w0 = cache.params()
        hdr, _ = render_nrc(scene, cam, prm, cache, ap.NrcRenderParams.defaults(
            train_paths=512, batch=512, train_steps=steps, train_passes=passes))
        images.append(hdr)
        moves.append(float(np.abs(cache.params() - w0).sum()))
        scene.close()
        cache.close()
    print(f"[nrc-render] training passes: weight change {moves}", flush=True)
    assert np.array_equal(images[0], images[1])
    assert moves[0] == 0.0 and moves[1] == 0.0 and moves[3] > moves[2] > 0.0


def srgb_rel_mse(a, b):
    m = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    a = np.maximum(a.reshape(-1, 3).astype(np.float64) @ m.T, 0.0)
    b = np.maximum(b.reshape(-1, 3).astype(np.float64) @ m.T, 0.0)
    return float((((a - b) ** 2) / (b ** 2 + 1e-2)).mean())


def test_trained_cache_gains_over_the_plain_tracer(gpu):
    """What the trained network gains: in the closed box (light bouncing
    between coloured walls), after the cache has trained on 128 frames, 32
    frames that end paths into it have far lower relMSE against a 4096-spp
    plain reference than 32 plain frames - and an untrained cache (random
    weights, the same renderer) is far worse than both, so the gain is the
    training's."""
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 96, 96)
    scene = ap.SceneRenderer(objs)
    P = dict(max_depth=8, rr_min_depth=3, wavelengths=4)
    ref = rg.render_gpu(scene, cam, ap.RenderParams.defaults(spp=4096, seed=1000, **P))["hdr"]
    plain = rg.render_gpu(scene, cam, ap.RenderParams.defaults(spp=32, seed=5, **P))["hdr"]
    frozen = ap.NrcRenderParams.defaults(train_paths=0, train_steps=0)
    untrained = ap.RadianceCache(seed=1)
    raw, _ = render_nrc(scene, cam, ap.RenderParams.defaults(spp=32, seed=5, **P), untrained, frozen)
    cache = ap.RadianceCache(seed=1)
    render_nrc(scene, cam, ap.RenderParams.defaults(spp=128, seed=77, **P), cache,
               ap.NrcRenderParams.defaults(train_paths=2048, batch=4096))
    nrc, _ = render_nrc(scene, cam, ap.RenderParams.defaults(spp=32, seed=5, **P), cache, frozen)
    e_plain, e_nrc, e_raw = (srgb_rel_mse(x, ref) for x in (plain, nrc, raw))
    print(f"[nrc-render] box 96x96, 32 spp: relMSE plain {e_plain:.4g}, trained cache {e_nrc:.4g} "
          f"({e_plain / e_nrc:.2f}x lower), untrained cache {e_raw:.4g}", flush=True)
    assert e_nrc < 0.5 * e_plain
    assert e_raw > 2.0 * e_nrc


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_guided_frozen_cache_matches_oracle(gpu, wavelengths):
    objs = ap.synth_box_scene(1, 2, 15.0)
    cam = ap.Camera.make([0, 0, 3.4], [0, 0, -1], 40.0, 40, 40)
    scene = ap.SceneRenderer(objs)
    margin = np.float32(scene.info().line_margin)
    bvh = ol.OracleBVH(objs)
    guide = ol.Guide()
    kw = dict(max_depth=10, rr_min_depth=3, seed=13, wavelengths=wavelengths, use_guiding=1)
    n = cam.width * cam.height
    # A guided frame without the cache builds the probes on both sides.
    prm0 = ap.RenderParams.defaults(spp=1, frame_index=48, **kw)
    hdr0 = ap.DeviceBuffer(n * 12)
    hdr0.zero()
    scene.render(cam, prm0, hdr0)
    ap.synchronize()
    bvh.render(cam, prm0, cull_margin=margin, guide=guide)
    assert np.array_equal(scene.guiding_export()[0].view(np.uint32), guide.export()[0].view(np.uint32))
    # A guided frame with the frozen cache.
    cache = ap.RadianceCache(seed=1)
    w = ncpu.init_params(47, scale=0.7)
    cache.set_params(w)
    nprm = ap.NrcRenderParams.defaults(train_paths=256, train_steps=0, unbiased_every=4, spread=0.01)
    prm1 = ap.RenderParams.defaults(spp=1, frame_index=49, **kw)
    g_hdr, g_sums = render_nrc(scene, cam, prm1, cache, nprm)
    st, g_s = scene.nrc_last_frame()
    lo, hi = cache.bounds()
    o_hdr, o_sums, o_s, o_nq = bvh.render_nrc(cam, prm1, w, lo, hi, spread=nprm.spread, train_paths=nprm.train_paths,
                                              unbiased_every=nprm.unbiased_every, cull_margin=margin, guide=guide)
    rel = np.abs(g_sums - o_sums) / (np.abs(o_sums) + 1e-6)
    gd, gc, gv, gb = scene.guiding_export()
    od, oc, ov, ob = guide.export()
    trel = np.abs(gd - od) / (np.abs(od) + 1e-6)
    print(f"[nrc-render] guided, {wavelengths} wavelengths: queries {st.queries} vs {o_nq}, samples {st.samples} vs "
          f"{len(o_s)}, pixel rel p99.5 {np.quantile(rel, 0.995):.3g} max {rel.max():.3g}; probes {len(gc)} vs "
          f"{len(oc)}, texel rel p99.9 {np.quantile(trel, 0.999):.3g}", flush=True)
    assert st.queries == o_nq and st.queries > 0
    assert st.samples == len(o_s) and st.samples > 0
    assert np.array_equal(g_s[:, 19].view(np.uint32), o_s[:, 19].view(np.uint32))
    assert np.array_equal(g_s[:, :13], o_s[:, :13])
    assert np.quantile(rel, 0.995) < 1e-4 and rel.max() < 2e-2
    assert np.array_equal(gb, ob) and np.array_equal(gv, ov) and len(gc) == len(oc)
    assert np.quantile(trel, 0.999) < 1e-3
    cache.close()
    scene.close()


def test_frames_of_one_call_equal_frames_one_call_each(gpu):
    """The frames of one ap_render_nrc call run two deep (the next frame's
    camera rays and closest hits beside this frame's last bounces, the
    cache's training beside the next frame's bounces, each frame set with
    its own path state and counters); one frame per call runs none of that
    overlap. Both give the same images, FP64 sums, trained weights and
    last-frame samples, with and without guiding."""
    objs = ap.synth_heightfield_scene(32, 3, 5, 20.0)
    cam = ap.Camera.make([0.0, 0.8, 1.4], [0.0, -0.8, -1.4], 45.0, 64, 48)
    n = cam.width * cam.height
    nprm = ap.NrcRenderParams.defaults(train_paths=1024, batch=1024, train_steps=2)
    frames = 5
    for guided in (0, 1):
        kw = dict(max_depth=16, rr_min_depth=4, seed=3, wavelengths=4, use_guiding=guided)
        out = []
        for one_call_each in (False, True):
            scene = ap.SceneRenderer(objs)
            cache = ap.RadianceCache(seed=9)
            d_hdr = ap.DeviceBuffer(n * 12)
            d_hdr.zero()
            d_sums = ap.DeviceBuffer.from_numpy(np.zeros(n * 3, np.float64))
            if one_call_each:
                for f in range(frames):
                    scene.render_nrc(cam, ap.RenderParams.defaults(spp=1, frame_index=f, **kw), cache, nprm, d_hdr,
                                     sum_xyz=d_sums)
            else:
                scene.render_nrc(cam, ap.RenderParams.defaults(spp=frames, **kw), cache, nprm, d_hdr, sum_xyz=d_sums)
            ap.synchronize()
            st, smp = scene.nrc_last_frame()
            out.append((d_hdr.to_numpy(np.float32, n * 3), d_sums.to_numpy(np.float64, n * 3), cache.params(),
                        st.queries, st.samples, smp))
            cache.close()
            scene.close()
        (h1, s1, w1, q1, n1, p1), (h2, s2, w2, q2, n2, p2) = out
        print(f"[nrc-render] pipelined frames, guided {guided}: queries {q1} / {q2}, samples {n1} / {n2}", flush=True)
        assert q1 == q2 and n1 == n2 and n1 > 0
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
        assert np.array_equal(h1, h2) and np.array_equal(s1, s2)
        assert np.array_equal(w1, w2)
