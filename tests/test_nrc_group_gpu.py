"""GPU checks of grouped radiance caches (ap_nrc_create_grouped and the
_grouped entry points): G independent networks of hidden width 32 or 64 in
BF16 or FP16, served by one launch per kernel, against the oracle's
width-generic FP64 network with the same rounding points
(or_nrc_forward_w / or_nrc_grad_w in oracle/or_nrc.cpp), one network at a
time.

Bars (the MFMA accumulates in FP32 in its own order, the oracle in FP64):
  * forward: relative error below 1e-2 for 99.5% of outputs, median below
    2e-4 (FP16 keeps 11 bits, BF16 8), for every network, with uneven and
    empty row segments and each network's own bounds;
  * gradient: per network, loss within 1e-3 relative, cosine similarity
    above 0.9999, relative L2 error below 1e-2, with the FP16 loss scale
    2^(ceil(log2(3 batch)) - 4) on both sides;
  * Adam on the GPU's own gradient equals the oracle's Adam bit for bit;
  * networks train independently: two networks fed different targets in
    one batch each fit their own.
"""
import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import test_nrc_cpu as cpu

pytestmark = pytest.mark.gpu

SHAPES = [(32, "fp16"), (64, "fp16"), (32, "bf16")]
MODE = {"bf16": ol.NRC_BF16, "fp16": ol.NRC_FP16}


def group_bounds(g):
    lo = np.array([-1.0 - 0.3 * g, -1.0, -1.2 + 0.1 * g], np.float32)
    return lo, lo + np.float32(2.0 + 0.5 * g)


def make_cache(groups, width, precision, seed=5):
    c = ap.RadianceCache(seed=seed, groups=groups, width=width, precision=precision)
    for g in range(groups):
        c.set_bounds(*group_bounds(g), group=g)
    p = np.stack([cpu.init_params_w(width, seed=100 + g) for g in range(groups)])
    c.set_params(p)
    return c, p


def samples_for(q, target):
    s = np.zeros((len(q), 20), np.float32)
    s[:, :16] = q
    s[:, 16:19] = target
    return s


@pytest.mark.parametrize("width,precision", SHAPES)
def test_grouped_forward_matches_oracle(gpu, width, precision):
    print(f"[nrc-group] forward width {width} {precision}", flush=True)
    c, p = make_cache(4, width, precision)
    assert c.groups == 4 and c.width == width and c.n_params == ol.nrc_params_w(width)
    assert np.array_equal(c.params(), p)
    sizes = [100, 0, 777, 64]
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    q = cpu.random_queries(int(seg[-1]), seed=50)
    got = c.infer_grouped(q, seg)
    for g in range(4):
        if sizes[g] == 0:
            continue
        rows = slice(int(seg[g]), int(seg[g + 1]))
        want = ol.nrc_forward_w(p[g], width, ol.nrc_encode(q[rows], *group_bounds(g)), MODE[precision])
        rel = np.abs(got[rows] - want) / (np.abs(want) + 1e-3)
        print(f"[nrc-group]   network {g}: median rel err {np.median(rel):.3g}, "
              f"p99.5 {np.quantile(rel, 0.995):.3g}", flush=True)
        assert np.median(rel) < 2e-4 and np.quantile(rel, 0.995) < 1e-2
    c.close()


@pytest.mark.parametrize("width,precision", SHAPES)
def test_grouped_gradient_matches_oracle(gpu, width, precision):
    print(f"[nrc-group] gradient width {width} {precision}", flush=True)
    G, B = 3, 1024
    c, p = make_cache(G, width, precision)
    rng = np.random.default_rng(51)
    q = cpu.random_queries(G * B, seed=52).reshape(G, B, 16)
    t = rng.uniform(0, 2, (G, B, 3)).astype(np.float32)
    s = np.stack([samples_for(q[g], t[g]) for g in range(G)])
    loss, grad = c.gradient(s)
    scale = c.loss_scale(B)
    for g in range(G):
        feat = ol.nrc_encode(q[g], *group_bounds(g))
        oloss, og = ol.nrc_grad_w(p[g], width, feat, t[g], MODE[precision], scale=scale)
        gg = grad[g].astype(np.float64)
        cos = float(gg @ og / (np.linalg.norm(gg) * np.linalg.norm(og)))
        rel = float(np.linalg.norm(gg - og) / np.linalg.norm(og))
        print(f"[nrc-group]   network {g}: loss {loss[g]:.6g} vs {oloss:.6g}, cosine {cos:.7f}, rel L2 {rel:.3g}",
              flush=True)
        assert abs(loss[g] / oloss - 1.0) < 1e-3
        assert cos > 0.9999 and rel < 1e-2
        assert np.all(grad[g][ol.nrc_params_w(width) - 16 * width + 3 * width:] == 0.0)
    c.close()


def test_grouped_adam_is_the_oracle_adam(gpu):
    G, B, W = 2, 512, 32
    c, p = make_cache(G, W, "fp16")
    q = cpu.random_queries(G * B, seed=53)
    t = np.random.default_rng(54).uniform(0, 2, (G * B, 3)).astype(np.float32)
    d = ap.DeviceBuffer.from_numpy(samples_for(q, t))
    c.train_step(d, B, lr=2e-3)
    ap.synchronize()
    g = c.last_gradient()
    got = c.params()
    for k in range(G):
        want, _, _ = ol.nrc_adam(p[k], np.zeros_like(p[k]), np.zeros_like(p[k]), g[k], 1, np.float32(2e-3))
        assert np.array_equal(got[k], want)
    c.close()


def test_networks_train_independently(gpu):
    """Two FP16 width-32 networks trained in the same launches on different
    targets each fit their own target and not the other's."""
    G, B = 2, 8192
    c = ap.RadianceCache(seed=7, groups=G, width=32, precision="fp16")
    c.set_bounds([-1, -1, -1], [1, 1, 1])
    q = cpu.random_queries(B, seed=55)
    t0 = np.stack([0.5 + 0.4 * q[:, 0], 0.3 + 0.2 * q[:, 3] ** 2, 0.2 + 0.1 * q[:, 7]], 1).astype(np.float32)
    t1 = np.stack([0.2 + 0.1 * q[:, 1], 0.6 - 0.3 * q[:, 2], 0.4 + 0.2 * q[:, 6] ** 2], 1).astype(np.float32)
    s = np.stack([samples_for(q, t0), samples_for(q, t1)])
    swapped = np.stack([samples_for(q, t1), samples_for(q, t0)])
    first, _ = c.gradient(s)
    d = ap.DeviceBuffer.from_numpy(s)
    for _ in range(200):
        c.train_step(d, B, lr=3e-3)
    ap.synchronize()
    final, _ = c.gradient(s)
    cross, _ = c.gradient(swapped)
    print(f"[nrc-group] independent training: loss {first} -> {final}, swapped targets {cross}", flush=True)
    assert np.all(final < 0.05 * first)
    assert np.all(cross > 10.0 * final)
    c.close()


def test_grouped_cache_argument_checks(gpu):
    c = ap.RadianceCache(seed=1, groups=2, width=32, precision="fp16")
    with pytest.raises(ap.ApertureError):
        c.infer(cpu.random_queries(64))
    with pytest.raises(ap.ApertureError):
        c.infer_grouped(cpu.random_queries(64), [0, 40, 30])
    with pytest.raises(ap.ApertureError):
        c.set_bounds([0, 0, 0], [1, 1, 1], group=2)
    c.close()
    with pytest.raises(ap.ApertureError):
        ap.RadianceCache(seed=1, groups=2, width=48)
    one = ap.RadianceCache(seed=1, groups=1, width=32, precision="fp16")
    assert one.infer(cpu.random_queries(64)).shape == (64, 3)
    one.close()


def test_many_tiles_per_block_equal_single_networks(gpu):
    """48 networks x 4096 rows (196,608 rows: more 64-row tiles than the
    training kernel's 2048 blocks, so blocks walk several tiles and reload
    weights between networks, as a 48-scene batch's training step does):
    each network's gradient and loss equal, bit for bit, those of the same
    network alone on its own rows."""
    G, B = 48, 4096
    c, p = make_cache(G, 32, "fp16")
    rng = np.random.default_rng(61)
    q = cpu.random_queries(G * B, seed=62).reshape(G, B, 16)
    t = rng.uniform(0, 2, (G, B, 3)).astype(np.float32)
    s = np.stack([samples_for(q[g], t[g]) for g in range(G)])
    loss, grad = c.gradient(s)
    for g in (0, 1, 31, 47):
        one = ap.RadianceCache(seed=5, groups=1, width=32, precision="fp16")
        one.set_bounds(*group_bounds(g))
        one.set_params(p[g])
        l1, g1 = one.gradient(s[g])
        print(f"[nrc-group] 48 networks, network {g}: loss {loss[g]:.6g} vs alone {l1:.6g}", flush=True)
        assert np.array_equal(g1, grad[g]) and l1 == loss[g]
        assert np.abs(g1).max() > 0.0
        one.close()
    c.close()
