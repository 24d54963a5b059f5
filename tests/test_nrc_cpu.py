"""CPU tests pinning the oracle's radiance-cache network (oracle/or_nrc.cpp):

  * in FP64 mode the analytic gradient of the relative L2 loss equals central
    finite differences for parameters of every layer (pins forward and
    backward algebra, the ReLU masks and the loss normalisation);
  * the encoding equals a numpy restatement (frequencies 2^k pi, k = 0..5);
  * Adam equals the textbook update computed in numpy;
  * BF16 mode stays within BF16 precision of the FP64 network;
  * a few hundred Adam steps fit a smooth target function (training works).
"""
import numpy as np

import oracle_lib as ol

G = np.random.default_rng(4)


def init_params(seed=0, scale=1.0):
    g = np.random.default_rng(seed)
    p = np.zeros(ol.NRC_PARAMS, np.float32)
    for l in range(5):
        p[l * 4096:(l + 1) * 4096] = g.normal(0, scale * np.sqrt(2.0 / 64), 4096)
    out = p[5 * 4096:].reshape(16, 64)
    out[:3] = g.normal(0, scale * np.sqrt(1.0 / 64), (3, 64))
    return p


def random_queries(n, seed=1):
    g = np.random.default_rng(seed)
    q = np.zeros((n, 16), np.float32)
    q[:, :3] = g.uniform(-1, 1, (n, 3))
    d = g.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = g.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9] = g.uniform(0.05, 1, n)
    q[:, 10:13] = g.uniform(0.05, 0.95, (n, 3))
    return q


BMIN, BMAX = np.array([-1.0, -1.0, -1.0], np.float32), np.array([1.0, 1.0, 1.0], np.float32)


def test_encoding_matches_numpy():
    q = random_queries(500)
    f = ol.nrc_encode(q, BMIN, BMAX)
    p = (q[:, :3] - BMIN) / (BMAX - BMIN)
    for a in range(3):
        for k in range(6):
            # The FP32 argument p * (2^k pi), then sin/cos within 2 ulp.
            arg = (p[:, a] * (np.float32(2 ** k) * np.float32(np.pi))).astype(np.float64)
            assert np.allclose(f[:, a * 12 + 2 * k], np.sin(arg), atol=3e-7)
            assert np.allclose(f[:, a * 12 + 2 * k + 1], np.cos(arg), atol=3e-7)
    assert np.array_equal(f[:, 36:39], q[:, 3:6]) and np.array_equal(f[:, 39:42], q[:, 6:9])
    assert np.array_equal(f[:, 42], q[:, 9]) and np.array_equal(f[:, 43:46], q[:, 10:13])
    assert (f[:, 46] == 1.0).all() and (f[:, 47:] == 0.0).all()


def test_fp64_gradient_matches_finite_differences():
    p = init_params(3)
    feat = ol.nrc_encode(random_queries(8, seed=5), BMIN, BMAX)
    target = G.uniform(0, 2, (8, 3)).astype(np.float32)
    _, grad = ol.nrc_grad(p, feat, target, round_bf16=False)
    # The relative loss stops the gradient through its denominator y^2 + 0.01:
    # differentiate the loss with the denominators held at the base point.
    y = ol.nrc_forward(p, feat, round_bf16=False).astype(np.float64)
    den = y * y + 0.01
    # Two parameters with non-zero gradient from every layer.
    picks = []
    for lo, hi in [(l * 4096, (l + 1) * 4096) for l in range(5)] + [(5 * 4096, 5 * 4096 + 3 * 64)]:
        nz = lo + np.flatnonzero(grad[lo:hi])
        assert len(nz) > 0
        picks += list(nz[:: max(1, len(nz) // 2)][:2])
    for k in picks:
        h = 1e-5 * max(1.0, abs(float(p[k])))  # small enough not to cross a ReLU kink
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        # Parameters are float32 in the interface: step by exactly representable amounts.
        hp = np.float32(p[k] + h) - p[k]
        pp = p.copy(); pp[k] = p[k] + hp
        pm = p.copy(); pm[k] = p[k] - hp
        lp = ol.nrc_loss_fixed(pp, feat, target, den)
        lm = ol.nrc_loss_fixed(pm, feat, target, den)
        fd = (lp - lm) / (2 * float(hp))
        assert abs(fd - grad[k]) <= 1e-5 * abs(grad[k]) + 1e-9, (k, fd, grad[k])
    # The padded output rows never receive gradient.
    assert np.all(grad[5 * 4096 + 3 * 64:] == 0.0)


def test_adam_matches_numpy():
    p = init_params(7)
    m = G.normal(0, 0.01, p.size).astype(np.float32)
    v = np.abs(G.normal(0, 0.01, p.size)).astype(np.float32)
    g = G.normal(0, 0.1, p.size).astype(np.float32)
    t, lr, b1, b2, eps = 3, np.float32(1e-3), np.float32(0.9), np.float32(0.99), np.float32(1e-8)
    p2, m2, v2 = ol.nrc_adam(p, m, v, g, t, lr, b1, b2, eps)
    mm = b1 * m + (np.float32(1) - b1) * g
    vv = b2 * v + (np.float32(1) - b2) * (g * g)
    p1 = p2_ = np.float32(1)
    for _ in range(t):  # beta^t as the FP32 product of t factors
        p1, p2_ = np.float32(p1 * b1), np.float32(p2_ * b2)
    c1 = np.float32(1) - p1
    c2 = np.float32(1) - p2_
    assert abs(float(c1) - (1 - 0.9 ** 3)) < 1e-7 and abs(float(c2) - (1 - 0.99 ** 3)) < 1e-7
    pp = p - lr * (mm / c1) / (np.sqrt(vv / c2) + eps)
    assert np.array_equal(m2, mm) and np.array_equal(v2, vv)
    assert np.allclose(p2, pp, rtol=1e-6, atol=1e-9)


def test_bf16_mode_tracks_fp64():
    p = init_params(9)
    feat = ol.nrc_encode(random_queries(256, seed=9), BMIN, BMAX)
    y64 = ol.nrc_forward(p, feat, round_bf16=False)
    y16 = ol.nrc_forward(p, feat, round_bf16=True)
    assert np.abs(y16 - y64).max() < 0.05 * np.abs(y64).max()


def test_training_fits_a_smooth_target():
    q = random_queries(2048, seed=11)
    feat = ol.nrc_encode(q, BMIN, BMAX)
    target = np.stack([0.5 + 0.4 * q[:, 0], 0.3 + 0.2 * q[:, 3] ** 2, 0.2 + 0.1 * q[:, 7]], 1).astype(np.float32)
    p = init_params(13, scale=0.5)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    first = None
    for t in range(1, 201):
        idx = G.choice(len(q), 256, replace=False)
        loss, g = ol.nrc_grad(p, feat[idx], target[idx], round_bf16=True)
        first = loss if first is None else first
        p, m, v = ol.nrc_adam(p, m, v, g.astype(np.float32), t, lr=np.float32(3e-3))
    final, _ = ol.nrc_grad(p, feat, target, round_bf16=True)
    assert final < 0.1 * first, (first, final)


def init_params_w(width, seed=0):
    g = np.random.default_rng(seed)
    n = ol.nrc_params_w(width)
    p = np.zeros(n, np.float32)
    p[:width * 64] = g.normal(0, np.sqrt(2.0 / 64), width * 64)
    p[width * 64:width * 64 + 4 * width * width] = g.normal(0, np.sqrt(2.0 / width), 4 * width * width)
    out = p[width * 64 + 4 * width * width:].reshape(16, width)
    out[:3] = g.normal(0, np.sqrt(1.0 / width), (3, width))
    return p


def test_f16_rounding_matches_numpy():
    g = np.random.default_rng(12)
    x = np.concatenate([g.normal(0, 1, 20000), g.normal(0, 1e-5, 5000), g.normal(0, 1e4, 5000),
                        [65504.0, 65519.0, 65520.0, 1e6, -1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0, -0.0],
                        np.float32(1.0) + np.arange(-4, 5) * 2.0 ** -11]).astype(np.float32)
    got = ol.f16_round(x)
    with np.errstate(over="ignore"):  # values past 65520 round to infinity on purpose
        want = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_general_width64_bf16_equals_the_specialised_network():
    p = init_params(41)
    feat = ol.nrc_encode(random_queries(64, seed=42), BMIN, BMAX)
    t = G.uniform(0, 2, (64, 3)).astype(np.float32)
    assert ol.nrc_params_w(64) == ol.NRC_PARAMS and ol.nrc_params_w(32) == 32 * 64 + 4 * 32 * 32 + 16 * 32
    assert np.array_equal(ol.nrc_forward_w(p, 64, feat, ol.NRC_BF16), ol.nrc_forward(p, feat, round_bf16=True))
    l1, g1 = ol.nrc_grad_w(p, 64, feat, t, ol.NRC_BF16)
    l2, g2 = ol.nrc_grad(p, feat, t, round_bf16=True)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_width32_gradient_matches_finite_differences():
    p = init_params_w(32, seed=43)
    feat = ol.nrc_encode(random_queries(16, seed=44), BMIN, BMAX)
    t = G.uniform(0.1, 1.5, (16, 3)).astype(np.float32)
    loss, grad = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_NONE)
    y = ol.nrc_forward_w(p, 32, feat, ol.NRC_NONE).astype(np.float64)
    # The denominator y^2 + 0.01 is a stop-gradient: hold it at the base point.
    den = y ** 2 + 0.01
    assert abs(ol.nrc_loss_fixed_w(p, 32, feat, t, den) / loss - 1.0) < 1e-6
    rng = np.random.default_rng(45)
    picks = np.concatenate([rng.integers(0, 2048, 6), rng.integers(2048, 2048 + 4096, 6),
                            rng.integers(2048 + 4096, 2048 + 4096 + 96, 4)])
    for k in picks:
        q1, q2 = p.copy(), p.copy()
        q1[k] += np.float32(1e-5)
        q2[k] -= np.float32(1e-5)
        fd = (ol.nrc_loss_fixed_w(q1, 32, feat, t, den) - ol.nrc_loss_fixed_w(q2, 32, feat, t, den)) / \
            (float(q1[k]) - float(q2[k]))
        assert abs(fd - grad[k]) < 1e-4 * max(1e-3, abs(grad[k])) + 1e-6, (k, fd, grad[k])
    assert np.all(grad[2048 + 4096 + 3 * 32:] == 0.0)
    assert ol.nrc_params_w(32) == 6656


def test_loss_scale_is_exact_and_keeps_fp16_gradients_out_of_underflow():
    """The loss-gradient scale multiplies before rounding and divides after:
    without rounding a power of two changes no bit. In FP16 a batch of 4096
    puts the unscaled loss gradient (~4e-5) among FP16's subnormals; the
    scale 2^(ceil(log2(3 * 4096)) - 4) = 1024 lifts it out, so that the
    gradient agrees with one scaled 64 times further (no underflow at all)
    far more closely than the unscaled one does, and stays within FP16's
    forward rounding of the FP64 gradient."""
    p = init_params_w(32, seed=46)
    n = 4096
    feat = ol.nrc_encode(random_queries(n, seed=47), BMIN, BMAX)
    t = G.uniform(0.1, 1.5, (n, 3)).astype(np.float32)
    l1, g1 = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_NONE)
    l2, g2 = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_NONE, scale=1024.0)
    assert l1 == l2 and np.array_equal(g1, g2)
    _, g_raw = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_FP16)
    _, g_sc = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_FP16, scale=1024.0)
    _, g_hi = ol.nrc_grad_w(p, 32, feat, t, ol.NRC_FP16, scale=1024.0 * 64)

    def rel(g, h):
        return float(np.linalg.norm(g - h) / np.linalg.norm(h))

    assert rel(g_sc, g_hi) < 1e-7 and rel(g_raw, g_hi) > 100.0 * rel(g_sc, g_hi), (rel(g_sc, g_hi), rel(g_raw, g_hi))
    assert rel(g_sc, g1) < 5e-3


def test_fp16_mode_saturates():
    """Weights beyond FP16's range become +-65504, not infinities, so the
    network stays finite (an infinity times a zero activation would be NaN)."""
    p = init_params_w(32, seed=48)
    p[5] = 1e6
    feat = ol.nrc_encode(random_queries(64, seed=49), BMIN, BMAX)
    y = ol.nrc_forward_w(p, 32, feat, ol.NRC_FP16)
    assert np.isfinite(y).all()
    q = p.copy()
    q[5] = 65504.0
    assert np.array_equal(y, ol.nrc_forward_w(q, 32, feat, ol.NRC_FP16))
