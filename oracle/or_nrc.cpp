// ORACLE — test infrastructure only: CPU restatement of the learned
// radiance cache's network (the checker for csrc/ap_nrc.hip).
//
// The reference has no radiance cache; this follows the published neural
// radiance cache (Mueller, Rousselle, Novak, Keller 2021, "Real-time Neural
// Radiance Caching for Path Tracing"): a fully connected ReLU network without
// biases, relative L2 loss L = mean((y - t)^2 / (sg(y)^2 + 0.01)), Adam
// (Kingma & Ba 2015). Sizes and encoding are the core's:
//   input 64 = position in the cache bounds, 6 frequencies x {sin, cos} per
//   axis (36), outgoing direction (3), normal (3), roughness (1), albedo (3),
//   a constant 1, zero padding; 5 hidden layers of 64; 3 outputs (RGB).
// Parameters (21504 floats): 5 hidden matrices W[o][i] 64x64, then the output
// matrix 16x64 whose rows 3..15 stay zero. The width-w variant (the *_w
// functions; w = 32 or 64) has W0 w x 64, W1..W4 w x w and the output 16 x w:
// w*64 + 4*w*w + 16*w parameters.
//
// Arithmetic is FP64. With round_bf16 the values the core stores in BF16 are
// rounded to BF16 (round to nearest even) at the same points: weights,
// encoded inputs, hidden activations, output deltas and hidden deltas; the
// *_w functions take a precision (0 none, 1 BF16, 2 FP16 with round to
// nearest even, subnormals and overflow to infinity). The accumulation order
// differs from the core's MFMA (FP32), so the core is compared with a stated
// tolerance. Without rounding the functions are the exact FP64 network,
// pinned by finite differences in tests/test_nrc_cpu.py.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ap_fmath.h"
#include "or_api.h"

namespace {

constexpr int IN = 64, H = 64, NH = 5, OUTP = 16, NOUT = 3;
constexpr int PARAMS = NH * H * IN + OUTP * H;  // 21504

// FP32 -> IEEE binary16 -> FP32, round to nearest even.
double f16(double x) {
    float f = static_cast<float>(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const int32_t exp = static_cast<int32_t>((u >> 23) & 0xffu);
    uint32_t mant = u & 0x7fffffu;
    uint32_t h;
    if (exp == 0xff) {
        h = sign | 0x7c00u | (mant ? 0x200u : 0u);
    } else {
        const int32_t e = exp - 127 + 15;
        if (e >= 31) {
            h = sign | 0x7c00u;
        } else if (e <= 0) {
            if (e < -10) {
                h = sign;
            } else {
                mant |= 0x800000u;
                const int shift = 14 - e;
                uint32_t v = mant >> shift;
                const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
                if (rem > half || (rem == half && (v & 1u))) ++v;
                h = sign | v;
            }
        } else {
            uint32_t v = (static_cast<uint32_t>(e) << 10) | (mant >> 13);
            const uint32_t rem = mant & 0x1fffu;
            if (rem > 0x1000u || (rem == 0x1000u && (v & 1u))) ++v;
            h = sign | v;  // a carry into the exponent rounds up correctly, to infinity at the top
        }
    }
    // binary16 -> float
    const uint32_t hs = (h & 0x8000u) << 16, he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
    uint32_t out;
    if (he == 0) {
        if (hm == 0) {
            out = hs;
        } else {
            int k = 0;
            uint32_t m = hm;
            while (!(m & 0x400u)) { m <<= 1; ++k; }
            out = hs | (static_cast<uint32_t>(127 - 15 - k + 1) << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (he == 31) {
        out = hs | 0x7f800000u | (hm << 13);
    } else {
        out = hs | ((he - 15 + 127) << 23) | (hm << 13);
    }
    float r;
    std::memcpy(&r, &out, 4);
    return r;
}

double bf16(double x) {
    float f = static_cast<float>(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return std::nan("");
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Net {
    std::vector<double> w[NH + 1];
    Net(const float* params, bool rnd) {
        size_t off = 0;
        for (int l = 0; l <= NH; ++l) {
            const int rows = l < NH ? H : OUTP;
            w[l].resize(static_cast<size_t>(rows) * IN);
            for (size_t k = 0; k < w[l].size(); ++k) w[l][k] = rnd ? bf16(params[off + k]) : params[off + k];
            off += w[l].size();
        }
    }
};

// Forward of one sample; acts[0] = input, acts[1..5] = hidden activations.
void forward(const Net& net, const float* feat, bool rnd, double acts[NH + 1][H], double y[NOUT]) {
    for (int i = 0; i < IN; ++i) acts[0][i] = rnd ? bf16(feat[i]) : feat[i];
    for (int l = 0; l < NH; ++l) {
        for (int o = 0; o < H; ++o) {
            double s = 0.0;
            for (int i = 0; i < IN; ++i) s += net.w[l][static_cast<size_t>(o) * IN + i] * acts[l][i];
            s = s > 0.0 ? s : 0.0;
            acts[l + 1][o] = rnd ? bf16(s) : s;
        }
    }
    for (int o = 0; o < NOUT; ++o) {
        double s = 0.0;
        for (int i = 0; i < H; ++i) s += net.w[NH][static_cast<size_t>(o) * IN + i] * acts[NH][i];
        y[o] = rnd ? static_cast<float>(s) : s;
    }
}

// FP16 rounding saturated at +-65504 (what the FP16 networks store).
double f16sat(double x) {
    const double r = f16(x);
    return std::isinf(r) ? std::copysign(65504.0, r) : r;
}

// Width-w network with a rounding mode (0 none, 1 BF16, 2 FP16).
double rnd(double x, int mode) { return mode == 1 ? bf16(x) : (mode == 2 ? f16sat(x) : x); }

struct NetW {
    int w;
    std::vector<double> m[NH + 1];  // m[l][o * in + i]
    int in(int l) const { return l == 0 ? IN : w; }
    int out(int l) const { return l < NH ? w : OUTP; }
    NetW(const float* params, int width, int mode) : w(width) {
        size_t off = 0;
        for (int l = 0; l <= NH; ++l) {
            m[l].resize(static_cast<size_t>(out(l)) * in(l));
            for (size_t k = 0; k < m[l].size(); ++k) m[l][k] = rnd(params[off + k], mode);
            off += m[l].size();
        }
    }
};

void forward_w(const NetW& net, const float* feat, int mode, std::vector<std::vector<double>>& acts, double y[NOUT]) {
    acts.assign(NH + 1, std::vector<double>(std::max(IN, net.w), 0.0));
    for (int i = 0; i < IN; ++i) acts[0][i] = rnd(feat[i], mode);
    for (int l = 0; l < NH; ++l)
        for (int o = 0; o < net.w; ++o) {
            double s = 0.0;
            for (int i = 0; i < net.in(l); ++i) s += net.m[l][static_cast<size_t>(o) * net.in(l) + i] * acts[l][i];
            s = s > 0.0 ? s : 0.0;
            acts[l + 1][o] = rnd(s, mode);
        }
    for (int o = 0; o < NOUT; ++o) {
        double s = 0.0;
        for (int i = 0; i < net.w; ++i) s += net.m[NH][static_cast<size_t>(o) * net.w + i] * acts[NH][i];
        y[o] = mode ? static_cast<float>(s) : s;
    }
}

}  // namespace

extern "C" {

size_t or_nrc_params_w(int width) {
    return static_cast<size_t>(width) * IN + 4u * width * width + static_cast<size_t>(OUTP) * width;
}

void or_nrc_forward_w(const float* params, int width, const float* feat, size_t n, int mode, float* out) {
    const NetW net(params, width, mode);
    std::vector<std::vector<double>> acts;
    double y[NOUT];
    for (size_t r = 0; r < n; ++r) {
        forward_w(net, feat + 64 * r, mode, acts, y);
        for (int o = 0; o < NOUT; ++o) out[NOUT * r + o] = static_cast<float>(y[o]);
    }
}

// Gradient of the mean relative L2 loss of the width-w network (as
// or_nrc_grad), with the loss gradient scaled by `scale` before its rounding
// and the weight gradient divided by it again.
double or_nrc_grad_w(const float* params, int width, const float* feat, const float* target, size_t n, int mode,
                     double scale, double* grad) {
    const NetW net(params, width, mode);
    const size_t np = or_nrc_params_w(width);
    for (size_t k = 0; k < np; ++k) grad[k] = 0.0;
    size_t off[NH + 1];
    {
        size_t o = 0;
        for (int l = 0; l <= NH; ++l) { off[l] = o; o += static_cast<size_t>(net.out(l)) * net.in(l); }
    }
    double loss = 0.0;
    const double norm = 1.0 / (static_cast<double>(n) * NOUT);
    std::vector<std::vector<double>> acts;
    double y[NOUT];
    const int w = net.w;
    for (size_t r = 0; r < n; ++r) {
        forward_w(net, feat + 64 * r, mode, acts, y);
        double dy[OUTP] = {0.0};
        for (int o = 0; o < NOUT; ++o) {
            const double t = target[NOUT * r + o];
            const double den = y[o] * y[o] + 0.01;
            const double e = y[o] - t;
            loss += e * e / den * norm;
            dy[o] = rnd(2.0 * e / den * norm * scale, mode);
        }
        for (int o = 0; o < NOUT; ++o)
            for (int i = 0; i < w; ++i) grad[off[NH] + static_cast<size_t>(o) * w + i] += dy[o] * acts[NH][i];
        std::vector<double> delta(w), prev(w);
        for (int i = 0; i < w; ++i) {
            double s = 0.0;
            for (int o = 0; o < NOUT; ++o) s += net.m[NH][static_cast<size_t>(o) * w + i] * dy[o];
            s = acts[NH][i] > 0.0 ? s : 0.0;
            delta[i] = rnd(s, mode);
        }
        for (int l = NH - 1; l >= 0; --l) {
            const int nin = net.in(l);
            for (int o = 0; o < w; ++o)
                for (int i = 0; i < nin; ++i) grad[off[l] + static_cast<size_t>(o) * nin + i] += delta[o] * acts[l][i];
            if (l == 0) break;
            for (int i = 0; i < w; ++i) {
                double s = 0.0;
                for (int o = 0; o < w; ++o) s += net.m[l][static_cast<size_t>(o) * w + i] * delta[o];
                s = acts[l][i] > 0.0 ? s : 0.0;
                prev[i] = rnd(s, mode);
            }
            delta = prev;
        }
    }
    for (size_t k = 0; k < np; ++k) grad[k] /= scale;
    return loss;
}

// Loss of the width-w network with the denominators held at `den` (FP64).
double or_nrc_loss_fixed_w(const float* params, int width, const float* feat, const float* target, const double* den,
                           size_t n, int mode) {
    const NetW net(params, width, mode);
    std::vector<std::vector<double>> acts;
    double y[NOUT], loss = 0.0;
    const double norm = 1.0 / (static_cast<double>(n) * NOUT);
    for (size_t r = 0; r < n; ++r) {
        forward_w(net, feat + 64 * r, mode, acts, y);
        for (int o = 0; o < NOUT; ++o) {
            const double e = y[o] - target[NOUT * r + o];
            loss += e * e / den[NOUT * r + o] * norm;
        }
    }
    return loss;
}

// FP16 round trip of FP32 values (round to nearest even), for its CPU pins.
void or_f16_round(const float* x, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) out[i] = static_cast<float>(f16(x[i]));
}

// Encoding of query records (16 floats: pos 3, dir 3, normal 3, roughness,
// albedo 3, pad 3) into 64 features, with the core's FP32 operations.
void or_nrc_encode(const float* q, size_t n, const float* bmin, const float* bmax, float* out) {
    const float kPi = 3.14159265358979323846f;
    for (size_t r = 0; r < n; ++r) {
        const float* s = q + 16 * r;
        float* f = out + 64 * r;
        for (int k = 0; k < 64; ++k) f[k] = 0.0f;
        for (int a = 0; a < 3; ++a) {
            const float p = (s[a] - bmin[a]) / (bmax[a] - bmin[a]);
            for (int k = 0; k < 6; ++k) {
                const float arg = p * (static_cast<float>(1 << k) * kPi);
                float sn, cs;
                apfm::sincos(arg, sn, cs);
                f[a * 12 + 2 * k] = sn;
                f[a * 12 + 2 * k + 1] = cs;
            }
        }
        for (int k = 0; k < 3; ++k) {
            f[36 + k] = s[3 + k];
            f[39 + k] = s[6 + k];
            f[43 + k] = s[10 + k];
        }
        f[42] = s[9];
        f[46] = 1.0f;
    }
}

void or_nrc_forward(const float* params, const float* feat, size_t n, int round_bf16, float* out) {
    const Net net(params, round_bf16 != 0);
    double acts[NH + 1][H], y[NOUT];
    for (size_t r = 0; r < n; ++r) {
        forward(net, feat + 64 * r, round_bf16 != 0, acts, y);
        for (int o = 0; o < NOUT; ++o) out[NOUT * r + o] = static_cast<float>(y[o]);
    }
}

// Gradient of the mean relative L2 loss over n samples w.r.t. the FP32
// master parameters (through the BF16 copies when round_bf16); returns the loss.
double or_nrc_grad(const float* params, const float* feat, const float* target, size_t n, int round_bf16,
                   double* grad) {
    const bool rnd = round_bf16 != 0;
    const Net net(params, rnd);
    for (int k = 0; k < PARAMS; ++k) grad[k] = 0.0;
    double loss = 0.0;
    const double norm = 1.0 / (static_cast<double>(n) * NOUT);
    double acts[NH + 1][H], y[NOUT];
    for (size_t r = 0; r < n; ++r) {
        forward(net, feat + 64 * r, rnd, acts, y);
        double dy[OUTP] = {0.0};
        for (int o = 0; o < NOUT; ++o) {
            const double t = target[NOUT * r + o];
            const double den = y[o] * y[o] + 0.01;
            const double e = y[o] - t;
            loss += e * e / den * norm;
            dy[o] = 2.0 * e / den * norm;
            if (rnd) dy[o] = bf16(dy[o]);
        }
        double* gout = grad + NH * H * IN;
        for (int o = 0; o < NOUT; ++o)
            for (int i = 0; i < H; ++i) gout[static_cast<size_t>(o) * IN + i] += dy[o] * acts[NH][i];
        double delta[H];
        for (int i = 0; i < H; ++i) {
            double s = 0.0;
            for (int o = 0; o < NOUT; ++o) s += net.w[NH][static_cast<size_t>(o) * IN + i] * dy[o];
            s = acts[NH][i] > 0.0 ? s : 0.0;
            delta[i] = rnd ? bf16(s) : s;
        }
        for (int l = NH - 1; l >= 0; --l) {
            double* g = grad + static_cast<size_t>(l) * H * IN;
            for (int o = 0; o < H; ++o)
                for (int i = 0; i < IN; ++i) g[static_cast<size_t>(o) * IN + i] += delta[o] * acts[l][i];
            if (l == 0) break;
            double prev[H];
            for (int i = 0; i < H; ++i) {
                double s = 0.0;
                for (int o = 0; o < H; ++o) s += net.w[l][static_cast<size_t>(o) * IN + i] * delta[o];
                s = acts[l][i] > 0.0 ? s : 0.0;
                prev[i] = rnd ? bf16(s) : s;
            }
            std::memcpy(delta, prev, sizeof(delta));
        }
    }
    return loss;
}

// Loss with the per-output denominators held at `den` (the stop-gradient of
// the relative L2 loss made explicit), for finite-difference checks.
double or_nrc_loss_fixed(const float* params, const float* feat, const float* target, const double* den, size_t n,
                         int round_bf16) {
    const Net net(params, round_bf16 != 0);
    double acts[NH + 1][H], y[NOUT], loss = 0.0;
    const double norm = 1.0 / (static_cast<double>(n) * NOUT);
    for (size_t r = 0; r < n; ++r) {
        forward(net, feat + 64 * r, round_bf16 != 0, acts, y);
        for (int o = 0; o < NOUT; ++o) {
            const double e = y[o] - target[NOUT * r + o];
            loss += e * e / den[NOUT * r + o] * norm;
        }
    }
    return loss;
}

// One Adam step (t = step number from 1) on FP32 master parameters; the
// bias corrections 1 - beta^t take beta^t as the FP32 product of t factors
// (the core keeps that product on the device).
void or_nrc_adam_n(float* params, float* m, float* v, const float* grad, size_t n, int t, float lr, float beta1,
                   float beta2, float eps) {
    float p1 = 1.0f, p2 = 1.0f;
    for (int k = 0; k < t; ++k) {
        p1 = p1 * beta1;
        p2 = p2 * beta2;
    }
    const float c1 = 1.0f - p1;
    const float c2 = 1.0f - p2;
    for (size_t k = 0; k < n; ++k) {
        const float g = grad[k];
        m[k] = beta1 * m[k] + (1.0f - beta1) * g;
        v[k] = beta2 * v[k] + (1.0f - beta2) * (g * g);
        const float mh = m[k] / c1;
        const float vh = v[k] / c2;
        params[k] = params[k] - lr * mh / (std::sqrt(vh) + eps);
    }
}

void or_nrc_adam(float* params, float* m, float* v, const float* grad, int t, float lr, float beta1, float beta2,
                 float eps) {
    or_nrc_adam_n(params, m, v, grad, PARAMS, t, lr, beta1, beta2, eps);
}

}  // extern "C"
