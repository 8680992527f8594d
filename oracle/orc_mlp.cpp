// CPU oracle for the collision surrogate — TEST INFRASTRUCTURE ONLY.
//
// The surrogate has no counterpart in the reference; it only orders the
// waypoint checks of link() (every verdict still comes from the exact
// checker).  This file is the CPU statement of the network the GPU trains:
//   x (64, BF16) -> 128 ReLU -> 128 ReLU -> 128 ReLU -> 1 logit
// with BF16 operands and FP32 accumulation, binary cross-entropy on the
// logit, and Adam.  Rounding points mirror the GPU kernels exactly:
//   weights and activations are rounded to BF16 before every product,
//   back-propagated activation gradients are rounded to BF16 before the
//   GEMMs that consume them; biases, logits, loss and weight gradients are FP32.
// With `exact` set, no rounding is applied and everything is FP64 — that
// mode is differentiated numerically in the tests to pin the formulas.
//
// Parameter layout (floats), shared with the GPU:
//   W1[64][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] w4[128] b4[1]
// (W[i][o]: input i, output o).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../mpt_amd/csrc/hd.hpp"

namespace {

constexpr int IN = 64, H = 128;
constexpr int64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
                  OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H, OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H,
                  NPARAM = OFF_B4 + 1;

template <typename T>
struct Net {
    bool exact;
    T r(T v) const {
        if (exact) return v;
        return (T)mptg::bf16_to_float(mptg::float_to_bf16((float)v));
    }
};

template <typename T>
T loss_and_grad(const float* p, const uint16_t* X, const float* y, int64_t N, T* g, bool exact, T* logits) {
    Net<T> net{exact};
    auto P = [&](int64_t i) { return net.r((T)p[i]); };
    std::vector<T> W1(IN * H), W2(H * H), W3(H * H), w4(H), b1(H), b2(H), b3(H);
    for (int i = 0; i < IN * H; ++i) W1[i] = P(OFF_W1 + i);
    for (int i = 0; i < H * H; ++i) {
        W2[i] = P(OFF_W2 + i);
        W3[i] = P(OFF_W3 + i);
    }
    for (int i = 0; i < H; ++i) {
        w4[i] = P(OFF_W4 + i);
        b1[i] = (T)p[OFF_B1 + i];
        b2[i] = (T)p[OFF_B2 + i];
        b3[i] = (T)p[OFF_B3 + i];
    }
    T b4 = (T)p[OFF_B4];
    std::vector<T> h1((size_t)N * H), h2((size_t)N * H), h3((size_t)N * H), z(N);
    std::vector<T> x((size_t)N * IN);
    for (int64_t i = 0; i < N * IN; ++i) x[i] = (T)mptg::bf16_to_float(X[i]);
    auto layer = [&](const std::vector<T>& in, int nin, const std::vector<T>& W, const std::vector<T>& b,
                     std::vector<T>& out) {
        for (int64_t r = 0; r < N; ++r)
            for (int o = 0; o < H; ++o) {
                T acc = 0;
                for (int i = 0; i < nin; ++i) acc += in[r * nin + i] * W[i * H + o];
                T v = acc + b[o];
                out[r * H + o] = net.r(v > 0 ? v : 0);
            }
    };
    layer(x, IN, W1, b1, h1);
    layer(h1, H, W2, b2, h2);
    layer(h2, H, W3, b3, h3);
    T loss = 0;
    std::vector<T> dz(N);
    for (int64_t r = 0; r < N; ++r) {
        T acc = 0;
        for (int k = 0; k < H; ++k) acc += h3[r * H + k] * w4[k];
        z[r] = acc + b4;
        if (logits) logits[r] = z[r];
        T zz = z[r], yy = (T)y[r];
        loss += std::max(zz, (T)0) - zz * yy + std::log1p(std::exp(-std::fabs(zz)));
        T sig = 1 / (1 + std::exp(-zz));
        dz[r] = (sig - yy) / (T)N;
    }
    loss /= (T)N;
    if (!g) return loss;
    std::fill(g, g + NPARAM, (T)0);
    // output layer
    for (int64_t r = 0; r < N; ++r) {
        for (int k = 0; k < H; ++k) g[OFF_W4 + k] += h3[r * H + k] * dz[r];
        g[OFF_B4] += dz[r];
    }
    std::vector<T> d3((size_t)N * H), d2((size_t)N * H), d1((size_t)N * H);
    for (int64_t r = 0; r < N; ++r)
        for (int k = 0; k < H; ++k) d3[r * H + k] = net.r(h3[r * H + k] > 0 ? dz[r] * w4[k] : 0);
    auto back = [&](const std::vector<T>& dout, const std::vector<T>& hin, int nin, const std::vector<T>& W,
                    int64_t offW, int64_t offB, std::vector<T>* din) {
        for (int64_t r = 0; r < N; ++r)
            for (int o = 0; o < H; ++o) {
                T d = dout[r * H + o];
                if (d == 0) continue;
                g[offB + o] += d;
                for (int i = 0; i < nin; ++i) g[offW + i * H + o] += hin[r * nin + i] * d;
            }
        if (!din) return;
        for (int64_t r = 0; r < N; ++r)
            for (int i = 0; i < nin; ++i) {
                T acc = 0;
                for (int o = 0; o < H; ++o) acc += dout[r * H + o] * W[i * H + o];
                (*din)[r * nin + i] = net.r(hin[r * nin + i] > 0 ? acc : 0);
            }
    };
    back(d3, h2, H, W3, OFF_W3, OFF_B3, &d2);
    back(d2, h1, H, W2, OFF_W2, OFF_B2, &d1);
    back(d1, x, IN, W1, OFF_W1, OFF_B1, nullptr);
    return loss;
}

}  // namespace

extern "C" {

int64_t orc_mlp_nparams(void) { return NPARAM; }

void orc_bf16_round(const float* x, int64_t n, uint16_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = mptg::float_to_bf16(x[i]);
}

// FP32 mode (GPU rounding points).  grad may be null.  Returns the mean loss.
float orc_mlp_loss_grad(const float* params, const uint16_t* X, const float* y, int64_t N, float* grad,
                        float* logits) {
    return loss_and_grad<float>(params, X, y, N, grad, false, logits);
}

// FP64, no rounding: the smooth function the FP32 mode approximates.
double orc_mlp_loss_grad_exact(const float* params, const uint16_t* X, const float* y, int64_t N, double* grad) {
    return loss_and_grad<double>(params, X, y, N, grad, true, nullptr);
}

// One Adam step (t = step number starting at 1) on FP32 master weights.
void orc_adam_step(float* params, float* m, float* v, const float* grad, int64_t n, float lr, float beta1,
                   float beta2, float eps, int32_t t) {
    float c1 = 1.0f - std::pow(beta1, (float)t);
    float c2 = 1.0f - std::pow(beta2, (float)t);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0f - beta1) * grad[i];
        v[i] = beta2 * v[i] + (1.0f - beta2) * grad[i] * grad[i];
        float mh = m[i] / c1;
        float vh = v[i] / c2;
        params[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"
