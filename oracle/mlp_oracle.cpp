// CPU oracle of the surrogate network (test infrastructure only; nothing in
// libgrrt.so links it).  The reference has no learned component, so this
// file restates the mathematics of the network the library trains: a
// 5 -> 256 -> 256 -> 256 -> 256 -> 9 multilayer perceptron with SiLU
// activations, its loss (softmax cross-entropy on 3 class logits plus
// weighted squared errors on ln r, ln g and the escape direction) and the
// exact gradient by reverse-mode differentiation.
//
// Two arithmetic modes:
//   emulate = 0  plain FP64 everywhere (pinned against central finite
//                differences of the same FP64 forward in the tests);
//   emulate = 1  the rounding points of the GPU path: BF16 (round to nearest
//                even) weights, BF16 activations of hidden layers 0-2,
//                inputs entering layer 0 and the last hidden layer's
//                activations entering the output layer as BF16 hi + lo pairs
//                (kept exact here), the BF16 copy of that last layer in the
//                output-layer weight gradient, BF16 dZ in the weight and
//                activation gradients, FP32-exact inputs in the layer-0
//                weight gradient; sums in FP64.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 5, kHid = 256, kOut = 9;
constexpr int oW0 = 0, oB0 = oW0 + kHid * kIn;
constexpr int oW1 = oB0 + kHid, oB1 = oW1 + kHid * kHid;
constexpr int oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid;
constexpr int oW4 = oB3 + kHid, oB4 = oW4 + kOut * kHid;
constexpr int kNumParams = oB4 + kOut;

// BF16 with round-to-nearest-even of the FP32 value (NaN kept quiet).
float bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
        u |= 0x00400000u;
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
    }
    u &= 0xFFFF0000u;
    float y;
    std::memcpy(&y, &u, 4);
    return y;
}

double rnd(double x, bool em) { return em ? static_cast<double>(bf16(static_cast<float>(x))) : x; }

double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }
double silu(double z) { return z * sigmoid(z); }
double dsilu(double z) {
    const double s = sigmoid(z);
    return s * (1.0 + z * (1.0 - s));
}

struct Net {
    const float* p;
    bool em;
    int K(int l) const { return l == 0 ? kIn : kHid; }
    int N(int l) const { return l == 4 ? kOut : kHid; }
    int wo(int l) const { return l == 0 ? oW0 : l == 1 ? oW1 : l == 2 ? oW2 : l == 3 ? oW3 : oW4; }
    int bo(int l) const { return l == 0 ? oB0 : l == 1 ? oB1 : l == 2 ? oB2 : l == 3 ? oB3 : oB4; }
    double W(int l, int n, int k) const { return rnd(p[wo(l) + n * K(l) + k], em); }
    double b(int l, int n) const { return p[bo(l) + n]; }
};

// Forward for one row; z[l][n] pre-activations, h[l][n] (rounded)
// activations of the hidden layers (h[3] unrounded: the output layer's
// input), y outputs.
void forward_row(const Net& net, const float* x, double z[4][kHid], double h[4][kHid], double y[kOut]) {
    double xin[kIn];
    for (int i = 0; i < kIn; ++i) {
        if (net.em) {
            const float hi = bf16(x[i]);
            const float lo = bf16(x[i] - hi);
            xin[i] = static_cast<double>(hi) + static_cast<double>(lo);
        } else {
            xin[i] = x[i];
        }
    }
    for (int n = 0; n < kHid; ++n) {
        double s = net.b(0, n);
        for (int i = 0; i < kIn; ++i) s += net.W(0, n, i) * xin[i];
        z[0][n] = s;
        h[0][n] = rnd(silu(s), net.em);
    }
    for (int l = 1; l <= 3; ++l)
        for (int n = 0; n < kHid; ++n) {
            double s = net.b(l, n);
            for (int k = 0; k < kHid; ++k) s += net.W(l, n, k) * h[l - 1][k];
            z[l][n] = s;
            h[l][n] = l < 3 ? rnd(silu(s), net.em) : silu(s);
        }
    for (int o = 0; o < kOut; ++o) {
        double s = net.b(4, o);
        for (int k = 0; k < kHid; ++k) s += net.W(4, o, k) * h[3][k];
        y[o] = s;
    }
}

// Loss of one row and its gradient with respect to the outputs (not yet
// divided by the batch size).  t = (class or -1, ln r, ln g, dir[4], escaped).
double loss_row(const double* y, const float* t, const float* w, double* dy) {
    for (int o = 0; o < kOut; ++o) dy[o] = 0.0;
    const int cls = static_cast<int>(t[0]);
    if (cls < 0) return 0.0;
    const double mx = std::fmax(y[0], std::fmax(y[1], y[2]));
    double e[3], S = 0.0;
    for (int k = 0; k < 3; ++k) S += (e[k] = std::exp(y[k] - mx));
    double L = -(y[cls] - mx - std::log(S));
    for (int k = 0; k < 3; ++k) dy[k] = e[k] / S - (k == cls ? 1.0 : 0.0);
    if (cls == 2) {
        const double er = y[3] - t[1], eg = y[4] - t[2];
        L += w[0] * er * er + w[1] * eg * eg;
        dy[3] = 2.0 * w[0] * er;
        dy[4] = 2.0 * w[1] * eg;
    }
    if (t[7] > 0.5f) {
        for (int k = 0; k < 4; ++k) {
            const double d = y[5 + k] - t[3 + k];
            L += w[2] * d * d;
            dy[5 + k] = 2.0 * w[2] * d;
        }
    }
    return L;
}

}  // namespace

extern "C" {

int mlp_num_params() { return kNumParams; }

float mlp_bf16(float x) { return bf16(x); }

void mlp_forward(const float* params, const float* x, long n, int emulate, double* y) {
    const Net net{params, emulate != 0};
    std::vector<double> z(4 * kHid), h(4 * kHid);
    for (long r = 0; r < n; ++r)
        forward_row(net, x + r * kIn, reinterpret_cast<double(*)[kHid]>(z.data()),
                    reinterpret_cast<double(*)[kHid]>(h.data()), y + r * kOut);
}

// Mean loss over n rows and its gradient (kNumParams doubles).
double mlp_gradients(const float* params, const float* x, const float* t, long n, const float* w, int emulate,
                     double* grad) {
    const Net net{params, emulate != 0};
    const bool em = net.em;
    for (int i = 0; i < kNumParams; ++i) grad[i] = 0.0;
    double z[4][kHid], h[4][kHid], y[kOut], dy[kOut];
    double dh[kHid], dz[kHid], dzr[kHid];
    double L = 0.0;
    const double invB = 1.0 / static_cast<double>(n);
    for (long r = 0; r < n; ++r) {
        const float* xr = x + r * kIn;
        forward_row(net, xr, z, h, y);
        L += loss_row(y, t + r * 8, w, dy);
        for (int o = 0; o < kOut; ++o) {
            // The GPU forms dY in FP32 with the 1/B factor applied.
            dy[o] = em ? static_cast<double>(static_cast<float>(dy[o] * invB)) : dy[o] * invB;
        }
        // Output layer.
        for (int o = 0; o < kOut; ++o) {
            for (int k = 0; k < kHid; ++k) grad[oW4 + o * kHid + k] += dy[o] * rnd(h[3][k], em);
            grad[oB4 + o] += dy[o];
        }
        for (int k = 0; k < kHid; ++k) {
            double s = 0.0;
            for (int o = 0; o < kOut; ++o) s += dy[o] * net.W(4, o, k);
            dh[k] = s;
        }
        for (int l = 3; l >= 1; --l) {
            for (int nn = 0; nn < kHid; ++nn) {
                dz[nn] = dh[nn] * dsilu(z[l][nn]);
                dzr[nn] = rnd(dz[nn], em);
            }
            const int wo = net.wo(l), bo = net.bo(l);
            for (int nn = 0; nn < kHid; ++nn) {
                for (int k = 0; k < kHid; ++k) grad[wo + nn * kHid + k] += dzr[nn] * h[l - 1][k];
                grad[bo + nn] += dz[nn];
            }
            for (int k = 0; k < kHid; ++k) {
                double s = 0.0;
                for (int nn = 0; nn < kHid; ++nn) s += dzr[nn] * net.W(l, nn, k);
                dh[k] = s;
            }
        }
        for (int m = 0; m < kHid; ++m) {
            const double d0 = dh[m] * dsilu(z[0][m]);
            for (int i = 0; i < kIn; ++i) grad[oW0 + m * kIn + i] += d0 * static_cast<double>(xr[i]);
            grad[oB0 + m] += d0;
        }
    }
    return L * invB;
}

// One Adam step (Kingma & Ba 2015, Algorithm 1) on FP32 state.
void mlp_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, long t) {
    const double bc1 = 1.0 - std::pow(static_cast<double>(b1), static_cast<double>(t));
    const double bc2 = 1.0 - std::pow(static_cast<double>(b2), static_cast<double>(t));
    for (long i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
        const double mh = m[i] / bc1, vh = v[i] / bc2;
        w[i] = static_cast<float>(w[i] - lr * mh / (std::sqrt(vh) + eps));
    }
}

}  // extern "C"
