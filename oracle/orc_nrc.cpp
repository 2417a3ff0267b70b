// ORACLE -- test infrastructure only (see orc_common.h).
//
// FP64 restatement of the radiance-cache network: input encoding, forward pass, relative L2 loss
// and its parameter gradient.  With bf16 = 1 it rounds at the points where the device kernels
// store BF16 (encoded inputs, weights, hidden activations, pre-activation gradients) so that the
// device's FP32-accumulated results can be compared tightly; with bf16 = 0 it is the exact network.
//
// Encoding follows the reference's encoder (renderernrc.cpp:50-64, 108-121): triangle waves
// 2|fmod(x / 2^i, 2) - 1| - 1, i = 0..11, of the three position coordinates; then the query's
// spherical angles of wo and n, exp(-roughness) and the albedo, and a constant 1.
// Network layout: W^T per layer, rows = output units, 64 inputs per row (5 x 64 x 64, then 16 x 64).
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int NP = 5 * 4096 + 16 * 64;

float bf16r(float x) {  // round to nearest even at 8 significand bits
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float y;
    std::memcpy(&y, &u, 4);
    return y;
}

void encode(const float* q, float* f) {
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 12; ++i) {
            const double x = double(q[a]) / std::pow(2.0, i);
            f[a * 12 + i] = float(2 * std::fabs(std::fmod(x, 2.0) - 1) - 1);
        }
    for (int k = 0; k < 4; ++k) f[36 + k] = q[4 + k];
    f[40] = float(std::exp(-double(q[3])));
    for (int k = 0; k < 3; ++k) f[41 + k] = q[8 + k];
    f[44] = 1.0f;
    for (int k = 45; k < 64; ++k) f[k] = 0.0f;
}

struct Net {
    std::vector<double> w;  // NP, rounded if bf16
    bool r;
    Net(const float* w32, bool bf16) : w(NP), r(bf16) {
        for (int i = 0; i < NP; ++i) w[i] = bf16 ? bf16r(w32[i]) : w32[i];
    }
    double rnd(double x) const { return r ? double(bf16r(float(x))) : x; }
    const double* layer(int l) const { return w.data() + l * 4096; }
    // x[l] = input of layer l (64 values), l = 0..5; y = 3 outputs
    void forward(const float* q, double x[6][64], double y[3]) const {
        float f[64];
        encode(q, f);
        for (int k = 0; k < 64; ++k) x[0][k] = rnd(f[k]);
        for (int l = 0; l < 5; ++l) {
            const double* W = layer(l);
            for (int o = 0; o < 64; ++o) {
                double z = 0;
                for (int i = 0; i < 64; ++i) z += W[o * 64 + i] * x[l][i];
                if (r) z = double(float(z));
                x[l + 1][o] = rnd(z > 0 ? z : 0.0);
            }
        }
        const double* W = layer(5);
        for (int o = 0; o < 3; ++o) {
            double z = 0;
            for (int i = 0; i < 64; ++i) z += W[o * 64 + i] * x[5][i];
            y[o] = z;
        }
    }
};

}  // namespace

extern "C" {

void orc_nrc_encode(const float* q, int n, float* feat) {
    for (int s = 0; s < n; ++s) encode(q + size_t(s) * 16, feat + size_t(s) * 64);
}

void orc_nrc_forward(const float* w, const float* q, int n, int bf16, double* y) {
    const Net net(w, bf16 != 0);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        double x[6][64];
        net.forward(q + size_t(s) * 16, x, y + size_t(s) * 3);
    }
}

// Mean relative L2 loss over n samples and its gradient (NP doubles, zero-initialised here).
double orc_nrc_grad(const float* w, const float* q, const float* t, int n, int bf16, double* grad) {
    const Net net(w, bf16 != 0);
    const int nt = omp_get_max_threads();
    std::vector<std::vector<double>> gacc(nt, std::vector<double>(NP, 0.0));
    std::vector<double> lacc(nt, 0.0);
    const double inv_n = 1.0 / n;
#pragma omp parallel
    {
        const int tid = omp_get_thread_num();
        std::vector<double>& G = gacc[tid];
#pragma omp for schedule(static)
        for (int s = 0; s < n; ++s) {
            double x[6][64], y[3], g[64], d[64];
            net.forward(q + size_t(s) * 16, x, y);
            for (int k = 0; k < 64; ++k) g[k] = 0;
            for (int c = 0; c < 3; ++c) {
                const double den = y[c] * y[c] + 0.01, e = y[c] - t[size_t(s) * 3 + c];
                if (bf16) {
                    const float den32 = float(y[c]) * float(y[c]) + 0.01f, e32 = float(y[c]) - t[size_t(s) * 3 + c];
                    g[c] = net.rnd(double(2.0f * e32 / den32 * float(inv_n)));
                } else {
                    g[c] = 2 * e / den * inv_n;
                }
                lacc[tid] += e * e / den * inv_n;
            }
            for (int l = 5; l >= 0; --l) {
                const double* W = net.layer(l);
                const int rows = l < 5 ? 64 : 3;
                for (int o = 0; o < rows; ++o)
                    for (int i = 0; i < 64; ++i) G[l * 4096 + o * 64 + i] += g[o] * x[l][i];
                if (l == 0) break;
                for (int i = 0; i < 64; ++i) {
                    double acc = 0;
                    for (int o = 0; o < rows; ++o) acc += W[o * 64 + i] * g[o];
                    if (bf16) acc = double(float(acc));
                    d[i] = x[l][i] > 0 ? net.rnd(acc) : 0.0;
                }
                for (int i = 0; i < 64; ++i) g[i] = d[i];
            }
        }
    }
    double loss = 0;
    for (int i = 0; i < NP; ++i) grad[i] = 0;
    for (int k = 0; k < nt; ++k) {
        loss += lacc[k];
        for (int i = 0; i < NP; ++i) grad[i] += gacc[k][i];
    }
    return loss;
}

}  // extern "C"
