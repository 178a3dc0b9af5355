// oracle_nrc.cpp - CPU ORACLE (test infrastructure only) of the neural
// radiance cache network: encoding, forward, relative-L2 loss gradient and
// the Adam step.  The cache has no counterpart in the reference; this file
// states the method the GPU kernels implement.
//
// Network: 13 inputs -> features s * 24 + 2 k + c = sin (c = 0) / cos (c = 1)
// of 2^k pi x_s, k = 0..11, feature 312 = 1, features 313..319 = 0 -> layer 0
// (320 x 64) -> ReLU -> 4 x (64 x 64, ReLU) -> 64 x 3 (linear).  Weights are
// stored per layer [in][out], row-major, in one array of 37056.
// Loss over n samples: (1 / 3n) sum (y - t)^2 / (y^2 + 0.01) with y in the
// denominator held constant for the gradient.
//
// mode 0 ("exact"): all arithmetic in double, nothing rounded.
// mode 1 ("bf16"): the rounding points of the GPU kernels - weights,
// encoded features, hidden activations, output gradients and
// pre-activation gradients rounded to bf16 (round to nearest even), sums in
// double; the output gradient is formed in float as the GPU does.
//
// Pinned by (tests/test_cpu_nrc.py): central finite differences of the
// exact-mode loss (gradient of every layer), an independent numpy matrix
// restatement of the forward pass, sin^2 + cos^2 = 1 and the closed-form
// features at dyadic inputs, and the Adam recurrences on hand-checked steps.
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kIn = 13, kFreq = 12, kEnc = 312, kX = 320, kW = 64, kOut = 3;
constexpr int kW0 = 0, kWH = kX * kW, kW5 = kWH + 4 * kW * kW, kNW = kW5 + kW * kOut;

float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;  // inf / nan unchanged
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float y;
    std::memcpy(&y, &u, 4);
    return y;
}
double q(double x, int mode) { return mode ? (double)bf16_round((float)x) : x; }

void encode(const float* x, int mode, double* f) {
    for (int s = 0; s < kIn; ++s)
        for (int k = 0; k < kFreq; ++k) {
            const double a = std::ldexp(M_PI, k) * (double)x[s];
            f[s * 24 + 2 * k] = q(std::sin(a), mode);
            f[s * 24 + 2 * k + 1] = q(std::cos(a), mode);
        }
    f[kEnc] = 1.0;
    for (int i = kEnc + 1; i < kX; ++i) f[i] = 0.0;
}

struct Net {
    std::vector<double> w;  // rounded per mode
    explicit Net(const double* wsrc, int mode) : w(wsrc, wsrc + kNW) {
        if (mode)
            for (double& v : w) v = (double)bf16_round((float)v);
    }
    const double* layer(int l) const {
        return l == 0 ? &w[kW0] : l < 5 ? &w[kWH + (l - 1) * kW * kW] : &w[kW5];
    }
};

// Activations of one sample: a[0] = features (320), a[1..5] = hidden (64).
struct Acts {
    double a0[kX];
    double h[5][kW];
    double y[kOut];
};

void forward(const Net& net, const float* x, int mode, Acts* A) {
    encode(x, mode, A->a0);
    const double* in = A->a0;
    int nin = kX;
    for (int l = 0; l < 5; ++l) {
        const double* W = net.layer(l);
        for (int o = 0; o < kW; ++o) {
            double z = 0.0;
            for (int i = 0; i < nin; ++i) z += in[i] * W[i * kW + o];
            A->h[l][o] = q(z > 0.0 ? z : 0.0, mode);
        }
        in = A->h[l];
        nin = kW;
    }
    const double* W = net.layer(5);
    for (int o = 0; o < kOut; ++o) {
        double z = 0.0;
        for (int i = 0; i < kW; ++i) z += A->h[4][i] * W[i * kOut + o];
        A->y[o] = mode ? (double)(float)z : z;
    }
}

}  // namespace

extern "C" {

int orc_nrc_num_weights() { return kNW; }

void orc_nrc_encode(const float* x, int mode, double* out) { encode(x, mode, out); }

void orc_nrc_forward(const double* w, const float* in, uint64_t n, int mode, double* out) {
    const Net net(w, mode);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)n; ++s) {
        Acts A;
        forward(net, in + s * kIn, mode, &A);
        for (int o = 0; o < kOut; ++o) out[s * kOut + o] = A.y[o];
    }
}

// grad: kNW doubles (overwritten); loss: mean relative L2.
void orc_nrc_grad(const double* w, const float* in, const float* target, uint64_t n, int mode,
                  double* grad, double* loss) {
    const Net net(w, mode);
    std::memset(grad, 0, sizeof(double) * kNW);
    double lsum = 0.0;
#pragma omp parallel
    {
        std::vector<double> g(kNW, 0.0);
        double lloc = 0.0;
#pragma omp for schedule(static)
        for (int64_t s = 0; s < (int64_t)n; ++s) {
            Acts A;
            forward(net, in + s * kIn, mode, &A);
            double gy[kOut];
            for (int o = 0; o < kOut; ++o) {
                const double t = target[s * kOut + o];
                if (mode) {
                    const float yf = (float)A.y[o], tf = (float)t;
                    const float dd = yf - tf;
                    const float den = yf * yf + 0.01f;
                    lloc += (double)(dd * dd / den);
                    const float gscale = 2.0f / (3.0f * (float)n);
                    gy[o] = (double)bf16_round(gscale * dd / den);
                } else {
                    const double dd = A.y[o] - t, den = A.y[o] * A.y[o] + 0.01;
                    lloc += dd * dd / den;
                    gy[o] = 2.0 * dd / den / (3.0 * (double)n);
                }
            }
            // output layer
            double* G5 = &g[kW5];
            for (int i = 0; i < kW; ++i)
                for (int o = 0; o < kOut; ++o) G5[i * kOut + o] += A.h[4][i] * gy[o];
            double dz[kW], dnext[kW];
            const double* W5 = net.layer(5);
            for (int i = 0; i < kW; ++i) {
                double d = 0.0;
                for (int o = 0; o < kOut; ++o) d += W5[i * kOut + o] * gy[o];
                dz[i] = A.h[4][i] > 0.0 ? q(d, mode) : 0.0;
            }
            // hidden layers 4..1 and layer 0
            for (int l = 4; l >= 0; --l) {
                const double* inp = l == 0 ? A.a0 : A.h[l - 1];
                const int nin = l == 0 ? kX : kW;
                double* G = l == 0 ? &g[kW0] : &g[kWH + (l - 1) * kW * kW];
                for (int i = 0; i < nin; ++i) {
                    const double xi = inp[i];
                    if (xi == 0.0) continue;
                    for (int o = 0; o < kW; ++o) G[i * kW + o] += xi * dz[o];
                }
                if (l == 0) break;
                const double* W = net.layer(l);
                for (int i = 0; i < kW; ++i) {
                    double d = 0.0;
                    for (int o = 0; o < kW; ++o) d += W[i * kW + o] * dz[o];
                    dnext[i] = A.h[l - 1][i] > 0.0 ? q(d, mode) : 0.0;
                }
                std::memcpy(dz, dnext, sizeof dz);
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < kNW; ++k) grad[k] += g[k];
            lsum += lloc;
        }
    }
    if (loss) *loss = n ? lsum / (3.0 * (double)n) : 0.0;
}

// Adam with explicit bias corrections, float arithmetic in the GPU's order.
void orc_nrc_adam(float* w, float* m, float* v, const float* g, uint64_t nw, float lr,
                  float beta1, float beta2, float eps, float bias1, float bias2) {
    for (uint64_t k = 0; k < nw; ++k) {
        const float mk = beta1 * m[k] + (1.0f - beta1) * g[k];
        const float vk = beta2 * v[k] + (1.0f - beta2) * (g[k] * g[k]);
        m[k] = mk;
        v[k] = vk;
        const float mhat = mk / bias1;
        const float vhat = vk / bias2;
        w[k] = w[k] - lr * mhat / (std::sqrt(vhat) + eps);
    }
}

}  // extern "C"
