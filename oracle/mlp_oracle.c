/*
 * mlp_oracle.c - CPU restatement, in double precision, of the dispatch cost
 * model's forward pass, masked squared-error loss, gradient and Adam step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
 * pi_compute_amd/csrc/mlp.hip.  The network has no counterpart in the
 * reference, whose dispatch is a fixed limb threshold
 * (src/arithmetic/int_ntt_multiplier.cpp:56-59); what it must compute is fixed
 * by the parameter layout and loss stated in include/pintt_mlp.h.
 * The gradient is pinned in the tests by central finite differences of
 * orc_mlp_loss, so it does not rest on this file's own algebra.
 */
#include <math.h>
#include <stddef.h>

#define IN 8
#define H 32

static int oW1(void) { return 0; }
static int oB1(void) { return H * IN; }
static int oW2(void) { return H * IN + H; }
static int oB2(void) { return H * IN + H + H * H; }
static int oW3(void) { return H * IN + 2 * H + H * H; }
static int oB3(int k) { return oW3() + H * k; }

size_t orc_mlp_param_count(int k) { return (size_t)(oW3() + (H + 1) * k); }

static void hidden(const double* w, const double* x, double* z1, double* h1, double* z2, double* h2) {
    for (int i = 0; i < H; ++i) {
        double s = w[oB1() + i];
        for (int j = 0; j < IN; ++j) s += w[oW1() + i * IN + j] * x[j];
        z1[i] = s;
        h1[i] = s > 0 ? s : 0;
    }
    for (int i = 0; i < H; ++i) {
        double s = w[oB2() + i];
        for (int j = 0; j < H; ++j) s += w[oW2() + i * H + j] * h1[j];
        z2[i] = s;
        h2[i] = s > 0 ? s : 0;
    }
}

static double out(const double* w, int k, int K, const double* h2) {
    double s = w[oB3(K) + k];
    for (int j = 0; j < H; ++j) s += w[oW3() + k * H + j] * h2[j];
    return s;
}

/* x[n][8] -> y[n][K] */
void orc_mlp_forward(const double* w, int K, const double* x, size_t n, double* y) {
    double z1[H], h1[H], z2[H], h2[H];
    for (size_t s = 0; s < n; ++s) {
        hidden(w, x + s * IN, z1, h1, z2, h2);
        for (int k = 0; k < K; ++k) y[s * K + k] = out(w, k, K, h2);
    }
}

/* L = 1/(2n) sum_s (y[s][path[s]] - t[s])^2 */
double orc_mlp_loss(const double* w, int K, const double* x, const int* path, const double* t, size_t n) {
    double z1[H], h1[H], z2[H], h2[H], L = 0;
    for (size_t s = 0; s < n; ++s) {
        hidden(w, x + s * IN, z1, h1, z2, h2);
        double e = out(w, path[s], K, h2) - t[s];
        L += 0.5 * e * e;
    }
    return L / (double)n;
}

/* Backpropagation (chain rule through the two ReLU layers). */
double orc_mlp_gradient(const double* w, int K, const double* x, const int* path, const double* t, size_t n,
                        double* g) {
    const size_t np = orc_mlp_param_count(K);
    for (size_t i = 0; i < np; ++i) g[i] = 0;
    double z1[H], h1[H], z2[H], h2[H], d1[H], d2[H], L = 0;
    for (size_t s = 0; s < n; ++s) {
        const double* xs = x + s * IN;
        hidden(w, xs, z1, h1, z2, h2);
        const int p = path[s];
        const double e = out(w, p, K, h2) - t[s];
        L += 0.5 * e * e;
        const double dy = e / (double)n;
        g[oB3(K) + p] += dy;
        for (int j = 0; j < H; ++j) g[oW3() + p * H + j] += dy * h2[j];
        for (int i = 0; i < H; ++i) d2[i] = z2[i] > 0 ? w[oW3() + p * H + i] * dy : 0;
        for (int i = 0; i < H; ++i) {
            g[oB2() + i] += d2[i];
            for (int j = 0; j < H; ++j) g[oW2() + i * H + j] += d2[i] * h1[j];
        }
        for (int j = 0; j < H; ++j) {
            double acc = 0;
            for (int i = 0; i < H; ++i) acc += w[oW2() + i * H + j] * d2[i];
            d1[j] = z1[j] > 0 ? acc : 0;
        }
        for (int i = 0; i < H; ++i) {
            g[oB1() + i] += d1[i];
            for (int j = 0; j < IN; ++j) g[oW1() + i * IN + j] += d1[i] * xs[j];
        }
    }
    return L / (double)n;
}

/* One Adam step, t >= 1 (Kingma & Ba 2015, Algorithm 1, with the bias
 * corrections folded into the step size and eps added to sqrt(v)). */
void orc_mlp_adam(double* w, const double* g, double* m, double* v, size_t np, int t, double lr) {
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double step = lr * sqrt(1.0 - pow(b2, t)) / (1.0 - pow(b1, t));
    for (size_t i = 0; i < np; ++i) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        w[i] -= step * m[i] / (sqrt(v[i]) + eps);
    }
}
