// mlp.hip - training and inference kernels of the dispatch cost model (gfx950).
//
// 4 -> 32 (ReLU) -> 32 (ReLU) -> K, FP32.  The network is tiny (<= 1480
// parameters), so every kernel stages the parameters in LDS and gives one
// sample to one lane; activations and deltas live feature-major ([unit][n])
// so that the gradient reduction reads consecutive samples.
//   k_mlp_forward   y = f(x)
//   k_mlp_backprop  per sample: forward, masked squared error, deltas
//   k_mlp_reduce    one workgroup per parameter (+1 for the loss): sum over
//                   samples in a fixed order (deterministic gradients)
//   k_mlp_adam      one Adam step
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "pintt_mlp.h"

namespace pintt_mlp {
namespace {

constexpr int IN = PINTT_MLP_INPUTS;
constexpr int H = PINTT_MLP_HIDDEN;
constexpr int KMAX = PINTT_MLP_MAX_OUTPUTS;
constexpr int oW1 = 0, oB1 = H * IN, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H;
__host__ __device__ constexpr int oB3(int k) { return oW3 + H * k; }
__host__ __device__ constexpr int nparams(int k) { return oW3 + (H + 1) * k; }
constexpr int kMaxParams = oW3 + (H + 1) * KMAX;
constexpr int kThreads = 256;

thread_local std::string t_error;

int fail(int code, const std::string& what) {
    t_error = what;
    return code;
}

#define MLP_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(PINTT_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ void stage_params(float* lds, const float* __restrict__ params, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) lds[i] = params[i];
    __syncthreads();
}

// Hidden layers of one sample; h1, h2 hold post-ReLU activations.
__device__ __forceinline__ void hidden(const float* w, const float* x, float* h1, float* h2) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float s = w[oB1 + i];
#pragma unroll
        for (int j = 0; j < IN; ++j) s += w[oW1 + i * IN + j] * x[j];
        h1[i] = s > 0.f ? s : 0.f;
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float s = w[oB2 + i];
#pragma unroll
        for (int j = 0; j < H; ++j) s += w[oW2 + i * H + j] * h1[j];
        h2[i] = s > 0.f ? s : 0.f;
    }
}

__device__ __forceinline__ float output(const float* w, int k, int K, const float* h2) {
    float s = w[oB3(K) + k];
#pragma unroll
    for (int j = 0; j < H; ++j) s += w[oW3 + k * H + j] * h2[j];
    return s;
}

// x: [IN][n], y: [K][n]
__global__ void __launch_bounds__(kThreads) k_mlp_forward(const float* __restrict__ params, int K,
                                                          const float* __restrict__ x, uint32_t n,
                                                          float* __restrict__ y) {
    __shared__ float w[kMaxParams];
    stage_params(w, params, nparams(K));
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xs[IN], h1[H], h2[H];
#pragma unroll
    for (int j = 0; j < IN; ++j) xs[j] = x[(size_t)j * n + s];
    hidden(w, xs, h1, h2);
    for (int k = 0; k < K; ++k) y[(size_t)k * n + s] = output(w, k, K, h2);
}

struct Acts {
    float* h1;    // [H][n]
    float* h2;    // [H][n]
    float* d1;    // [H][n]  dL/d(pre-activation 1)
    float* d2;    // [H][n]  dL/d(pre-activation 2)
    float* dy;    // [K][n]  dL/dy
    float* loss;  // [n]
};

__global__ void __launch_bounds__(kThreads) k_mlp_backprop(const float* __restrict__ params, int K,
                                                           const float* __restrict__ x,
                                                           const int* __restrict__ path,
                                                           const float* __restrict__ target, uint32_t n,
                                                           float inv_n, Acts a) {
    __shared__ float w[kMaxParams];
    stage_params(w, params, nparams(K));
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xs[IN], h1[H], h2[H];
#pragma unroll
    for (int j = 0; j < IN; ++j) xs[j] = x[(size_t)j * n + s];
    hidden(w, xs, h1, h2);
    const int ps = path[s];
    const float e = output(w, ps, K, h2) - target[s];
    a.loss[s] = 0.5f * e * e * inv_n;
    const float g = e * inv_n;
    for (int k = 0; k < K; ++k) a.dy[(size_t)k * n + s] = k == ps ? g : 0.f;
    float d2[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        d2[i] = h2[i] > 0.f ? w[oW3 + ps * H + i] * g : 0.f;
        a.h1[(size_t)i * n + s] = h1[i];
        a.h2[(size_t)i * n + s] = h2[i];
        a.d2[(size_t)i * n + s] = d2[i];
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float t = 0.f;
#pragma unroll
        for (int j = 0; j < H; ++j) t += w[oW2 + j * H + i] * d2[j];
        a.d1[(size_t)i * n + s] = h1[i] > 0.f ? t : 0.f;
    }
}

// Workgroup b computes parameter b's gradient (b == nparams(K): the loss).
__global__ void __launch_bounds__(kThreads) k_mlp_reduce(int K, const float* __restrict__ x, Acts a, uint32_t n,
                                                         float* __restrict__ grad, float* __restrict__ loss_out) {
    __shared__ float part[kThreads];
    const int b = blockIdx.x;
    const float* u = nullptr;  // delta row
    const float* v = nullptr;  // input row (nullptr: bias)
    if (b < oB1) {
        u = a.d1 + (size_t)(b / IN) * n;
        v = x + (size_t)(b % IN) * n;
    } else if (b < oW2) {
        u = a.d1 + (size_t)(b - oB1) * n;
    } else if (b < oB2) {
        u = a.d2 + (size_t)((b - oW2) / H) * n;
        v = a.h1 + (size_t)((b - oW2) % H) * n;
    } else if (b < oW3) {
        u = a.d2 + (size_t)(b - oB2) * n;
    } else if (b < oB3(K)) {
        u = a.dy + (size_t)((b - oW3) / H) * n;
        v = a.h2 + (size_t)((b - oW3) % H) * n;
    } else if (b < nparams(K)) {
        u = a.dy + (size_t)(b - oB3(K)) * n;
    } else {
        u = a.loss;
    }
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) s += v ? u[i] * v[i] : u[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (b < nparams(K)) grad[b] = part[0];
        else *loss_out = part[0];
    }
}

__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int count, float step, float b1, float b2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= step * mi / (sqrtf(vi) + eps);
}

// Device buffers of one call, freed on every exit path.
struct Buffers {
    std::vector<void*> ptrs;
    ~Buffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        hipError_t e = hipMalloc((void**)p, count * sizeof(T) + 16);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

std::vector<float> feature_major(const float* x, size_t n, int width) {
    std::vector<float> t((size_t)width * n);
    for (size_t s = 0; s < n; ++s)
        for (int j = 0; j < width; ++j) t[(size_t)j * n + s] = x[s * width + j];
    return t;
}

int check_args(const float* params, int K, const float* x, size_t n) {
    if (!params || !x || n == 0 || K < 1 || K > KMAX) return fail(PINTT_ERR_ARGS, "bad arguments");
    if (n > (1u << 26)) return fail(PINTT_ERR_TOO_LARGE, "more than 2^26 samples");
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return fail(PINTT_ERR_NO_DEVICE, "no HIP device");
    return PINTT_OK;
}

// Uploads the training data and allocates the activations.
struct Training {
    Buffers buf;
    float *p, *g, *x, *t, *loss;
    int* path;
    Acts a;
    int K;
    uint32_t n;
    int setup(const float* params, int K_, const float* xh, const int* ph, const float* th, size_t n_) {
        K = K_;
        n = (uint32_t)n_;
        for (size_t s = 0; s < n_; ++s)
            if (ph[s] < 0 || ph[s] >= K) return fail(PINTT_ERR_ARGS, "path index out of range");
        const int np = nparams(K);
        if (buf.alloc(&p, np) || buf.alloc(&g, np) || buf.alloc(&x, (size_t)IN * n) || buf.alloc(&t, n) ||
            buf.alloc(&path, n) || buf.alloc(&loss, 1) || buf.alloc(&a.h1, (size_t)H * n) ||
            buf.alloc(&a.h2, (size_t)H * n) || buf.alloc(&a.d1, (size_t)H * n) || buf.alloc(&a.d2, (size_t)H * n) ||
            buf.alloc(&a.dy, (size_t)K * n) || buf.alloc(&a.loss, n))
            return fail(PINTT_ERR_NO_MEMORY, "hipMalloc failed");
        std::vector<float> xt = feature_major(xh, n, IN);
        MLP_HIP(hipMemcpy(p, params, np * sizeof(float), hipMemcpyHostToDevice));
        MLP_HIP(hipMemcpy(x, xt.data(), xt.size() * sizeof(float), hipMemcpyHostToDevice));
        MLP_HIP(hipMemcpy(t, th, n * sizeof(float), hipMemcpyHostToDevice));
        MLP_HIP(hipMemcpy(path, ph, n * sizeof(int), hipMemcpyHostToDevice));
        return PINTT_OK;
    }
    int gradient() {
        const dim3 grid((n + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(k_mlp_backprop, grid, dim3(kThreads), 0, 0, p, K, x, path, t, n, 1.0f / n, a);
        MLP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_reduce, dim3(nparams(K) + 1), dim3(kThreads), 0, 0, K, x, a, n, g, loss);
        MLP_HIP(hipGetLastError());
        return PINTT_OK;
    }
};

}  // namespace
}  // namespace pintt_mlp

using namespace pintt_mlp;

extern "C" {

size_t gpu_mlp_param_count(int outputs) {
    return (outputs < 1 || outputs > KMAX) ? 0 : (size_t)nparams(outputs);
}

const char* gpu_mlp_last_error(void) { return t_error.c_str(); }

int gpu_mlp_forward(const float* params, int K, const float* x, size_t n, float* y) {
    if (int rc = check_args(params, K, x, n)) return rc;
    if (!y) return fail(PINTT_ERR_ARGS, "null output");
    Buffers buf;
    float *dp, *dx, *dy;
    if (buf.alloc(&dp, nparams(K)) || buf.alloc(&dx, (size_t)IN * n) || buf.alloc(&dy, (size_t)K * n))
        return fail(PINTT_ERR_NO_MEMORY, "hipMalloc failed");
    std::vector<float> xt = feature_major(x, n, IN);
    MLP_HIP(hipMemcpy(dp, params, nparams(K) * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dx, xt.data(), xt.size() * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, 0, dp, K, dx,
                       (uint32_t)n, dy);
    MLP_HIP(hipGetLastError());
    std::vector<float> yt((size_t)K * n);
    MLP_HIP(hipMemcpy(yt.data(), dy, yt.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; ++s)
        for (int k = 0; k < K; ++k) y[s * K + k] = yt[(size_t)k * n + s];
    return PINTT_OK;
}

int gpu_mlp_gradient(const float* params, int K, const float* x, const int* path, const float* target, size_t n,
                     float* grad, float* loss) {
    if (int rc = check_args(params, K, x, n)) return rc;
    if (!path || !target || !grad) return fail(PINTT_ERR_ARGS, "null pointer");
    Training tr;
    if (int rc = tr.setup(params, K, x, path, target, n)) return rc;
    if (int rc = tr.gradient()) return rc;
    MLP_HIP(hipMemcpy(grad, tr.g, nparams(K) * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) MLP_HIP(hipMemcpy(loss, tr.loss, sizeof(float), hipMemcpyDeviceToHost));
    return PINTT_OK;
}

int gpu_mlp_train(float* params, int K, const float* x, const int* path, const float* target, size_t n, int steps,
                  float lr, float* loss_history) {
    if (int rc = check_args(params, K, x, n)) return rc;
    if (!path || !target || steps < 0) return fail(PINTT_ERR_ARGS, "bad arguments");
    Training tr;
    if (int rc = tr.setup(params, K, x, path, target, n)) return rc;
    const int np = nparams(K);
    float *m, *v, *hist = nullptr;
    if (tr.buf.alloc(&m, np) || tr.buf.alloc(&v, np) || (steps && tr.buf.alloc(&hist, steps)))
        return fail(PINTT_ERR_NO_MEMORY, "hipMalloc failed");
    MLP_HIP(hipMemset(m, 0, np * sizeof(float)));
    MLP_HIP(hipMemset(v, 0, np * sizeof(float)));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double b1t = 1.0, b2t = 1.0;
    for (int it = 0; it < steps; ++it) {
        if (int rc = tr.gradient()) return rc;
        MLP_HIP(hipMemcpyAsync(hist + it, tr.loss, sizeof(float), hipMemcpyDeviceToDevice, 0));
        b1t *= b1;
        b2t *= b2;
        // Bias correction folded into the step size (Kingma & Ba, Algorithm 1 note).
        const float step = (float)(lr * std::sqrt(1.0 - b2t) / (1.0 - b1t));
        hipLaunchKernelGGL(k_mlp_adam, dim3((np + kThreads - 1) / kThreads), dim3(kThreads), 0, 0, tr.p, tr.g, m,
                           v, np, step, b1, b2, eps);
        MLP_HIP(hipGetLastError());
    }
    MLP_HIP(hipMemcpy(params, tr.p, np * sizeof(float), hipMemcpyDeviceToHost));
    if (loss_history && steps) MLP_HIP(hipMemcpy(loss_history, hist, steps * sizeof(float), hipMemcpyDeviceToHost));
    return PINTT_OK;
}

}  // extern "C"
