/* pintt_mlp.h - C ABI of the learned dispatch cost model (gfx950 kernels).
 *
 * The reference chooses the CPU or GPU path with a fixed limb threshold
 * (src/arithmetic/int_ntt_multiplier.cpp:56-59, src/arithmetic/gpu_ntt_multiplier.cpp:67-72,
 * default 10000 from src/main.cpp:63).  These entry points have no counterpart
 * there: they train and evaluate the small FP32 network that replaces that
 * threshold (8 inputs -> 32 ReLU -> 32 ReLU -> K outputs, one predicted
 * log-time per path).  The threshold rule stays the fallback.
 *
 * Parameter vector (float, this order, row-major, row = output unit):
 *   W1[32][8], b1[32], W2[32][32], b2[32], W3[K][32], b3[K]
 * Samples are sample-major: x[n][8], y[n][K].
 * Loss (training): L = 1/(2n) * sum_s (y[s][path[s]] - target[s])^2, i.e. each
 * sample supervises only the output of the path it was measured on.
 * All functions use the current HIP device, block until done, and return a
 * pintt_status (pintt_gpu.h); gpu_mlp_last_error() describes the last failure.
 */
#ifndef PINTT_MLP_H
#define PINTT_MLP_H

#include <stddef.h>

#include "pintt_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PINTT_MLP_INPUTS 8
#define PINTT_MLP_HIDDEN 32
#define PINTT_MLP_MAX_OUTPUTS 8

/* 1344 + 33 * outputs, or 0 when outputs is outside [1, 8]. */
size_t gpu_mlp_param_count(int outputs);

/* Inference: y = f(x) for n samples. */
int gpu_mlp_forward(const float* params, int outputs, const float* x, size_t n, float* y);

/* Full-batch gradient of L (grad has gpu_mlp_param_count(outputs) entries)
 * and the loss itself; deterministic (fixed reduction order). */
int gpu_mlp_gradient(const float* params, int outputs, const float* x, const int* path, const float* target,
                     size_t n, float* grad, float* loss);

/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, step size
 * lr) on device-resident data, updating params in place.  loss_history, if not
 * null, receives the loss before each step (steps entries). */
int gpu_mlp_train(float* params, int outputs, const float* x, const int* path, const float* target, size_t n,
                  int steps, float lr, float* loss_history);

const char* gpu_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PINTT_MLP_H */
