// no_device.cpp - the C ABI (include/pintt_gpu.h, include/pintt_mlp.h) of a
// CPU-only build, for hosts without ROCm: linked into the host programs in
// place of libpintt.so (`make -C pi_compute_amd cpu`, g++ only).
//
// Every entry point answers as libpintt.so does on a host with no gfx950
// device: PINTT_ERR_NO_DEVICE, computing nothing (there is no CPU product path
// behind the ABI).  GMP-only runs never reach it; `pi_compute_hip --gpu`
// fails at HipNttMultiplier's construction, like the reference's GPU
// multipliers without a device (src/arithmetic/int_ntt_multiplier.cpp:26-38).
// Three answers describe layouts rather than compute and stay exact: the
// transform-length rule, the chunk-cut rule and the cost model's parameter
// count.
//
// Reference counterpart: the PI_CUDA_ENABLED gate around every GPU source
// (src/arithmetic/gpu_ntt_multiplier.h:14, src/gpu/int_ntt.cu:10).
#include "pintt_gpu.h"
#include "pintt_mlp.h"

namespace {
const char kNoDevice[] = "CPU-only build: the gfx950 engine (libpintt.so) is not linked in";
int no_device() { return PINTT_ERR_NO_DEVICE; }
}  // namespace

extern "C" {

int gpu_ntt_init(int) { return no_device(); }
void gpu_ntt_shutdown(void) {}
int gpu_ntt_mul(const uint64_t*, size_t, const uint64_t*, size_t, uint64_t*) { return no_device(); }
int gpu_ntt_sqr(const uint64_t*, size_t, uint64_t*) { return no_device(); }
int gpu_ntt_mul_device(const uint64_t*, size_t, const uint64_t*, size_t, uint64_t*, void*) { return no_device(); }
int gpu_ntt_mul_device_noalloc(const uint64_t*, size_t, const uint64_t*, size_t, uint64_t*, void*) {
    return no_device();
}
int gpu_ntt_mul_batch(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*, const size_t*,
                      uint64_t* const*) {
    return no_device();
}
int gpu_ntt_mul_batch_device(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*, const size_t*,
                             uint64_t* const*, void*) {
    return no_device();
}
int gpu_ntt_graph_create(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*, const size_t*,
                         uint64_t* const*, void** graph) {
    if (graph) *graph = nullptr;
    return no_device();
}
int gpu_ntt_graph_launch(void*, void*) { return no_device(); }
size_t gpu_ntt_graph_nodes(void*) { return 0; }
void gpu_ntt_graph_destroy(void*) {}
int gpu_ntt_dev_alloc(size_t, uint64_t** d_ptr) {
    if (d_ptr) *d_ptr = nullptr;
    return no_device();
}
int gpu_ntt_dev_free(uint64_t* d_ptr) { return d_ptr ? no_device() : PINTT_OK; }
int gpu_ntt_upload(size_t count, uint64_t* const*, const uint64_t* const*, const size_t*) {
    return count ? no_device() : PINTT_OK;
}
int gpu_ntt_download(size_t count, uint64_t* const*, const uint64_t* const*, const size_t*) {
    return count ? no_device() : PINTT_OK;
}
int gpu_ntt_dev_copy(uint64_t*, const uint64_t*, size_t) { return no_device(); }
int gpu_ntt_add_batch_device(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*, const size_t*,
                             const int*, uint64_t* const*) {
    return no_device();
}
int gpu_ntt_add_batch_device_async(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*,
                                   const size_t*, const int*, uint64_t* const*, uint32_t*) {
    return no_device();
}
int gpu_ntt_add_batch_device_flags(size_t, const uint64_t* const*, const size_t*, const uint64_t* const*,
                                   const size_t*, const int*, uint64_t* const*, int*) {
    return no_device();
}
int gpu_ntt_decimal_leaves_device(size_t, const uint64_t* const*, const size_t*, const size_t*, const size_t*,
                                  uint8_t*, size_t) {
    return no_device();
}
int gpu_ntt_bs_leaves(uint64_t, uint64_t, size_t, uint64_t*, uint32_t*, int32_t*) { return no_device(); }
int gpu_ntt_synchronize(void) { return PINTT_OK; }
void gpu_ntt_set_split(int) {}
void gpu_ntt_set_graph_branches(int) {}
// The length rule of the engine: the least 2^k >= 2 an + 2 bn - 1 with
// 10 <= k <= 25, -1 above.
int gpu_ntt_transform_log2(size_t an, size_t bn) {
    if (an == 0 || bn == 0) return -1;
    const size_t need = 2 * an + 2 * bn - 1;
    int k = 10;
    while (k <= 25 && ((size_t)1 << k) < need) ++k;
    return k > 25 ? -1 : k;
}
int gpu_ntt_product_plan(size_t, size_t, int*, int*, size_t*) { return no_device(); }
size_t gpu_ntt_max_limbs(void) { return (size_t)1 << 24; }
void gpu_ntt_set_chunk_cost(pintt_chunk_cost_fn, void*) {}
// The engine's cut rule (engine.hip chunk_shape): chunks of ceil(an / m)
// limbs rounded up to 16, none shorter than b, the last not empty, and a
// transform shorter than the uncut product's.
int gpu_ntt_chunk_plan(size_t an, size_t bn, int chunks, size_t* chunk_limbs, int* log2n) {
    if (an < bn) {
        const size_t t = an;
        an = bn;
        bn = t;
    }
    const int full = gpu_ntt_transform_log2(an, bn);
    if (full < 0 || chunks < 1 || chunks > 4) return 0;
    if (chunks == 1) {
        if (chunk_limbs) *chunk_limbs = an;
        if (log2n) *log2n = full;
        return 1;
    }
    const size_t m = (size_t)chunks;
    size_t h = (an + m - 1) / m;
    h = (h + 15) & ~(size_t)15;
    if (h < bn || (m - 1) * h >= an) return 0;
    const int k = gpu_ntt_transform_log2(h, bn);
    if (k < 0 || k >= full) return 0;
    if (chunk_limbs) *chunk_limbs = h;
    if (log2n) *log2n = k;
    return 1;
}
const char* gpu_ntt_last_error(void) { return kNoDevice; }
int gpu_ntt_autotune(int, int) { return no_device(); }
int gpu_ntt_get_plan(int, int*, int*, int*, float*) { return no_device(); }
int gpu_ntt_set_batch_plan(int, int, int, int) { return no_device(); }
int gpu_ntt_get_batch_plan(int, int*, int*, int*) { return no_device(); }
int gpu_ntt_set_plan(int, int, int, int) { return no_device(); }
int gpu_ntt_list_plans(int, int*, int) { return 0; }
int gpu_ntt_save_plans(const char*) { return no_device(); }
int gpu_ntt_load_plans(const char*) { return no_device(); }
void gpu_ntt_set_timing(int) {}
int gpu_ntt_pass_times(double*, int) { return 0; }

// 4 -> 32 -> 32 -> K, weights and biases per layer.
size_t gpu_mlp_param_count(int outputs) {
    constexpr size_t in = PINTT_MLP_INPUTS, h = PINTT_MLP_HIDDEN;
    if (outputs < 1 || outputs > PINTT_MLP_MAX_OUTPUTS) return 0;
    return h * in + h + h * h + h + (h + 1) * (size_t)outputs;
}
int gpu_mlp_forward(const float*, int, const float*, size_t, float*) { return no_device(); }
int gpu_mlp_gradient(const float*, int, const float*, const int*, const float*, size_t, float*, float*) {
    return no_device();
}
int gpu_mlp_train(float*, int, const float*, const int*, const float*, size_t, int, float, float*) {
    return no_device();
}
const char* gpu_mlp_last_error(void) { return kNoDevice; }

}  // extern "C"
