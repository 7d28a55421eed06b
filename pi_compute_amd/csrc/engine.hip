// engine.hip - host side of the C ABI (include/pintt_gpu.h).
//
// Owns, per process, one device context: a stream, the twiddle tables (one per
// prime, grown to the largest transform seen), two residue arenas of 3 * N
// words, the carry scratch and the staging buffers of the host entry points.
// Products are serialised by a mutex; device-pointer products issued on other
// streams are ordered behind the previous product with an event.
//
// Reference counterparts: IntNttEngine::ensure_buffers / TwiddleCache::build
// (src/gpu/int_ntt.cu:110-186,282-299), which rebuilt the tables whenever the
// length changed, and the per-GPU mutex of the wrappers
// (src/arithmetic/int_ntt_multiplier.cpp:137-144).
#include "pintt_gpu.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ntt_kernels.h"

namespace pintt {
namespace {

thread_local std::string t_error;

struct Engine {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int tw_log2 = 0;
    uint2* tw_fwd[kPrimes] = {};
    uint2* tw_inv[kPrimes] = {};
    size_t arena_n = 0;  // per-prime capacity of bufA / bufB
    uint32_t* bufA = nullptr;
    uint32_t* bufB = nullptr;
    size_t carry_cap = 0;  // carry blocks covered by `status`
    uint64_t* status = nullptr;
    CarrySync* sync = nullptr;
    size_t io_cap = 0;  // limbs per staging buffer
    uint64_t* d_a = nullptr;
    uint64_t* d_b = nullptr;
    uint64_t* d_out = nullptr;
    bool timing = false;
    hipEvent_t ev[5] = {};
    bool have_times = false;
    // Tile plan per transform length: tuned on the device at first use
    // (unless disabled), or set explicitly.
    bool autotune = true;
    bool have_plan[kMaxLog2N + 1] = {};
    Plan plan[kMaxLog2N + 1] = {};
    float plan_ms[kMaxLog2N + 1] = {};
    size_t tune_cap = 0;  // limbs in each tuning operand buffer
    uint64_t* tune_in = nullptr;
    uint64_t* tune_out = nullptr;
};

Engine g;
std::mutex g_mu;

int fail(int code, const std::string& what) {
    t_error = what;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(e == hipErrorOutOfMemory ? PINTT_ERR_NO_MEMORY : PINTT_ERR_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define PINTT_HIP(call)                                     \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

uint32_t powmod(uint32_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = b % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
uint32_t invmod(uint32_t a, uint32_t p) { return powmod(a % p, p - 2, p); }
uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
uint2 shoup(uint32_t w, uint32_t p) { return make_uint2(w, (uint32_t)(((uint64_t)w << 32) / p)); }

CrtConst crt_constants(int log2n) {
    const uint32_t p0 = kP[0], p1 = kP[1], p2 = kP[2];
    auto scale = [&](uint32_t p) {  // 2^32 / N mod p
        return mulmod((uint32_t)((1ull << 32) % p), invmod((uint32_t)((1ull << log2n) % p), p), p);
    };
    CrtConst c;
    c.k0 = shoup(scale(p0), p0);
    uint32_t ip0_1 = invmod(p0 % p1, p1);
    c.k1 = shoup(mulmod(scale(p1), ip0_1, p1), p1);
    c.c1 = shoup(ip0_1, p1);
    uint32_t ip01_2 = invmod(mulmod(p0 % p2, p1 % p2, p2), p2);
    c.k2 = shoup(mulmod(scale(p2), ip01_2, p2), p2);
    c.c20 = shoup(ip01_2, p2);
    c.c21 = shoup(invmod(p1 % p2, p2), p2);
    return c;
}

int ensure_init(int device) {
    if (g.ready) {
        if (device < 0 || device == g.device) return PINTT_OK;
        return fail(PINTT_ERR_ARGS, "engine already initialised on another device");
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(PINTT_ERR_NO_DEVICE, "no HIP device available");
    int dev = device < 0 ? 0 : device;
    if (dev >= count) return fail(PINTT_ERR_NO_DEVICE, "device index out of range");
    PINTT_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    PINTT_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PINTT_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                             ", the kernels are built for gfx950 only");
    PINTT_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    PINTT_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    for (auto& e : g.ev) PINTT_HIP(hipEventCreate(&e));
    const char* at = std::getenv("PINTT_AUTOTUNE");
    g.autotune = !(at && at[0] == '0');
    g.device = dev;
    g.ready = true;
    return PINTT_OK;
}

template <class T>
int grow(T*& ptr, size_t count) {
    if (ptr) {
        PINTT_HIP(hipFree(ptr));
        ptr = nullptr;
    }
    PINTT_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
    return PINTT_OK;
}

int ensure_capacity(int log2n, size_t nwords) {
    int rc;
    if (log2n > g.tw_log2) {
        int tl = log2n < 16 ? 16 : log2n;
        PINTT_HIP(hipStreamSynchronize(g.stream));
        size_t cnt = (size_t)1 << tl;
        for (int i = 0; i < kPrimes; ++i) {
            if ((rc = grow(g.tw_fwd[i], cnt))) return rc;
            if ((rc = grow(g.tw_inv[i], cnt))) return rc;
            PINTT_HIP(launch_build_twiddles(g.tw_fwd[i], g.tw_inv[i], i, tl, g.stream));
        }
        g.tw_log2 = tl;
    }
    size_t n = (size_t)1 << log2n;
    if (n > g.arena_n) {
        PINTT_HIP(hipStreamSynchronize(g.stream));
        if ((rc = grow(g.bufA, kPrimes * n))) return rc;
        if ((rc = grow(g.bufB, kPrimes * n))) return rc;
        g.arena_n = n;
    }
    if (!g.sync) {
        if ((rc = grow(g.sync, 1))) return rc;
        PINTT_HIP(hipMemset(g.sync, 0, sizeof(CarrySync)));
    }
    size_t nblk = (nwords + kCarryWordsPerBlock - 1) / kCarryWordsPerBlock;
    if (nblk > g.carry_cap) {
        // Status words start at zero; epochs start at 1, so stale words never match.
        PINTT_HIP(hipStreamSynchronize(g.stream));
        if ((rc = grow(g.status, nblk))) return rc;
        PINTT_HIP(hipMemset(g.status, 0, nblk * sizeof(uint64_t)));
        g.carry_cap = nblk;
    }
    return PINTT_OK;
}

int log2_for(size_t an, size_t bn) {
    size_t need = 2 * an + 2 * bn - 1;
    int k = kMinLog2N;
    while (k <= kMaxLog2N && ((size_t)1 << k) < need) ++k;
    return k > kMaxLog2N ? -1 : k;
}

Twiddles twiddles() {
    Twiddles t;
    for (int i = 0; i < kPrimes; ++i) {
        t.fwd[i] = g.tw_fwd[i];
        t.inv[i] = g.tw_inv[i];
    }
    return t;
}

std::vector<Plan> candidate_plans(int log2n) {
    std::vector<Plan> out;
    for (int r = 5; r <= 11; ++r) {
        int c = log2n - r;
        if (c < 5 || c > 14) continue;
        for (int w = 3; w <= 6; ++w)
            for (int rows = 0; rows <= 8; ++rows) {
                Plan pl{log2n, r, c, w, rows};
                if (plan_valid(pl)) out.push_back(pl);
            }
    }
    return out;
}

// Enqueue one product on `s` with tile plan `pl`; device pointers throughout.
int enqueue_product(const Plan& pl, const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn,
                    uint64_t* d_out, bool square, hipStream_t s, bool timed) {
    const int log2n = pl.log2n;
    const size_t nwords = 2 * (an + bn);
    const Twiddles tw = twiddles();
    const size_t n = (size_t)1 << log2n;
    if (s != g.stream) PINTT_HIP(hipStreamWaitEvent(s, g.done, 0));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[0], s));
    PINTT_HIP(launch_forward_columns(pl, reinterpret_cast<const uint32_t*>(d_a), (uint32_t)(2 * an),
                                     reinterpret_cast<const uint32_t*>(d_b), (uint32_t)(2 * bn),
                                     g.bufA, g.bufB, square ? 1 : 2, tw, s));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[1], s));
    PINTT_HIP(launch_row_convolution(pl, g.bufA, g.bufB, square ? 1 : 0, tw, s));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[2], s));
    PINTT_HIP(launch_inverse_columns(pl, g.bufA, tw, g.sync, s));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[3], s));
    PINTT_HIP(launch_crt_carry(g.bufA, n, reinterpret_cast<uint32_t*>(d_out), nwords, crt_constants(log2n),
                               g.sync, g.status, s));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[4], s));
    PINTT_HIP(hipEventRecord(g.done, s));
    return PINTT_OK;
}

// Time every compiled tile plan for length 2^log2n on synthetic operands
// (an = bn = N/4 limbs) and keep the fastest.  Runs on the library stream and
// waits for it, so it must not happen inside a stream capture.
int tune(int log2n) {
    const size_t n = (size_t)1 << log2n;
    const size_t an = n / 4, bn = n / 4;
    int rc = ensure_capacity(log2n, 2 * (an + bn));
    if (rc) return rc;
    if (g.tune_cap < an) {
        PINTT_HIP(hipStreamSynchronize(g.stream));
        if ((rc = grow(g.tune_in, 2 * an))) return rc;
        if ((rc = grow(g.tune_out, 2 * an))) return rc;
        PINTT_HIP(hipMemset(g.tune_in, 0xA7, 2 * an * sizeof(uint64_t)));
        g.tune_cap = an;
    }
    Plan best = make_plan(log2n);
    float best_ms = 1e30f;
    for (const Plan& pl : candidate_plans(log2n)) {
        const int reps = log2n >= 22 ? 3 : 5;
        // warm-up, then `reps` timed products
        if ((rc = enqueue_product(pl, g.tune_in, an, g.tune_in + an, bn, g.tune_out, false, g.stream, false)))
            return rc;
        PINTT_HIP(hipEventRecord(g.ev[0], g.stream));
        for (int i = 0; i < reps; ++i)
            if ((rc = enqueue_product(pl, g.tune_in, an, g.tune_in + an, bn, g.tune_out, false, g.stream, false)))
                return rc;
        PINTT_HIP(hipEventRecord(g.ev[4], g.stream));
        PINTT_HIP(hipEventSynchronize(g.ev[4]));
        float ms = 0.f;
        PINTT_HIP(hipEventElapsedTime(&ms, g.ev[0], g.ev[4]));
        ms /= reps;
        if (ms < best_ms) {
            best_ms = ms;
            best = pl;
        }
    }
    g.plan[log2n] = best;
    g.plan_ms[log2n] = best_ms;
    g.have_plan[log2n] = true;
    return PINTT_OK;
}

int plan_for(int log2n, Plan* out) {
    if (!g.have_plan[log2n]) {
        if (g.autotune) {
            int rc = tune(log2n);
            if (rc) return rc;
        } else {
            g.plan[log2n] = make_plan(log2n);
            g.have_plan[log2n] = true;
        }
    }
    *out = g.plan[log2n];
    return PINTT_OK;
}

// Enqueue one product on `s`; operands and output are device pointers.
int run_product(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn, uint64_t* d_out,
                bool square, hipStream_t s) {
    const int log2n = log2_for(an, bn);
    if (log2n < 0) return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
    int rc = ensure_capacity(log2n, 2 * (an + bn));
    if (rc) return rc;
    Plan pl;
    if ((rc = plan_for(log2n, &pl))) return rc;
    if ((rc = enqueue_product(pl, d_a, an, d_b, bn, d_out, square, s, g.timing))) return rc;
    if (g.timing) g.have_times = true;
    return PINTT_OK;
}

int host_product(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out,
                 bool square) {
    if (!a || !b || !out || an == 0 || bn == 0) return fail(PINTT_ERR_ARGS, "null pointer or zero length");
    if ((out < a + an && a < out + an + bn) || (out < b + bn && b < out + an + bn))
        return fail(PINTT_ERR_ARGS, "output overlaps an operand");
    std::lock_guard<std::mutex> lock(g_mu);
    int rc = ensure_init(-1);
    if (rc) return rc;
    if (log2_for(an, bn) < 0) return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
    PINTT_HIP(hipSetDevice(g.device));
    size_t need = an + bn;
    if (need > g.io_cap) {
        PINTT_HIP(hipStreamSynchronize(g.stream));
        if ((rc = grow(g.d_a, need))) return rc;
        if ((rc = grow(g.d_b, need))) return rc;
        if ((rc = grow(g.d_out, need))) return rc;
        g.io_cap = need;
    }
    PINTT_HIP(hipMemcpyAsync(g.d_a, a, an * sizeof(uint64_t), hipMemcpyHostToDevice, g.stream));
    if (!square) PINTT_HIP(hipMemcpyAsync(g.d_b, b, bn * sizeof(uint64_t), hipMemcpyHostToDevice, g.stream));
    rc = run_product(g.d_a, an, square ? g.d_a : g.d_b, bn, g.d_out, square, g.stream);
    if (rc) return rc;
    PINTT_HIP(hipMemcpyAsync(out, g.d_out, need * sizeof(uint64_t), hipMemcpyDeviceToHost, g.stream));
    uint32_t err = 0;
    PINTT_HIP(hipMemcpyAsync(&err, &g.sync->error, sizeof(err), hipMemcpyDeviceToHost, g.stream));
    PINTT_HIP(hipStreamSynchronize(g.stream));
    if (err) return fail(PINTT_ERR_HIP, "carry look-back gave up (device state corrupted)");
    return PINTT_OK;
}

}  // namespace
}  // namespace pintt

using namespace pintt;

extern "C" {

int gpu_ntt_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    return ensure_init(device);
}

void gpu_ntt_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g.ready) return;
    (void)hipSetDevice(g.device);
    (void)hipStreamSynchronize(g.stream);
    for (int i = 0; i < kPrimes; ++i) {
        (void)hipFree(g.tw_fwd[i]);
        (void)hipFree(g.tw_inv[i]);
    }
    (void)hipFree(g.bufA);
    (void)hipFree(g.bufB);
    (void)hipFree(g.status);
    (void)hipFree(g.sync);
    (void)hipFree(g.d_a);
    (void)hipFree(g.d_b);
    (void)hipFree(g.d_out);
    (void)hipFree(g.tune_in);
    (void)hipFree(g.tune_out);
    for (auto& e : g.ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(g.done);
    (void)hipStreamDestroy(g.stream);
    g = Engine{};
}

int gpu_ntt_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out) {
    return host_product(a, an, b, bn, out, a == b && an == bn);
}

int gpu_ntt_sqr(const uint64_t* a, size_t an, uint64_t* out) {
    return host_product(a, an, a, an, out, true);
}

int gpu_ntt_mul_device(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn,
                       uint64_t* d_out, void* stream) {
    if (!d_a || !d_b || !d_out || an == 0 || bn == 0)
        return fail(PINTT_ERR_ARGS, "null pointer or zero length");
    std::lock_guard<std::mutex> lock(g_mu);
    int rc = ensure_init(-1);
    if (rc) return rc;
    PINTT_HIP(hipSetDevice(g.device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g.stream;
    return run_product(d_a, an, d_b, bn, d_out, d_a == d_b && an == bn, s);
}

int gpu_ntt_transform_log2(size_t an, size_t bn) {
    if (an == 0 || bn == 0) return -1;
    return log2_for(an, bn);
}

size_t gpu_ntt_max_limbs(void) { return ((size_t)1 << (kMaxLog2N - 1)); }

const char* gpu_ntt_last_error(void) { return t_error.c_str(); }

int gpu_ntt_autotune(int log2n_min, int log2n_max) {
    std::lock_guard<std::mutex> lock(g_mu);
    int rc = ensure_init(-1);
    if (rc) return rc;
    if (log2n_min < kMinLog2N) log2n_min = kMinLog2N;
    if (log2n_max > kMaxLog2N) log2n_max = kMaxLog2N;
    PINTT_HIP(hipSetDevice(g.device));
    for (int k = log2n_min; k <= log2n_max; ++k)
        if ((rc = tune(k))) return rc;
    return PINTT_OK;
}

int gpu_ntt_get_plan(int log2n, int* log2r, int* log2w, int* log2rows, float* ms) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    Plan pl = g.have_plan[log2n] ? g.plan[log2n] : make_plan(log2n);
    if (log2r) *log2r = pl.log2r;
    if (log2w) *log2w = pl.log2w;
    if (log2rows) *log2rows = pl.log2rows;
    if (ms) *ms = g.have_plan[log2n] ? g.plan_ms[log2n] : 0.f;
    return PINTT_OK;
}

int gpu_ntt_set_plan(int log2n, int log2r, int log2w, int log2rows) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    Plan pl{log2n, log2r, log2n - log2r, log2w, log2rows};
    if (log2r <= 0) pl = make_plan(log2n);
    if (!plan_valid(pl)) return fail(PINTT_ERR_ARGS, "tile plan not compiled into this library");
    g.plan[log2n] = pl;
    g.plan_ms[log2n] = 0.f;
    g.have_plan[log2n] = true;
    return PINTT_OK;
}

int gpu_ntt_list_plans(int log2n, int* out, int max_plans) {
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return 0;
    std::vector<Plan> c = candidate_plans(log2n);
    int k = 0;
    for (const Plan& pl : c) {
        if (k >= max_plans) break;
        out[3 * k] = pl.log2r;
        out[3 * k + 1] = pl.log2w;
        out[3 * k + 2] = pl.log2rows;
        ++k;
    }
    return k;
}

void gpu_ntt_set_timing(int enable) {
    std::lock_guard<std::mutex> lock(g_mu);
    g.timing = enable != 0;
    g.have_times = false;
}

int gpu_ntt_pass_times(double* ms, int max) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g.ready || !g.have_times || !ms) return 0;
    if (hipEventSynchronize(g.ev[4]) != hipSuccess) return 0;
    int k = 0;
    for (int i = 0; i < 4 && k < max; ++i) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, g.ev[i], g.ev[i + 1]) != hipSuccess) return k;
        ms[k++] = t;
    }
    return k;
}

}  // extern "C"
