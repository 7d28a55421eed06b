// hip_ntt_multiplier.cpp - see hip_ntt_multiplier.h.
#include "hip_ntt_multiplier.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <stdexcept>

#include "pintt_gpu.h"

namespace pi {

HipNttMultiplier::HipNttMultiplier(size_t threshold, int device)
    : threshold_(threshold), device_(device) {
    if (gpu_ntt_init(device) != PINTT_OK)
        throw std::runtime_error(std::string("HipNttMultiplier: ") + gpu_ntt_last_error());
}

std::string HipNttMultiplier::device_name() const { return "gfx950 device " + std::to_string(device_); }

bool HipNttMultiplier::use_gpu(const mpz_t a, const mpz_t b) const {
    return std::min(mpz_size(a), mpz_size(b)) >= threshold_;
}

// |result| = |a| * |b| (or |a|^2) through the C ABI, then the sign.  The ABI
// forbids output/operand overlap, so an aliased result is built in a temporary.
void HipNttMultiplier::gpu_product(mpz_t result, const mpz_t a, const mpz_t b, bool square) {
    const int sign = mpz_sgn(a) * mpz_sgn(b);
    if (sign == 0) {
        mpz_set_ui(result, 0);
        return;
    }
    const size_t an = mpz_size(a), bn = mpz_size(b);
    if (an + bn > gpu_ntt_max_limbs()) {
        stats_.fallback_calls.fetch_add(1);
        mpz_mul(result, a, b);
        return;
    }
    const bool alias = (result == a) || (result == b);
    mpz_t tmp;
    mpz_ptr dst = result;
    if (alias) {
        mpz_init(tmp);
        dst = tmp;
    }
    auto t0 = std::chrono::steady_clock::now();
    const mp_limb_t* ap = mpz_limbs_read(a);
    const mp_limb_t* bp = square ? ap : mpz_limbs_read(b);
    mp_limb_t* out = mpz_limbs_write(dst, an + bn);
    static_assert(sizeof(mp_limb_t) == sizeof(uint64_t), "64-bit limbs expected");
    int rc = square ? gpu_ntt_sqr(reinterpret_cast<const uint64_t*>(ap), an, reinterpret_cast<uint64_t*>(out))
                    : gpu_ntt_mul(reinterpret_cast<const uint64_t*>(ap), an,
                                  reinterpret_cast<const uint64_t*>(bp), bn, reinterpret_cast<uint64_t*>(out));
    if (rc != PINTT_OK) {
        stats_.fallback_calls.fetch_add(1);
        mpz_limbs_finish(dst, 0);
        mpz_mul(dst, a, b);
    } else {
        size_t n = an + bn;
        while (n > 0 && out[n - 1] == 0) --n;
        mpz_limbs_finish(dst, static_cast<mp_size_t>(n));
        if (sign < 0) mpz_neg(dst, dst);
        stats_.gpu_calls.fetch_add(1);
        stats_.gpu_ns.fetch_add(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                    std::chrono::steady_clock::now() - t0)
                                    .count());
    }
    if (alias) {
        mpz_swap(result, tmp);
        mpz_clear(tmp);
    }
}

void HipNttMultiplier::multiply(mpz_t result, const mpz_t a, const mpz_t b) {
    if (!use_gpu(a, b)) {
        stats_.cpu_calls.fetch_add(1);
        mpz_mul(result, a, b);
        return;
    }
    gpu_product(result, a, b, a == b);
}

void HipNttMultiplier::square(mpz_t result, const mpz_t a) {
    if (!use_gpu(a, a)) {
        stats_.cpu_calls.fetch_add(1);
        mpz_mul(result, a, a);
        return;
    }
    gpu_product(result, a, a, true);
}

void HipNttMultiplier::reset_stats() {
    stats_.gpu_calls = 0;
    stats_.cpu_calls = 0;
    stats_.fallback_calls = 0;
    stats_.gpu_ns = 0;
}

void HipNttMultiplier::print_stats() const {
    const uint64_t g = stats_.gpu_calls.load();
    std::printf("  HIP NTT multiplier: %llu GPU calls, %llu CPU (below threshold), %llu fallbacks",
                (unsigned long long)g, (unsigned long long)stats_.cpu_calls.load(),
                (unsigned long long)stats_.fallback_calls.load());
    if (g) std::printf(", %.3f ms per GPU call", stats_.gpu_ns.load() / 1e6 / g);
    std::printf("\n");
}

}  // namespace pi
