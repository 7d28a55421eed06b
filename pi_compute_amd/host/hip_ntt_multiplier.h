// hip_ntt_multiplier.h - Multiplier backed by the gfx950 NTT engine.
//
// Drop-in beside GmpMultiplier.  Semantics follow the reference GPU wrappers
// (src/arithmetic/int_ntt_multiplier.cpp:115-159,
//  src/arithmetic/gpu_ntt_multiplier.cpp:144-210):
//   * min(|a|, |b|) in limbs below the threshold -> mpz_mul on the host;
//   * sign = sgn(a) * sgn(b), zero operand -> result 0;
//   * engine refusal (too large, device error) -> mpz_mul, product still exact.
// Differences: operands cross as raw GMP limbs (no digit vectors), squaring
// takes one forward transform, and result may alias a or b.
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>

#include "multiplier.h"

namespace pi {

class HipNttMultiplier : public Multiplier {
public:
    // threshold: minimum limbs of the smaller operand for the GPU path
    // (reference default 10000, src/main.cpp:63).  device: HIP device index.
    // Throws std::runtime_error when no usable device exists, like the
    // reference constructors (src/arithmetic/int_ntt_multiplier.cpp:26-31).
    explicit HipNttMultiplier(size_t threshold = 10000, int device = 0);
    ~HipNttMultiplier() override = default;

    void multiply(mpz_t result, const mpz_t a, const mpz_t b) override;
    void square(mpz_t result, const mpz_t a) override;

    size_t threshold() const { return threshold_; }
    void set_threshold(size_t t) { threshold_ = t; }
    std::string device_name() const;

    struct Stats {
        std::atomic<uint64_t> gpu_calls{0};
        std::atomic<uint64_t> cpu_calls{0};       // below threshold
        std::atomic<uint64_t> fallback_calls{0};  // engine refused or failed
        std::atomic<uint64_t> gpu_ns{0};
    };
    const Stats& stats() const { return stats_; }
    void reset_stats();
    void print_stats() const;

private:
    bool use_gpu(const mpz_t a, const mpz_t b) const;
    void gpu_product(mpz_t result, const mpz_t a, const mpz_t b, bool square);

    size_t threshold_;
    int device_;
    mutable Stats stats_;
};

}  // namespace pi
