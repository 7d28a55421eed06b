// pi_engine.h - Chudnovsky pi over a Multiplier (the reference's PiEngine,
// src/engine/pi_engine.h:16-46, floating-point final step of
// src/engine/pi_engine.cpp:433-503).
//
//   terms   = terms_needed(D + 100)
//   (P,Q,R) = binary splitting over [0, terms)
//   pi      = 426880 * Q * sqrt(10005) / R       at (D+100)*3.3219281 + 64 bits
//   digits  = "3." + the first D decimals of floor(pi * 10^D), no newline
//
// Final step: FinalStep::Mpf is the reference's mpf sequence; FinalStep::Newton
// computes the same quotient with fixed-point Newton iterations (newton.h)
// whose products go through the Multiplier, 64 bits beyond that precision.
// Conversion::Multiplier splits the decimal conversion with Newton quotients
// through the Multiplier as well (decimal.h).  Either way the digits are the
// truncation of the same value: the error is below 10^-(D+90) of pi * 10^D.
#pragma once

#include <cstddef>
#include <string>

#include "binary_splitting.h"
#include "dispatch.h"
#include "multiplier.h"

namespace pi {

struct PiConfig {
    size_t digits = 1000;
    std::string output_file;  // empty: do not write
    bool verbose = false;
    unsigned num_threads = 0;  // 0: hardware concurrency
    bool level_batches = true;  // level-synchronous tree with one batch per level
    bool device_levels = false;  // upper levels with P, Q, R resident on the GPU (level_batches only)
    size_t device_from_limbs = 0;  // first device level: smallest operand at least this many limbs
    const Dispatcher* device_switch = nullptr;  // if set, it picks the first device level instead
    enum class FinalStep { Mpf, Newton } final_step = FinalStep::Mpf;
    enum class Conversion { Host, Multiplier } conversion = Conversion::Host;
    // Conversion::Multiplier: host conversion below this width (device-resident
    // conversion levels: at most 4096, the widest leaf the device converts)
    size_t conversion_leaf_digits = 1 << 15;
    unsigned bs_segments = 8;                  // device levels: segments pipelined with their host subtrees
    // Terms per host subtree at the bottom of the tree; 0: 128 with device-resident
    // levels (the device's small-product kernel takes the next level), else 256.
    unsigned long bs_leaf_span = 0;
    // sqrt(10005) and 10^D (FinalStep::Newton) and the conversion's powers of
    // ten and their reciprocals (Conversion::Multiplier) depend only on D, so
    // they are computed on another thread while binary splitting runs.
    bool overlap_constants = true;
    // Conversion::Multiplier with device-resident levels: the conversion's
    // levels also run on device-resident numbers (decimal.h), leaves on the host.
    bool device_conversion = true;
    // FinalStep::Newton with device-resident levels: the final step on
    // device-resident numbers too (pi_scaled_newton_device).
    bool device_final = true;
    // Device-resident levels: the bottom subtrees of binary splitting on the
    // device too (gpu_ntt_bs_leaves) instead of host threads.
    bool device_leaves = true;
};

struct PiResult {
    std::string digits;
    double elapsed_seconds;
    unsigned long terms_used;
    double bs_seconds, final_seconds, conv_seconds;
};

// The factors of the Newton final step that do not depend on Q and R:
// y = 2^(2 bits) / sqrt(10005) to `bits` bits (newton_inv_sqrt_ui) and 10^digits.
struct FinalConstants {
    mpz_t inv_sqrt, pow10;
    double inv_sqrt_s = 0, pow10_s = 0;
    FinalConstants();
    ~FinalConstants();
    FinalConstants(const FinalConstants&) = delete;
    FinalConstants& operator=(const FinalConstants&) = delete;
};
// device (bits a multiple of 64): the inverse square root as whole-limb
// device Newton steps (newton.h, device_inv_sqrt_limbs).
void final_constants(FinalConstants& c, size_t digits, size_t bits, Multiplier& m, bool device = false);

// floor(426880 * sqrt(10005) * Q * 10^digits / R) with `bits` of working
// precision, products through m (R > 0).  Used by FinalStep::Newton.  With
// `pre`, its constants (computed for the same digits and bits) are used.
struct NewtonStats {
    double inv_sqrt_s = 0, pow10_s = 0, reciprocal_s = 0, products_s = 0;
};
void pi_scaled_newton(mpz_t out, mpz_srcptr Q, mpz_srcptr R, size_t digits, size_t bits, Multiplier& m,
                      NewtonStats* stats = nullptr, const FinalConstants* pre = nullptr);

// Same quotient with every operand in device memory (the GPU engine's device
// entry points): Q, R and the constants are uploaded once, cut to w/64 + 2
// limbs, the reciprocal of R runs as whole-limb Newton steps (newton.h,
// newton_reciprocal_limbs) and only the result comes back.  w must be a
// multiple of 64 and pre computed with it.  Throws std::runtime_error if the
// device fails or a check on the way does.
void pi_scaled_newton_device(mpz_t out, mpz_srcptr Q, mpz_srcptr R, size_t w, const FinalConstants& pre);
// The same from Q and R already in device memory (binary_splitting.h,
// compute_device_qr): nothing of them crosses PCIe.
void pi_scaled_newton_device(mpz_t out, const DeviceQR& qr, size_t w, const FinalConstants& pre);
// ... and the result left in device memory too (for to_decimal_device).
struct DeviceInt {
    dev::Buffers mem;
    dev::Num x;
};
void pi_scaled_newton_device(DeviceInt& out, const DeviceQR& qr, size_t w, const FinalConstants& pre);

class PiEngine {
public:
    explicit PiEngine(Multiplier& m) : mult_(m) {}
    PiResult compute(const PiConfig& config);
    std::string compute_digits(size_t digits);

private:
    Multiplier& mult_;
};

}  // namespace pi
