// decimal.h - parallel divide-and-conquer conversion of a non-negative mpz to
// decimal (the role of the reference's BaseConverter::parallel_to_decimal,
// src/io/base_converter.h:37-41; same output: the digits of n, no sign).
//
// Powers 10^(2^i) are built once by squaring; a number known to hold at most
// `width` digits is split at the largest power below width, the halves are
// converted independently (threads at the top levels) and written into their
// slots of one preallocated string, left-padded with '0'.
#pragma once

#include <gmp.h>

#include <cstddef>
#include <memory>
#include <string>
#include <vector>

#include "device_ops.h"
#include "multiplier.h"

namespace pi {

std::string to_decimal(mpz_srcptr n, unsigned threads = 0);

// Same digits; the splits above `leaf_digits` run level by level, every node
// of a level divided by the same power 10^(2^j) at once: quotient estimate
// q = (top bits of x) * (Newton reciprocal of 10^(2^j)) / 2^s, remainder
// x - q * 10^(2^j), both products of a level as one multiply_batch, then at
// most a few +-1 corrections on the host.  Below leaf_digits the host
// conversion above takes over, one subtree per thread.
struct DecimalStats {
    double powers_s = 0, reciprocals_s = 0, products_s = 0, corrections_s = 0, leaves_s = 0;
    long min_correction = 0, max_correction = 0;  // quotient = estimate + correction, over all splits
};

// The powers 10^(2^j) and their Newton reciprocals that the conversion of a
// number of `width` digits uses: they depend only on width and leaf_digits, so
// a caller that knows the width early can build them beforehand (for pi,
// beside binary splitting).
struct DecimalPowers;

// Device copies of the powers 10^(2^j) and reciprocals with 2^j >= the leaf
// width (j >= jmin), made once by decimal_powers(..., device = true).
struct DevicePowers {
    dev::Buffers mem;
    std::vector<dev::Num> P, inv;  // indexed by j
    int jmin;
    DevicePowers(const DecimalPowers& d, int jmin);
};

struct DecimalPowers {
    size_t width = 0, leaf_digits = 0;
    std::vector<__mpz_struct> pw;   // 10^(2^j), j <= k with 2^k < width <= 2^(k+1)
    // 2^e / pw[j] - 9 .. 2^e / pw[j] - 1, e = 2 bits(pw[j]) + 64 rounded up to
    // whole limbs, where 2^j >= leaf_digits; else 0
    std::vector<__mpz_struct> inv;
    double powers_s = 0, reciprocals_s = 0;
    std::unique_ptr<DevicePowers> dev;  // device copies (device = true)
    DecimalPowers() = default;
    ~DecimalPowers();
    DecimalPowers(const DecimalPowers&) = delete;
    DecimalPowers& operator=(const DecimalPowers&) = delete;
};
// device: the reciprocals as whole-limb Newton steps on device numbers
// (newton.h, device_reciprocal_limbs), and device copies of what the device
// conversion reads; the powers go through m either way.
void decimal_powers(DecimalPowers& d, size_t width, Multiplier& m, size_t leaf_digits = 1 << 15,
                    bool device = false);

// `pre`: powers built for a width with the same k and for the same leaf_digits
// (others are ignored).  `device`: the levels run on device-resident numbers
// through the GPU engine's device entry points (m then only builds missing
// powers): n, the powers and the reciprocals are uploaded once, each level is
// two device product batches and flagged device subtractions, and only the
// leaves come back.  Throws std::runtime_error if the device fails.
std::string to_decimal(mpz_srcptr n, Multiplier& m, unsigned threads = 0, size_t leaf_digits = 1 << 15,
                       DecimalStats* stats = nullptr, const DecimalPowers* pre = nullptr, bool device = false);

// The same digits from a non-negative number already in device memory (xn
// limbs at x, top limbs may be zero), every level and the leaves on the
// device: nothing of it crosses PCIe but the digits.  pre must hold the
// powers for its width and leaf_digits (at most 4,096).  Throws
// std::runtime_error if they do not or the device fails.
std::string to_decimal_device(const uint64_t* x, size_t xn, unsigned threads, size_t leaf_digits,
                              DecimalStats* stats, const DecimalPowers& pre);

}  // namespace pi
