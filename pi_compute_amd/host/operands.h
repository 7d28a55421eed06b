// operands.h - seeded operands: splitmix64 stream (Vigna), limbs drawn least
// significant first, top limb masked to the bit length with its top bit set.
#pragma once

#include <gmp.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pi {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline void draw_operand(mpz_t r, SplitMix64& g, size_t bits) {
    const size_t n = (bits + 63) / 64;
    std::vector<uint64_t> limbs(n);
    for (auto& l : limbs) l = g.next();
    const size_t top = bits - 64 * (n - 1);
    if (top < 64) limbs[n - 1] &= (1ull << top) - 1;
    limbs[n - 1] |= 1ull << (top - 1);
    mpz_import(r, n, -1, 8, 0, 0, limbs.data());
}

}  // namespace pi
